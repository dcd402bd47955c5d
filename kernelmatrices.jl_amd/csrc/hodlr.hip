// HODLR build / symmetric factorization / solve / logdet / matvec on one
// device, composed from the batched primitives of la.h and assemble.h.
// Every stage takes a list of handles: independent problems of different
// sizes share each grouped launch (their levels are aligned from the leaves).
//
// Data layout in HBM (column-major):
//   pts    x[0..n) then y[0..n)
//   leaves leaf i (m_i x m_i, ld m_i) at leaf_mat_off[i]; Cholesky in place
//   P      n x Rtot, ld n.  Level l owns columns coff[l] .. coff[l]+rpad[l];
//          block (a0,b1,c0,d1) keeps U in rows a0..b1 and V in rows c0..d1
//          of those columns (ACA ranks below rpad are zero-padded).  Every
//          row belongs to exactly one child per level, so each level's panel
//          is one n x rpad matrix.  The factorization leaves there the
//          orthonormal Q factors, so P then holds M = blkdiag(Q_V, Q_U) of
//          every block (HODLR/src/utils.jl:29-35).
//   X, G   per block 2r x 2r: X = chol([I R1R2'; R2R1' I]) - I and
//          G = I - chol(...)^{-1}, so that W^{-1} t = t - M G M^T t and
//          W t = t + M X M^T t (HODLR/src/baseoverloads.jl:36-93 with
//          lrx_solterm = X (I+X)^{-1} = I - (I+X)^{-1}).
#include <algorithm>
#include <cmath>
#include <memory>

#include "assemble.h"
#include "hodlr.h"
#include "la.h"
#include "leaf.h"

namespace kmh {

static int floor_log2(int64_t n) {
    int l = -1;
    while (n) { n >>= 1; ++l; }
    return l;
}

// indices.jl:5-65
static void tree_indices(Hodlr& H) {
    const int fl = floor_log2(H.n);
    int lv;
    if (H.opts.level_kind == KMH_FIXED_LEVEL) {
        lv = std::min(std::max(0, H.opts.level), fl);
    } else {
        KMH_CHECK(H.opts.level >= 0, KMH_E_BAD_ARG, "Please provide a positive value for the level offset.");
        lv = std::min(std::max(0, fl - H.opts.level), fl);
    }
    H.L = lv;
    const int64_t nleaf = int64_t(1) << lv;
    const int64_t base = H.n / nleaf, rem = H.n - base * nleaf;
    H.leaf_off.assign(nleaf + 1, 0);
    for (int64_t i = 0; i < nleaf; ++i) H.leaf_off[i + 1] = H.leaf_off[i] + base + (i < rem ? 1 : 0);
    H.blocks.assign(lv, {});
    for (int j = 1; j <= lv; ++j) {
        const int64_t skp = int64_t(1) << j, half = int64_t(1) << (j - 1);
        for (int64_t i = 0; i < nleaf - 1; i += skp)
            H.blocks[j - 1].push_back({H.leaf_off[i + half], H.leaf_off[i + skp], H.leaf_off[i], H.leaf_off[i + half]});
    }
}

void check_kernel(const kmh_kernel& k, bool allow_deriv) {
    KMH_CHECK(k.family >= KMH_RATIONAL2 && k.family <= KMH_MATERN32, KMH_E_BAD_ARG, "unknown kernel family");
    const int dv = k.deriv;
    KMH_CHECK(dv == KMH_D0 || dv == KMH_D1 || dv == KMH_D2 || dv == KMH_D11 || dv == KMH_D12 || dv == KMH_D22,
              KMH_E_BAD_ARG, "unknown kernel derivative code");
    KMH_CHECK(allow_deriv || dv == KMH_D0, KMH_E_BAD_ARG, "the HODLR kernel must be a value kernel (deriv 0)");
    KMH_CHECK(dv == KMH_D0 || k.family == KMH_MATERN12 || k.family == KMH_MATERN32, KMH_E_BAD_ARG,
              "parameter derivatives are provided for the Matern families only");
    KMH_CHECK(std::isfinite(k.params[0]) && std::isfinite(k.params[1]), KMH_E_BAD_ARG, "non-finite kernel parameter");
    if (k.family != KMH_CAUCHY) KMH_CHECK(k.params[1] != 0.0, KMH_E_BAD_ARG, "kernel length scale must be nonzero");
}

// Read the per-handle Cholesky flags (one copy, one sync) and mark failures;
// extra doubles (e.g. per-handle sums) can ride along in the same copy.
static const char* const LANDMARK_FAIL = "Cholesky of the Nystrom landmark matrix K(P,P) failed.";

// Handles with a deferred landmark check (landmark_check) have their
// info[1] read too; a set info[1] reports the landmark Cholesky's failure.
static void collect_info(Ctx& c, const std::vector<Hodlr*>& hs, std::vector<Status>& st, const char* what,
                         const std::vector<const double*>& extra = {}, double* extra_out = nullptr) {
    std::vector<const int*> fp;
    std::vector<size_t> at;
    std::vector<char> lm;
    for (size_t i = 0; i < hs.size(); ++i)
        if (hs[i]) {
            fp.push_back(hs[i]->info.as<int>());
            at.push_back(i);
            lm.push_back(0);
            if (hs[i]->landmark_check) {
                fp.push_back(hs[i]->info.as<int>() + 1);
                at.push_back(i);
                lm.push_back(1);
                hs[i]->landmark_check = false;
            }
        }
    const size_t nd = extra.size(), ni = fp.size();
    if (ni + nd == 0) return;
    char* buf = static_cast<char*>(c.scratch(nd * sizeof(double) + ni * sizeof(int)));
    gather_doubles(c, extra, reinterpret_cast<double*>(buf));
    gather_ints(c, fp, reinterpret_cast<int*>(buf + nd * sizeof(double)));
    std::vector<char> h(nd * sizeof(double) + ni * sizeof(int));
    KMH_HIP(hipMemcpyAsync(h.data(), buf, h.size(), hipMemcpyDeviceToHost, c.q()));
    c.sync();
    c.scratch_free(buf);
    if (nd) std::memcpy(extra_out, h.data(), nd * sizeof(double));
    for (size_t q = 0; q < ni; ++q) {
        int flag;
        std::memcpy(&flag, h.data() + nd * sizeof(double) + q * sizeof(int), sizeof(int));
        const size_t i = at[q];
        if (flag && lm[q]) {
            // the landmark matrix failed first; what followed is void
            st[i].code = KMH_E_NOT_PD;
            st[i].msg = std::string("PosDefException: matrix is not positive definite; ") + LANDMARK_FAIL;
        } else if (flag && st[i].code == KMH_OK) {
            st[i].code = KMH_E_NOT_PD;
            st[i].msg = std::string("PosDefException: matrix is not positive definite; ") + what;
        }
    }
}

static void reset_info(Ctx& c, const std::vector<Hodlr*>& hs) {
    std::vector<int*> p;
    for (auto* h : hs)
        if (h) p.push_back(h->info.as<int>());
    zero_ints(c, p);
}

// ---------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------

static void validate(const BuildReq& r) {
    KMH_CHECK(r.d == 2, KMH_E_BAD_ARG, "points must be 2-dimensional");
    KMH_CHECK(r.n >= 1 && r.n < (int64_t(1) << 31), KMH_E_BAD_ARG, "invalid number of points");
    KMH_CHECK(r.pts != nullptr, KMH_E_BAD_ARG, "null points");
    check_kernel(r.kern, false);
    const auto& o = r.opts;
    KMH_CHECK(o.method == KMH_ACA || o.method == KMH_NYSTROM, KMH_E_BAD_ARG, "unknown compression method");
    KMH_CHECK(o.level_kind == KMH_FIXED_LEVEL || o.level_kind == KMH_LOG_LEVEL, KMH_E_BAD_ARG, "unknown level kind");
    KMH_CHECK(o.precision == KMH_FP64 || o.precision == KMH_MIXED || o.precision == KMH_MIXED_Q, KMH_E_BAD_ARG,
              "unsupported precision mode");
    KMH_CHECK(o.maxrank >= 0, KMH_E_BAD_ARG, "maxrank must be >= 0");
}

static void nystrom_setup(Hodlr& H) {
    const int r = H.opts.maxrank;
    KMH_CHECK(r >= 2, KMH_E_BAD_ARG, "Please supply a fixed off-diagonal rank that is >=2.");
    int64_t minleaf = H.n;
    for (int i = 0; i < H.nleaves(); ++i) minleaf = std::min(minleaf, H.leaf_size(i) - 1);
    KMH_CHECK(r < minleaf, KMH_E_BAD_ARG, "Your nystrom rank is too big. Reduce the HODLR level or nystrom rank.");
    // landmarks: Int64.(round.(LinRange(1, n, r)))  (constructor.jl:31)
    H.nyind.resize(r);
    for (int j = 0; j < r; ++j) {
        const double t = (double)j / (double)(r - 1);
        H.nyind[j] = (int64_t)std::nearbyint((1.0 - t) * 1.0 + t * (double)H.n) - 1;
    }
    H.nys_r = r;
    Ctx& c = *H.ctx;
    H.lm.alloc(&c, 2 * (size_t)r * sizeof(double));
    H.kpp.alloc(&c, (size_t)r * r * sizeof(double));
    H.kxp.alloc(&c, (size_t)H.n * r * sizeof(double));
    H.ynys.alloc(&c, (size_t)H.n * r * sizeof(double));
    H.rpad.assign(H.L, r);
    H.coff.assign(H.L + 1, 0);
    for (int l = 0; l < H.L; ++l) H.coff[l + 1] = H.coff[l] + r;
    H.Rtot = H.coff[H.L];
    H.ranks.assign(H.L, {});
    for (int l = 0; l < H.L; ++l) H.ranks[l].assign(H.blocks[l].size(), r);
}

// ACA of every block of every level of every ACA handle, all blocks
// stepping together.  A level whose block reached the panel's column
// capacity gets a panel twice as wide holding the columns found so far, and
// its stopped blocks continue where they stopped.
static void build_aca_batch(Ctx& c, const std::vector<Hodlr*>& hs, std::vector<Status>& st,
                            const std::vector<int>& idx) {
    struct Lv { int h; int l; int cap; int maxsz; int b0; DevBuf tmp; };
    std::vector<Lv> lvs;
    int nblk = 0;
    for (int i : idx) {
        Hodlr& H = *hs[i];
        for (int l = 0; l < H.L; ++l) {
            int ms_max = 1;
            for (const auto& b : H.blocks[l]) {
                const int h = (int)std::min(b.b1 - b.a0, b.d1 - b.c0);
                ms_max = std::max(ms_max, H.opts.maxrank == 0 ? h : std::min(H.opts.maxrank, h));
            }
            lvs.push_back({i, l, std::min(ms_max, 256), ms_max, nblk, DevBuf()});
            nblk += (int)H.blocks[l].size();
        }
    }
    if (lvs.empty()) return;
    DevBuf out;
    out.alloc(&c, 2 * sizeof(int) * nblk);
    int* d_rank = out.as<int>();
    int* d_stat = d_rank + nblk;
    std::vector<int> h_out(2 * nblk);
    std::vector<AcaDesc> ds;
    for (auto& v : lvs) {
        Hodlr& H = *hs[v.h];
        v.tmp.alloc(&c, (size_t)H.n * v.cap * sizeof(double));
        double* T = v.tmp.as<double>();
        const double* px = H.pts.as<double>();
        int bi = v.b0;
        for (const auto& b : H.blocks[v.l]) {
            const int nro = (int)(b.b1 - b.a0), ncl = (int)(b.d1 - b.c0);
            const int h = std::min(nro, ncl);
            const int ms = H.opts.maxrank == 0 ? h : std::min(H.opts.maxrank, h);
            ds.push_back({px, px + H.n, px, px + H.n, (int)b.a0, nro, (int)b.c0, ncl, T + b.a0, T + b.c0, (int)H.n,
                          ms, v.cap, 0, d_rank + bi, d_stat + bi, H.opts.tol, to_kdev(H.kern)});
            ++bi;
        }
    }
    AcaRun run(c, ds);
    for (;;) {
        run.run();
        KMH_HIP(hipMemcpyAsync(h_out.data(), d_rank, 2 * sizeof(int) * nblk, hipMemcpyDeviceToHost, c.q()));
        c.sync();
        bool again = false;
        for (auto& v : lvs) {
            Hodlr& H = *hs[v.h];
            const int nbl = (int)H.blocks[v.l].size();
            bool wider = false;
            for (int k = 0; k < nbl; ++k) {
                const int s = h_out[nblk + v.b0 + k];
                if (s == 1 && st[v.h].code == KMH_OK) {
                    st[v.h].code = KMH_E_ACA;
                    st[v.h].msg = "Incompatible maxrank or matrix dimensions.";
                }
                if (s == 2) wider = true;
            }
            if (!wider || st[v.h].code != KMH_OK) continue;
            KMH_CHECK(v.cap < v.maxsz, KMH_E_DEVICE, "ACA capacity logic error");
            const int cap = std::min(v.maxsz, v.cap * 2);
            if (cap > 4096) {
                st[v.h].code = KMH_E_BAD_ARG;
                st[v.h].msg = "ACA rank exceeds the supported maximum of 4096";
                continue;
            }
            DevBuf wide;
            wide.alloc(&c, (size_t)H.n * cap * sizeof(double));
            KMH_HIP(hipMemcpyAsync(wide.p, v.tmp.p, (size_t)H.n * v.cap * sizeof(double), hipMemcpyDeviceToDevice,
                                   c.q()));
            v.tmp = std::move(wide);
            v.cap = cap;
            double* T = v.tmp.as<double>();
            for (int k = 0; k < nbl; ++k) {
                if (h_out[nblk + v.b0 + k] != 2) continue;
                const auto& b = H.blocks[v.l][k];
                run.widen(v.b0 + k, T + b.a0, T + b.c0, cap);
            }
            again = true;
        }
        if (!again) break;
    }
    // ranks, panel widths, compaction into P
    std::vector<Copy2D> cp;
    int bi = 0;
    size_t q = 0;
    for (int i : idx) {
        Hodlr& H = *hs[i];
        H.ranks.assign(H.L, {});
        H.rpad.assign(H.L, 1);
        size_t q0 = q;
        for (int l = 0; l < H.L; ++l, ++q)
            for (size_t k = 0; k < H.blocks[l].size(); ++k, ++bi) {
                H.ranks[l].push_back(h_out[bi]);
                H.rpad[l] = std::max(H.rpad[l], h_out[bi]);
            }
        if (st[i].code != KMH_OK) continue;
        H.coff.assign(H.L + 1, 0);
        for (int l = 0; l < H.L; ++l) H.coff[l + 1] = H.coff[l] + H.rpad[l];
        H.Rtot = H.coff[H.L];
        H.P.alloc(&c, (size_t)H.n * H.Rtot * (H.p32() ? sizeof(float) : sizeof(double)));
        KMH_HIP(hipMemsetAsync(H.P.p, 0, H.P.bytes, c.q()));
        for (int l = 0; l < H.L; ++l) {
            const size_t Pl = (size_t)H.coff[l] * H.n;
            const double* T = lvs[q0 + l].tmp.as<double>();
            for (size_t k = 0; k < H.blocks[l].size(); ++k) {
                const auto& b = H.blocks[l][k];
                const int r = H.ranks[l][k];
                // the mixed form keeps the ACA panels in FP32
                Copy2D u{T + b.a0, (int)H.n, H.Pat(Pl + b.a0), (int)H.n, (int)(b.b1 - b.a0), r, 1.0};
                Copy2D v{T + b.c0, (int)H.n, H.Pat(Pl + b.c0), (int)H.n, (int)(b.d1 - b.c0), r, 1.0};
                u.ed = v.ed = H.p32();
                cp.push_back(u);
                cp.push_back(v);
            }
        }
    }
    copy2d(c, cp);
    c.sync();
}

std::vector<Hodlr*> hodlr_build_batch(Ctx& c, const std::vector<BuildReq>& reqs, std::vector<Status>& st) {
    c.set_device();
    const size_t R = reqs.size();
    st.assign(R, Status());
    std::vector<std::unique_ptr<Hodlr>> own(R);
    std::vector<Hodlr*> hs(R, nullptr);
    std::vector<std::vector<double>> soa(R);
    std::vector<Copy2D> dev_pts;
    std::vector<int*> info0;
    bool defer = R > 0;
    for (const auto& q : reqs) defer = defer && q.defer_checks && q.on_device;
    for (size_t i = 0; i < R; ++i) {
        try {
            const BuildReq& q = reqs[i];
            validate(q);
            own[i].reset(new Hodlr());
            Hodlr& H = *own[i];
            H.ctx = &c;
            H.n = q.n;
            H.kern = q.kern;
            H.opts = q.opts;
            tree_indices(H);
            H.info.alloc(&c, 256);
            H.pts.alloc(&c, 2 * (size_t)q.n * sizeof(double));
            double* px = H.pts.as<double>();
            if (!q.on_device) {
                soa[i].resize(2 * (size_t)q.n);
                for (int64_t j = 0; j < q.n; ++j) {
                    soa[i][j] = q.pts[2 * j];
                    soa[i][q.n + j] = q.pts[2 * j + 1];
                }
                KMH_HIP(hipMemcpyAsync(px, soa[i].data(), soa[i].size() * sizeof(double), hipMemcpyHostToDevice,
                                       c.q()));
            } else {
                dev_pts.push_back({q.pts, 2, px, (int)q.n, 2, (int)q.n, 1.0});
            }
            const int nl = H.nleaves();
            H.leaf_mat_off.assign(nl + 1, 0);
            for (int l = 0; l < nl; ++l) H.leaf_mat_off[l + 1] = H.leaf_mat_off[l] + H.leaf_size(l) * H.leaf_size(l);
            H.leaves.alloc(&c, (size_t)H.leaf_mat_off[nl] * sizeof(double));
            if (H.L > 0 && H.nystrom()) nystrom_setup(H);
            if (H.L == 0) {
                H.coff.assign(1, 0);
                H.Rtot = 0;
            }
            hs[i] = own[i].get();
            info0.push_back(H.info.as<int>());   // only handles that survive validation
            if (defer) info0.push_back(H.info.as<int>() + 1);
        } catch (const Error& e) {
            st[i].code = e.code;
            st[i].msg = e.what();
            own[i].reset();
        }
    }
    zero_ints(c, info0);
    copy2d(c, dev_pts, true);
    // landmark gathers (Nystrom)
    std::vector<Copy2D> g;
    std::vector<EvalDesc> ev;
    std::vector<SqMat> kpp;
    std::vector<int> nys, aca;
    for (size_t i = 0; i < R; ++i) {
        if (!hs[i]) continue;
        Hodlr& H = *hs[i];
        const double* px = H.pts.as<double>();
        const double* py = px + H.n;
        const KDev kd = to_kdev(H.kern);
        // the leaves (constructor.jl:39) are assembled when first needed
        // (ensure_leaves); the factorization evaluates leaves of at most
        // LEAF_FUSED_MAX rows itself
        if (H.L > 0 && H.nystrom()) {
            nys.push_back((int)i);
            const int r = H.nys_r;
            double* lx = H.lm.as<double>();
            double* ly = lx + r;
            for (int j = 0; j < r; ++j) {
                g.push_back({px + H.nyind[j], 1, lx + j, 1, 1, 1, 1.0});
                g.push_back({py + H.nyind[j], 1, ly + j, 1, 1, 1, 1.0});
            }
            ev.push_back({lx, ly, lx, ly, r, r, H.kpp.as<double>(), r, 0, 0, kd});
            ev.push_back({px, py, lx, ly, (int)H.n, r, H.kxp.as<double>(), (int)H.n, 0, 0, kd});
            kpp.push_back({H.kpp.as<double>(), r, r, H.info.as<int>() + (defer ? 1 : 0)});
            H.landmark_check = defer;
        } else if (H.L > 0) {
            aca.push_back((int)i);
        }
    }
    copy2d(c, g);
    eval_blocks(c, ev);
    if (!nys.empty()) {
        potrf(c, kpp);
        if (!defer) collect_info(c, hs, st, LANDMARK_FAIL);
        // Y = K(x,P) K(P,P)^{-1} (constructor.jl / factorizations.jl:92):
        // K(P,P)^{-1} = L^{-T} L^{-1} from two triangular solves on the r x r
        // identity, then one n x r x r product.
        std::vector<DiagOp> eye;
        std::vector<TrsmProb> tr;
        std::vector<GemmProb> yg;
        for (int i : nys) {
            Hodlr& H = *hs[i];
            const int r = H.nys_r;
            H.kinv.alloc(&c, (size_t)r * r * sizeof(double));   // kept: the factorization's leaf stage uses it
            double* ki = H.kinv.as<double>();
            eye.push_back({ki, r, r, 1.0});
            tr.push_back({H.kpp.as<double>(), r, r, ki, r, r});
            yg.push_back({(int)H.n, r, r, H.kxp.as<double>(), (int)H.n, ki, r, H.ynys.as<double>(), (int)H.n, 1.0, 0.0, 0});
        }
        set_identity(c, eye);
        trsm_left(c, false, tr);
        trsm_left(c, true, tr);
        gemm(c, false, false, yg);
    }
    if (!aca.empty()) build_aca_batch(c, hs, st, aca);
    if (!defer) c.sync();
    std::vector<Hodlr*> out(R, nullptr);
    for (size_t i = 0; i < R; ++i)
        if (hs[i] && st[i].code == KMH_OK) out[i] = own[i].release();
    return out;
}

Hodlr* hodlr_build(Ctx& c, const double* pts, bool dev, int64_t n, int d, const kmh_kernel& k,
                   const kmh_build_opts& o) {
    std::vector<Status> st;
    auto hs = hodlr_build_batch(c, {{pts, dev, n, d, k, o}}, st);
    if (st[0].code != KMH_OK) throw Error(st[0].code, st[0].msg);
    return hs[0];
}

// K's leaves (constructor.jl:39) in the handle's leaf storage, for the
// handles whose leaves are not assembled yet.
static void ensure_leaves(Ctx& c, const std::vector<Hodlr*>& hs) {
    std::vector<EvalDesc> ev;
    for (Hodlr* h : hs) {
        Hodlr& H = *h;
        if (H.leaves_ready) continue;
        const double* px = H.pts.as<double>();
        const double* py = px + H.n;
        const KDev kd = to_kdev(H.kern);
        for (int l = 0; l < H.nleaves(); ++l) {
            const int64_t s = H.leaf_off[l];
            const int m = (int)H.leaf_size(l);
            ev.push_back({px + s, py + s, px + s, py + s, m, m, H.leaf_ptr(l), m, 0, 0, kd});
        }
        H.leaves_ready = true;
    }
    eval_blocks(c, ev);
}

// Leaves that the fused leaf kernels take (every leaf at most LEAF_FUSED_MAX rows).
static bool fused_leaves(const Hodlr& H) {
    for (int i = 0; i < H.nleaves(); ++i)
        if (H.leaf_size(i) > LEAF_FUSED_MAX) return false;
    return true;
}

// ---------------------------------------------------------------------------
// factor pieces
// ---------------------------------------------------------------------------
enum { W_INV = 0, W_INVT = 1, W_MUL = 2, W_MULT = 3 };

struct Rhs {
    Hodlr* H;
    double* T;
    int64_t ldT;
    int ncols;
    int lv = 0;     // level whose factors apply (woodbury_batch)
    bool e32 = false;   // T stored in FP32 (the mixed form's panels)
};

// T advanced by e elements of its storage type
static double* tadv(double* T, size_t e, bool f32) { return f32 ? f64p(f32p(T) + e) : T + e; }

static void leaf_trsm_batch(Ctx& c, bool trans, const std::vector<Rhs>& rs) {
    std::vector<TrsmProb> pr;
    for (const auto& x : rs) {
        Hodlr& H = *x.H;
        for (int i = 0; i < H.nleaves(); ++i) {
            const int m = (int)H.leaf_size(i);
            TrsmProb t{H.leaf_ptr(i), m, m, tadv(x.T, H.leaf_off[i], x.e32), x.ncols, (int)x.ldT};
            t.eb = x.e32;
            pr.push_back(t);
        }
    }
    trsm_left(c, trans, pr);
}

// T <- op(L_leaf) T, blockwise (L lower with zero upper triangle)
static void leaf_mul(Hodlr& H, bool trans, double* T, int ncols) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    double* tmp = static_cast<double*>(c.scratch((size_t)n * ncols * sizeof(double)));
    copy2d(c, {{T, (int)n, tmp, (int)n, (int)n, ncols, 1.0}});
    std::vector<GemmProb> g;
    for (int i = 0; i < H.nleaves(); ++i) {
        const int m = (int)H.leaf_size(i);
        const int64_t s = H.leaf_off[i];
        g.push_back({m, ncols, m, H.leaf_ptr(i), m, tmp + s, (int)n, T + s, (int)n, 1.0, 0.0, 0});
    }
    gemm(c, trans, false, g);
    c.scratch_free(tmp);
}

// Apply the level-lv block factors of each handle to its T in place:
// S = M^T T, S2 = op(G or X) S, T -/+= M S2.  One grouped launch per step.
static void woodbury_batch(Ctx& c, int mode, const std::vector<Rhs>& rs) {
    size_t ssz = 0;
    for (const auto& x : rs)
        for (int r : x.H->ranks[x.lv]) ssz += (size_t)2 * r * x.ncols;
    if (ssz == 0) return;
    double* S = static_cast<double*>(c.scratch(2 * ssz * sizeof(double)));
    double* S2 = S + ssz;
    const bool inv = (mode == W_INV || mode == W_INVT);
    const bool tr = (mode == W_INVT || mode == W_MULT);
    const double sg = inv ? -1.0 : 1.0;
    std::vector<GemmProb> tn, op, up;
    size_t off = 0;
    for (const auto& x : rs) {
        Hodlr& H = *x.H;
        const int lv = x.lv;
        const int64_t n = H.n;
        const size_t P = (size_t)H.coff[lv] * n;
        const double* Mb = inv ? H.G.as<double>() : H.X.as<double>();
        const int nc = x.ncols;
        const uint8_t q32 = H.p32(), t32 = x.e32;
        for (size_t k = 0; k < H.blocks[lv].size(); ++k) {
            const auto& b = H.blocks[lv][k];
            const int r = H.ranks[lv][k];
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
            const double* Q1 = H.Pat(P + b.c0);
            const double* Q2 = H.Pat(P + b.a0);
            double* s = S + off;
            double* s2 = S2 + off;
            off += (size_t)2 * r * nc;
            // G and X have zero first r rows (G = I - chol(C)^-1, X = chol(C) - I
            // with chol(C) = [I, 0; Bm^T, L22]).  Untransposed, the first r
            // rows of M * (G M^T t) vanish: only the second child's rows change
            // and only the last r rows of G (X) are applied.  Transposed, the
            // first r columns of G^T (X^T) vanish: the first child's projection
            // Q_V^T t_V is not needed.
            const double* Mk = Mb + H.coff_blk[lv][k];
            GemmProb t2{r, nc, h2, Q2, (int)n, tadv(x.T, b.a0, t32), (int)x.ldT, s + r, 2 * r, 1.0, 0.0, 0};
            t2.ea = q32;
            t2.eb = t32;
            tn.push_back(t2);
            GemmProb u2{h2, nc, r, Q2, (int)n, s2 + r, 2 * r, tadv(x.T, b.a0, t32), (int)x.ldT, sg, 1.0, 0};
            u2.ea = q32;
            u2.ec = t32;
            up.push_back(u2);
            if (!tr) {
                GemmProb t1{r, nc, h1, Q1, (int)n, tadv(x.T, b.c0, t32), (int)x.ldT, s, 2 * r, 1.0, 0.0, 0};
                t1.ea = q32;
                t1.eb = t32;
                tn.push_back(t1);
                op.push_back({r, nc, 2 * r, Mk + r, 2 * r, s, 2 * r, s2 + r, 2 * r, 1.0, 0.0, 0});
            } else {
                op.push_back({2 * r, nc, r, Mk + r, 2 * r, s + r, 2 * r, s2, 2 * r, 1.0, 0.0, 0});
                GemmProb u1{h1, nc, r, Q1, (int)n, s2, 2 * r, tadv(x.T, b.c0, t32), (int)x.ldT, sg, 1.0, 0};
                u1.ea = q32;
                u1.ec = t32;
                up.push_back(u1);
            }
        }
    }
    gemm(c, true, false, tn);
    gemm(c, tr, false, op);
    gemm(c, false, false, up);
    c.scratch_free(S);
}

// ---------------------------------------------------------------------------
// symmetricfactorize!  (symfact.jl:6-50), across handles
// ---------------------------------------------------------------------------

// Per-block slot of the factor stage (doubles): a 2r x w coefficient
// matrix, w the cascade panel width (2r without one), then Bm (r x r).
static size_t level_slot(const Hodlr& H, int r) {
    const size_t w = H.AB ? (size_t)H.ab_w() : (size_t)2 * r;
    return (size_t)2 * r * w + (size_t)r * r;
}

// W_lv^{-1} applied to the shared Nystrom panels [A B] (n x 2r; with a
// carried right-hand side [y A B y], n x (2r + 2)).  With M = blkdiag(Q_V,
// Q_U), V = B rows of the first child and U = A rows of the second child,
// M^T [A B] has the blocks Q_V^T B_c = R_V and Q_U^T A_a = R_U already: only
// Q_V^T A_c and Q_U^T B_a need a product (and Q_V^T y_c, Q_U^T y_a, which sit
// beside them in the panels).  sbs[i] holds, per block of handle i in block
// order, a 2r x w slot (ld 2r, w the panel width) into which the panel QR
// wrote R_V (top, B columns) and R_U (bottom, A columns).
static void woodbury_shared_basis(Ctx& c, const std::vector<Hodlr*>& hs, const std::vector<int>& lvs,
                                  const std::vector<double*>& sbs) {
    size_t ssz = 0;
    for (size_t i = 0; i < hs.size(); ++i)
        for (int r : hs[i]->ranks[lvs[i]]) ssz += (size_t)2 * r * hs[i]->ab_w();
    if (ssz == 0) return;
    double* S2 = static_cast<double*>(c.scratch(ssz * sizeof(double)));
    std::vector<GemmProb> tn, op, up;
    std::vector<Copy2D> ycp;
    size_t off = 0;
    for (size_t i = 0; i < hs.size(); ++i) {
        Hodlr& H = *hs[i];
        const int lv = lvs[i];
        const int64_t n = H.n;
        const size_t P = (size_t)H.coff[lv] * n;
        const double* G = H.G.as<double>();
        const bool e = H.ab32(), eq = H.p32();
        double* AB = H.AB;
        double* s = sbs[i];
        const int w = H.ab_w(), ca = H.ab_a(), cb = H.ab_b();
        for (size_t k = 0; k < H.blocks[lv].size(); ++k) {
            const auto& b = H.blocks[lv][k];
            const int r = H.ranks[lv][k];
            KMH_CHECK(r == H.nys_r, KMH_E_BAD_ARG, "shared-basis panels need the full Nystrom rank");
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
            const double* Q1 = H.Pat(P + b.c0);
            const double* Q2 = H.Pat(P + b.a0);
            double* s2 = S2 + off;   // s = [Q_V^T (y) A_c, R_V; R_U, Q_U^T B_a (y)], 2r x w, ld 2r
            off += (size_t)2 * r * w;
            const int l2 = 2 * r;
            // Q_V^T of the first child's [y] A columns, Q_U^T of the second child's B [y] columns
            GemmProb t1{r, r + ca, h1, Q1, (int)n, tadv(AB, b.c0, e), (int)n, s, l2, 1.0, 0.0, 0};
            GemmProb t2{r, w - cb, h2, Q2, (int)n, tadv(AB, (size_t)n * cb + b.a0, e), (int)n,
                        s + r + (size_t)cb * l2, l2, 1.0, 0.0, 0};
            t1.ea = t2.ea = eq;
            t1.eb = t2.eb = e;
            tn.push_back(t1);
            tn.push_back(t2);
            if (H.carry) {
                // both y columns of s hold M^T y
                ycp.push_back({s + r + (size_t)(w - 1) * l2, l2, s + r, l2, r, 1, 1.0});
                ycp.push_back({s, l2, s + (size_t)(w - 1) * l2, l2, r, 1, 1.0});
            }
            // the first r rows of G are zero: only the second child's rows of
            // the panels change (see woodbury_batch)
            op.push_back({r, w, l2, G + H.coff_blk[lv][k] + r, l2, s, l2, s2 + r, l2, 1.0, 0.0, 0});
            GemmProb u2{h2, w, r, Q2, (int)n, s2 + r, l2, tadv(AB, b.a0, e), (int)n, -1.0, 1.0, 0};
            u2.ea = eq;
            u2.ec = e;
            up.push_back(u2);
            s += level_slot(H, r);
        }
    }
    gemm(c, true, false, tn);
    copy2d(c, ycp);
    gemm(c, false, false, op);
    gemm(c, false, false, up);
    c.scratch_free(S2);
}

// W_top^{-1} on the carried y (column 0 of the cascade panels) at the last
// level, which updates no panels: y_a -= Q_U G_low [Q_V^T y_c; Q_U^T y_a],
// then y (now W^{-1} y) to winvy.
static void carry_last_level(Ctx& c, const std::vector<Hodlr*>& hs, const std::vector<int>& lvs) {
    size_t vsz = 0;
    for (size_t i = 0; i < hs.size(); ++i)
        for (int r : hs[i]->ranks[lvs[i]]) vsz += (size_t)3 * r;
    if (vsz == 0) return;
    double* V = static_cast<double*>(c.scratch(vsz * sizeof(double)));
    std::vector<GemmProb> tn, op, up;
    std::vector<Copy2D> out;
    size_t off = 0;
    for (size_t i = 0; i < hs.size(); ++i) {
        Hodlr& H = *hs[i];
        const int lv = lvs[i];
        const int64_t n = H.n;
        const size_t P = (size_t)H.coff[lv] * n;
        const double* G = H.G.as<double>();
        double* y = H.AB;
        for (size_t k = 0; k < H.blocks[lv].size(); ++k) {
            const auto& b = H.blocks[lv][k];
            const int r = H.ranks[lv][k];
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0), l2 = 2 * r;
            double* v = V + off;   // M^T y (2r), then G_low M^T y (r)
            off += (size_t)3 * r;
            tn.push_back({r, 1, h1, H.Pat(P + b.c0), (int)n, y + b.c0, (int)n, v, l2, 1.0, 0.0, 0});
            tn.push_back({r, 1, h2, H.Pat(P + b.a0), (int)n, y + b.a0, (int)n, v + r, l2, 1.0, 0.0, 0});
            op.push_back({r, 1, l2, G + H.coff_blk[lv][k] + r, l2, v, l2, v + l2, r, 1.0, 0.0, 0});
            up.push_back({h2, 1, r, H.Pat(P + b.a0), (int)n, v + l2, r, y + b.a0, (int)n, -1.0, 1.0, 0});
        }
        H.winvy.alloc(&c, (size_t)n * sizeof(double));
        out.push_back({y, (int)n, H.winvy.as<double>(), (int)n, (int)n, 1, 1.0});
        H.has_winvy = true;
    }
    gemm(c, true, false, tn);
    gemm(c, false, false, op);
    gemm(c, false, false, up);
    copy2d(c, out);
    c.scratch_free(V);
}

// One level per handle (levels may differ across handles).
void factor_level_batch(Ctx& c, const std::vector<Hodlr*>& act, const std::vector<int>& lvs) {
    // Nystrom: this level's panels are row ranges of the cascade panels; the
    // panel QR reads them there and writes Q into the level's slab of P
    // per block a slot (level_slot) for the R factors (laid out as the
    // shared-basis coefficient matrix when a Woodbury update of Nystrom
    // panels follows, else R1 and R2 at its start, ld r), then Bm
    size_t rsz = 0;
    for (size_t ai = 0; ai < act.size(); ++ai)
        for (int r : act[ai]->ranks[lvs[ai]]) rsz += level_slot(*act[ai], r);
    double* Rb = static_cast<double*>(c.scratch(std::max<size_t>(1, rsz) * sizeof(double)));
    std::vector<QrPanel> qp;
    std::vector<Copy2D> qcopy;
    std::vector<GemmProb> bm;
    std::vector<DiagOp> plus1, minus1, eye_s, eye_g;
    std::vector<GemmProb> schur, g21;
    std::vector<Copy2D> lowleft;
    std::vector<SqMat> cm;
    std::vector<double*> cm_out;
    std::vector<TrsmProb> ginv;
    std::vector<Axpby2D> neg;
    std::vector<CapDesc> cap;
    size_t off = 0;
    for (size_t ai = 0; ai < act.size(); ++ai) {
        Hodlr& H = *act[ai];
        const int lv = lvs[ai];
        const int64_t n = H.n;
        const size_t P = (size_t)H.coff[lv] * n;
        const bool e = H.p32(), eab = H.ab32();
        double* X = H.X.as<double>();
        double* G = H.G.as<double>();
        int pb = H.nleaves();
        for (int l = 0; l < lv; ++l) pb += (int)H.blocks[l].size();
        const bool shared = H.AB && lv + 1 < H.L;   // woodbury_shared_basis follows
        for (size_t k = 0; k < H.blocks[lv].size(); ++k) {
            const auto& b = H.blocks[lv][k];
            const int r = H.ranks[lv][k];
            double* slot = Rb + off;
            const int ldr = shared ? 2 * r : r;
            const int ca = H.AB ? H.ab_a() : 0, cb = H.AB ? H.ab_b() : r;   // panel columns of the A and B halves
            double* R1 = shared ? slot + (size_t)cb * 2 * r : slot;      // R_V: top rows, B columns of the slot
            double* R2 = shared ? slot + r + (size_t)ca * 2 * r : slot + (size_t)r * r;   // R_U: bottom rows, A columns
            double* Bm = slot + level_slot(H, r) - (size_t)r * r;
            off += level_slot(H, r);
            // lrsymfact(V, U): V (first child) first (symfact.jl:31, HODLR/src/utils.jl:29-35)
            QrPanel qv{H.Pat(P + b.c0), (int)(b.d1 - b.c0), r, (int)n, R1, ldr};
            QrPanel qu{H.Pat(P + b.a0), (int)(b.b1 - b.a0), r, (int)n, R2, ldr};
            qv.e32 = qu.e32 = e;
            if (H.AB && e == eab) {   // V: B half of the first child's rows; U: A half of the second child's
                qv.src = tadv(H.AB, (size_t)n * cb + b.c0, e);
                qv.lds = (int)n;
                qu.src = tadv(H.AB, (size_t)n * ca + b.a0, e);
                qu.lds = (int)n;
            } else if (H.AB) {
                // KMH_MIXED_Q: the FP64 cascade panels rounded into the FP32 slab, QR in place
                Copy2D cv{H.AB + (size_t)n * cb + b.c0, (int)n, qv.A, (int)n, qv.m, r, 1.0};
                Copy2D cu{H.AB + (size_t)n * ca + b.a0, (int)n, qu.A, (int)n, qu.m, r, 1.0};
                cv.ed = cu.ed = 1;
                qcopy.push_back(cv);
                qcopy.push_back(cu);
            }
            qp.push_back(qv);
            qp.push_back(qu);
            double* Xb = X + H.coff_blk[lv][k];
            double* Gb = G + H.coff_blk[lv][k];
            const int l2 = 2 * r;
            if (r <= CAP_MAX) {
                // the whole capacitance computation in one workgroup (la.h)
                cap.push_back({R1, R2, ldr, r, Xb, Gb, l2, H.parts + pb + (int)k, H.info.as<int>()});
                continue;
            }
            bm.push_back({r, r, r, R1, ldr, R2, ldr, Bm, r, 1.0, 0.0, 0});
            // C = [I, Bm; Bm^T, I] (2r x 2r, Bm = R1 R2^T) has the Cholesky
            // factor [I, 0; Bm^T, L22] with L22 = chol(S), S = I - Bm^T Bm (the
            // first r columns of a blocked Cholesky of C are exact), so only
            // S is factored:
            //   X = chol(C) - I    = [0, 0; Bm^T, L22 - I]
            //   G = I - chol(C)^-1 = [0, 0; L22^-1 Bm^T, I - L22^-1]
            // The Woodbury products read only the last r rows of X and G, so
            // the zero first r rows are neither written nor read.
            double* X21 = Xb + r;
            double* X22 = Xb + r + (size_t)r * l2;
            double* G21 = Gb + r;
            double* G22 = Gb + r + (size_t)r * l2;
            lowleft.push_back({Bm, r, X21, l2, r, r, 1.0});   // X21 = Bm^T
            eye_s.push_back({X22, r, l2, 1.0});              // S = I ...
            schur.push_back({r, r, r, Bm, r, Bm, r, X22, l2, -1.0, 1.0, 0});   // ... - Bm^T Bm
            cm.push_back({X22, r, l2, H.info.as<int>()});
            cm_out.push_back(H.parts + pb + (int)k);
            eye_g.push_back({G22, r, l2, 1.0});
            ginv.push_back({X22, r, l2, G22, r, l2});        // G22 = L22^-1
            g21.push_back({r, r, r, G22, l2, X21, l2, G21, l2, 1.0, 0.0, 0});
            neg.push_back({G22, l2, G22, l2, r, r, 0.0, -1.0});
            plus1.push_back({G22, r, l2, 1.0});
            minus1.push_back({X22, r, l2, -1.0});
        }
    }
    copy2d(c, qcopy);
    qr_panels(c, qp);
    capacitance(c, cap);
    gemm(c, false, true, bm);                    // Bm = R1 R2^T
    copy2d(c, lowleft, true);                    // X21 = Bm^T
    set_identity(c, eye_s);
    gemm(c, true, false, schur);                 // S = I - Bm^T Bm
    potrf(c, cm);                                // L22 = chol(S): logdet C = logdet S
    log_diag_sums(c, cm, cm_out);
    set_identity(c, eye_g);
    trsm_left(c, false, ginv);                   // G22 = L22^-1
    gemm(c, false, false, g21);                  // G21 = L22^-1 Bm^T
    axpby2d(c, neg);
    add_diag(c, plus1);                          // G22 = I - L22^-1
    add_diag(c, minus1);                         // X22 = L22 - I
    // Woodbury update of every higher level (symfact.jl:34-37)
    std::vector<Rhs> wr;
    std::vector<Hodlr*> nys;
    std::vector<int> nys_lv;
    std::vector<double*> nys_r;
    std::vector<Hodlr*> last;
    std::vector<int> last_lv;
    off = 0;
    for (size_t ai = 0; ai < act.size(); ++ai) {
        Hodlr* h = act[ai];
        Hodlr& H = *h;
        const int lv = lvs[ai];
        double* r0 = Rb + off;
        for (int r : H.ranks[lv]) off += level_slot(H, r);
        if (lv + 1 >= H.L) {
            if (H.carry) {
                last.push_back(h);
                last_lv.push_back(lv);
            }
            continue;
        }
        if (H.AB) {
            nys.push_back(h);
            nys_lv.push_back(lv);
            nys_r.push_back(r0);
        } else {
            wr.push_back({h, H.Pat((size_t)H.coff[lv + 1] * H.n), H.n, (int)(H.Rtot - H.coff[lv + 1]), lv, H.p32()});
        }
    }
    woodbury_batch(c, W_INV, wr);
    woodbury_shared_basis(c, nys, nys_lv, nys_r);
    carry_last_level(c, last, last_lv);
    c.scratch_free(Rb);
}

// Leaf stage: storage, leaf Cholesky, leaf solves on the panels.
void factor_begin_batch(Ctx& c, const std::vector<Hodlr*>& hs) {
    c.set_device();
    reset_info(c, hs);
    // Leaves of at most LEAF_FUSED_MAX rows: evaluated, factored and their
    // diagonal-block inverses kept by one fused kernel, then the panels'
    // leaf solves by the second (leaf.h).  Larger leaves are assembled and
    // go through potrf and the blocked triangular solve.
    std::vector<Hodlr*> assemble;
    std::vector<char> fz(hs.size());
    size_t linv_sz = 0;
    for (size_t hi = 0; hi < hs.size(); ++hi) {
        Hodlr& H = *hs[hi];
        fz[hi] = fused_leaves(H);
        if (!fz[hi]) assemble.push_back(hs[hi]);
        else
            for (int i = 0; i < H.nleaves(); ++i) linv_sz += leaf_linv_size((int)H.leaf_size(i));
    }
    ensure_leaves(c, assemble);
    struct Hold {
        Ctx& c;
        void* p;
        ~Hold() { if (p) c.scratch_free(p); }
    } linv{c, linv_sz ? c.scratch(linv_sz * sizeof(double)) : nullptr};
    size_t loff = 0;
    std::vector<SqMat> lm;
    std::vector<double*> lm_out;
    std::vector<LeafChol> lc;
    std::vector<LeafSolve> ls;
    std::vector<Copy2D> abinit;
    std::vector<Rhs> lt;
    std::vector<GemmProb> bhalf;
    std::vector<Copy2D> ycopy;
    for (size_t hi = 0; hi < hs.size(); ++hi) {
        Hodlr* h = hs[hi];
        Hodlr& H = *h;
        H.release_transient();
        H.carry = false;
        H.has_winvy = false;
        H.factor_started = true;
        H.leaves_ready = false;   // the leaf storage receives L
        const int64_t n = H.n;
        const int L = H.L, nl = H.nleaves();
        H.nparts = nl + H.nblocks();
        H.parts = static_cast<double*>(c.scratch((size_t)(H.nparts + 1) * sizeof(double)));
        double* hlinv = static_cast<double*>(linv.p) + loff;   // this handle's diagonal-block inverses
        if (fz[hi]) {
            const double* px = H.pts.as<double>();
            const double* py = px + n;
            const KDev kd = to_kdev(H.kern);
            size_t o = 0;
            for (int i = 0; i < nl; ++i) {
                const int m = (int)H.leaf_size(i);
                const int64_t s = H.leaf_off[i];
                lc.push_back({px + s, py + s, m, H.leaf_ptr(i), hlinv + o, H.parts + i, H.info.as<int>(), kd});
                o += leaf_linv_size(m);
            }
            loff += o;
        } else {
            for (int i = 0; i < nl; ++i) {
                lm.push_back({H.leaf_ptr(i), (int)H.leaf_size(i), (int)H.leaf_size(i), H.info.as<int>()});
                lm_out.push_back(H.parts + i);
            }
        }
        // the panels' leaf solves: fused for FP64 panels beside fused leaves
        auto leaf_solves = [&](double* T, int ncols, bool e32) {
            if (!fz[hi] || e32) {
                lt.push_back({h, T, n, ncols, 0, e32});
                return;
            }
            size_t o = 0;
            for (int i = 0; i < nl; ++i) {
                const int m = (int)H.leaf_size(i);
                ls.push_back({H.leaf_ptr(i), hlinv + o, m, T + H.leaf_off[i], (int)n, ncols});
                o += leaf_linv_size(m);
            }
        };
        // per-block factor storage
        H.coff_blk.assign(L, {});
        size_t xtot = 0;
        for (int l = 0; l < L; ++l)
            for (size_t k = 0; k < H.blocks[l].size(); ++k) {
                H.coff_blk[l].push_back((int64_t)xtot);
                const size_t r2 = 2 * (size_t)H.ranks[l][k];
                xtot += r2 * r2;
            }
        H.X.alloc(&c, std::max<size_t>(1, xtot) * sizeof(double));
        H.G.alloc(&c, std::max<size_t>(1, xtot) * sizeof(double));
        if (L == 0) continue;
        if (H.nystrom()) {
            // Every block's U and V are row ranges of A = K(x,P) and
            // B = K(x,P)K(P,P)^{-1}, and every W_l is block-diagonal over whole
            // nodes, so the transformed panels of all higher levels stay row
            // ranges of W_l^{-1}...W_1^{-1} L^{-1} [A B]: the cascade updates
            // these 2r columns instead of every higher level's panel (same W).
            const int r = H.nys_r;
            const bool e = H.ab32();
            H.P.alloc(&c, (size_t)n * H.Rtot * (H.p32() ? sizeof(float) : sizeof(double)));
            // L^{-1} B = (L^{-1} A) K(P,P)^{-1} (B = A K(P,P)^{-1} row by row,
            // and L^{-1} acts on rows): the leaf solves take the A half only,
            // one n x r x r product gives the B half
            H.carry = (H.carry_dev || H.carry_host) && !H.p32();
            const int w = H.ab_w(), ca = H.ab_a(), cb = H.ab_b();
            H.AB = static_cast<double*>(c.scratch((size_t)n * w * (e ? sizeof(float) : sizeof(double))));
            Copy2D a{H.kxp.as<double>(), (int)n, tadv(H.AB, (size_t)n * ca, e), (int)n, (int)n, r, 1.0};
            a.ed = e;
            abinit.push_back(a);
            if (H.carry) {
                if (H.carry_dev) abinit.push_back({H.carry_dev, (int)n, H.AB, (int)n, (int)n, 1, 1.0});
                else KMH_HIP(hipMemcpyAsync(H.AB, H.carry_host, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c.q()));
                ycopy.push_back({H.AB, (int)n, H.AB + (size_t)n * (w - 1), (int)n, (int)n, 1, 1.0});
            }
            leaf_solves(H.AB, r + ca, e);   // [y] and the K(x,P) half
            GemmProb bh{(int)n, r, r, tadv(H.AB, (size_t)n * ca, e), (int)n, H.kinv.as<double>(), r,
                        tadv(H.AB, (size_t)n * cb, e), (int)n, 1.0, 0.0, 0};
            bh.ea = bh.ec = e;
            bhalf.push_back(bh);
        } else {
            if (H.p32()) {
                // the mixed form keeps U/V beside the Q slabs for the refinement's matvec
                H.UV.alloc(&c, H.P.bytes);
                Copy2D u{H.P.as<double>(), (int)n, H.UV.as<double>(), (int)n, (int)n, (int)H.Rtot, 1.0};
                u.es = u.ed = 1;
                abinit.push_back(u);
            }
            // leaf inverse on every level's U and V (symfact.jl:21)
            leaf_solves(H.P.as<double>(), (int)H.Rtot, H.p32());
        }
    }
    // leaf Cholesky (symfact.jl:17)
    leaf_chol(c, lc);
    potrf(c, lm);
    log_diag_sums(c, lm, lm_out);
    copy2d(c, abinit);
    leaf_trsm_batch(c, false, lt);   // (symfact.jl:21)
    leaf_solve(c, ls);
    gemm(c, false, false, bhalf);
    copy2d(c, ycopy);   // L^{-1} y into the last panel column too
}

// logdet accumulation, failure flags, release of the transient state.
void factor_end_batch(Ctx& c, const std::vector<Hodlr*>& hs, std::vector<Status>& st) {
    std::vector<SumJob> sj;
    for (Hodlr* h : hs) sj.push_back({h->parts, h->nparts, h->parts + h->nparts});
    sum_vectors(c, sj);
    std::vector<double> sums(hs.size());
    std::vector<const double*> sp;
    for (Hodlr* h : hs) sp.push_back(h->parts + h->nparts);
    st.assign(hs.size(), Status());
    collect_info(c, hs, st, "Cholesky factorization failed.", sp, sums.data());
    std::vector<int*> failed;
    for (size_t j = 0; j < hs.size(); ++j) {
        Hodlr& H = *hs[j];
        H.release_transient();
        H.factor_started = false;
        if (st[j].code != KMH_OK) {
            failed.push_back(H.info.as<int>());
        } else {
            H.logdet = 2.0 * sums[j];
            H.factorized = true;
        }
    }
    zero_ints(c, failed);
}

void hodlr_factorize_batch(Ctx& c, const std::vector<Hodlr*>& hs_in, std::vector<Status>& st) {
    c.set_device();
    st.assign(hs_in.size(), Status());
    std::vector<Hodlr*> hs;
    std::vector<size_t> pos;
    for (size_t i = 0; i < hs_in.size(); ++i) {
        Hodlr* h = hs_in[i];
        if (!h) { st[i].code = KMH_E_BAD_ARG; st[i].msg = "null HODLR handle"; continue; }
        if (h->factorized) { st[i].code = KMH_E_STATE; st[i].msg = "The matrix is already factorized."; continue; }
        if (h->factor_started) {
            // the panels were transformed in place by the part that ran
            st[i].code = KMH_E_STATE;
            st[i].msg = "A previous factorization of this matrix failed part-way; build it again.";
            continue;
        }
        hs.push_back(h);
        pos.push_back(i);
    }
    if (hs.empty()) return;
    factor_begin_batch(c, hs);
    int maxL = 0;
    for (Hodlr* h : hs) maxL = std::max(maxL, h->L);
    for (int lv = 0; lv < maxL; ++lv) {
        std::vector<Hodlr*> act;
        for (Hodlr* h : hs)
            if (h->L > lv) act.push_back(h);
        factor_level_batch(c, act, std::vector<int>(act.size(), lv));
    }
    std::vector<Status> st2;
    factor_end_batch(c, hs, st2);
    for (size_t j = 0; j < hs.size(); ++j) st[pos[j]] = st2[j];
}

void hodlr_factorize(Hodlr& H) {
    std::vector<Status> st;
    hodlr_factorize_batch(*H.ctx, {&H}, st);
    if (st[0].code != KMH_OK) throw Error(st[0].code, st[0].msg);
}

// ---------------------------------------------------------------------------
// solve and factor applies
// ---------------------------------------------------------------------------

// W^-T W^-1 T in place: the factorization's own solve
static void direct_solve_batch(Ctx& c, const std::vector<Hodlr*>& hs, const std::vector<double*>& T,
                               const std::vector<int>& nrhs) {
    std::vector<Rhs> rs;
    int maxL = 0;
    for (size_t i = 0; i < hs.size(); ++i) {
        KMH_CHECK(hs[i]->factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        rs.push_back({hs[i], T[i], hs[i]->n, nrhs[i]});
        maxL = std::max(maxL, hs[i]->L);
    }
    // W^{-1}: leaves, then levels 1..L; W^{-T}: levels L..1, then leaves
    leaf_trsm_batch(c, false, rs);
    for (int l = 0; l < maxL; ++l) {
        std::vector<Rhs> a;
        for (auto x : rs)
            if (x.H->L > l) { x.lv = l; a.push_back(x); }
        woodbury_batch(c, W_INV, a);
    }
    for (int l = maxL - 1; l >= 0; --l) {
        std::vector<Rhs> a;
        for (auto x : rs)
            if (x.H->L > l) { x.lv = l; a.push_back(x); }
        woodbury_batch(c, W_INVT, a);
    }
    leaf_trsm_batch(c, true, rs);
}

static void operator_matvec(Hodlr& H, const double* x, double* y, int nrhs);

// Iterative refinement of the mixed form's solves against the FP64 HODLR
// operator: x <- x + (W W^T)^{-1} (b - K x) until the largest column's
// relative residual is <= KMH_IR_TOL, stops improving (less than halved) or
// KMH_IR_MAX corrections were made.  The factor W of the mixed form is an
// FP32-accurate factorization of K, so each correction gains several digits.
constexpr double KMH_IR_TOL = 1e-12;
constexpr int KMH_IR_MAX = 8;

static void refine_solve_batch(Ctx& c, const std::vector<Hodlr*>& hs, const std::vector<double*>& T,
                               const std::vector<int>& nrhs) {
    const size_t N = hs.size();
    std::vector<double*> B0(N), R(N);
    size_t ncol = 0;
    std::vector<size_t> col0(N);
    for (size_t i = 0; i < N; ++i) {
        KMH_CHECK(hs[i]->has_uv(), KMH_E_STATE, "refinement needs the HODLR operator's panels");
        const size_t bytes = (size_t)hs[i]->n * nrhs[i] * sizeof(double);
        B0[i] = static_cast<double*>(c.scratch(bytes));
        R[i] = static_cast<double*>(c.scratch(bytes));
        col0[i] = ncol;
        ncol += nrhs[i];
    }
    double* dots = static_cast<double*>(c.scratch(2 * ncol * sizeof(double)));
    std::vector<Copy2D> keep;
    for (size_t i = 0; i < N; ++i) keep.push_back({T[i], (int)hs[i]->n, B0[i], (int)hs[i]->n, (int)hs[i]->n, nrhs[i], 1.0});
    copy2d(c, keep);
    for (size_t i = 0; i < N; ++i)
        column_dots(c, B0[i], (int)hs[i]->n, B0[i], (int)hs[i]->n, (int)hs[i]->n, nrhs[i], dots + ncol + col0[i]);
    direct_solve_batch(c, hs, T, nrhs);
    std::vector<char> active(N, 1);
    std::vector<double> prev(N, INFINITY), h(2 * ncol);
    for (size_t i = 0; i < N; ++i) hs[i]->ir_iters = 0;
    for (int it = 0;; ++it) {
        // R = b - K x for the active systems
        std::vector<Axpby2D> res;
        for (size_t i = 0; i < N; ++i) {
            if (!active[i]) continue;
            operator_matvec(*hs[i], T[i], R[i], nrhs[i]);
            res.push_back({B0[i], (int)hs[i]->n, R[i], (int)hs[i]->n, (int)hs[i]->n, nrhs[i], -1.0, 1.0});
        }
        axpby2d(c, res);
        for (size_t i = 0; i < N; ++i)
            if (active[i]) column_dots(c, R[i], (int)hs[i]->n, R[i], (int)hs[i]->n, (int)hs[i]->n, nrhs[i], dots + col0[i]);
        KMH_HIP(hipMemcpyAsync(h.data(), dots, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        std::vector<Hodlr*> ah;
        std::vector<double*> aR;
        std::vector<int> an;
        std::vector<Axpby2D> upd;
        for (size_t i = 0; i < N; ++i) {
            if (!active[i]) continue;
            double rel = 0.0;
            for (int j = 0; j < nrhs[i]; ++j) {
                const double bb = h[ncol + col0[i] + j];
                rel = std::max(rel, bb > 0 ? std::sqrt(h[col0[i] + j] / bb) : 0.0);
            }
            hs[i]->ir_relres = rel;
            if (rel <= KMH_IR_TOL || it >= KMH_IR_MAX || rel > 0.5 * prev[i]) {
                active[i] = 0;
                continue;
            }
            prev[i] = rel;
            hs[i]->ir_iters = it + 1;
            ah.push_back(hs[i]);
            aR.push_back(R[i]);
            an.push_back(nrhs[i]);
            upd.push_back({R[i], (int)hs[i]->n, T[i], (int)hs[i]->n, (int)hs[i]->n, nrhs[i], 1.0, 1.0});
        }
        if (ah.empty()) break;
        direct_solve_batch(c, ah, aR, an);
        axpby2d(c, upd);
    }
    c.scratch_free(dots);
    for (size_t i = N; i-- > 0;) {
        c.scratch_free(R[i]);
        c.scratch_free(B0[i]);
    }
}

// K^{-1} T: the factorization's solve, refined to FP64 for the mixed form
void hodlr_solve_batch(Ctx& c, const std::vector<Hodlr*>& hs, const std::vector<double*>& T,
                       const std::vector<int>& nrhs) {
    std::vector<Hodlr*> h64, h32;
    std::vector<double*> t64, t32;
    std::vector<int> n64, n32;
    for (size_t i = 0; i < hs.size(); ++i) {
        KMH_CHECK(hs[i]->factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        if (hs[i]->p32()) { h32.push_back(hs[i]); t32.push_back(T[i]); n32.push_back(nrhs[i]); }
        else { h64.push_back(hs[i]); t64.push_back(T[i]); n64.push_back(nrhs[i]); }
    }
    if (!h64.empty()) direct_solve_batch(c, h64, t64, n64);
    if (!h32.empty()) refine_solve_batch(c, h32, t32, n32);
}

void hodlr_factor_apply_batch(Ctx& c, int op, const std::vector<Hodlr*>& hs, const std::vector<double*>& T,
                              const std::vector<int>& nrhs) {
    KMH_CHECK(op == KMH_WINV || op == KMH_WTINV, KMH_E_BAD_ARG, "batched factor apply: W^-1 or W^-T only");
    std::vector<Rhs> rs;
    int maxL = 0;
    for (size_t i = 0; i < hs.size(); ++i) {
        KMH_CHECK(hs[i]->factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        rs.push_back({hs[i], T[i], hs[i]->n, nrhs[i]});
        maxL = std::max(maxL, hs[i]->L);
    }
    auto at = [&](int l) {
        std::vector<Rhs> a;
        for (auto x : rs)
            if (x.H->L > l) { x.lv = l; a.push_back(x); }
        return a;
    };
    if (op == KMH_WINV) {
        leaf_trsm_batch(c, false, rs);
        for (int l = 0; l < maxL; ++l) woodbury_batch(c, W_INV, at(l));
    } else {
        for (int l = maxL - 1; l >= 0; --l) woodbury_batch(c, W_INVT, at(l));
        leaf_trsm_batch(c, true, rs);
    }
}

void hodlr_factor_apply(Hodlr& H, int op, double* T, int nrhs) {
    KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
    Ctx& c = *H.ctx;
    const std::vector<Rhs> rs{{&H, T, H.n, nrhs, 0}};
    auto at = [&](int l) { return std::vector<Rhs>{{&H, T, H.n, nrhs, l}}; };
    switch (op) {
        case KMH_WINV:
            leaf_trsm_batch(c, false, rs);
            for (int l = 0; l < H.L; ++l) woodbury_batch(c, W_INV, at(l));
            break;
        case KMH_WTINV:
            for (int l = H.L - 1; l >= 0; --l) woodbury_batch(c, W_INVT, at(l));
            leaf_trsm_batch(c, true, rs);
            break;
        case KMH_W:
            for (int l = H.L - 1; l >= 0; --l) woodbury_batch(c, W_MUL, at(l));
            leaf_mul(H, false, T, nrhs);
            break;
        case KMH_WT:
            leaf_mul(H, true, T, nrhs);
            for (int l = 0; l < H.L; ++l) woodbury_batch(c, W_MULT, at(l));
            break;
        default:
            throw Error(KMH_E_BAD_ARG, "unknown factor operation");
    }
}

void hodlr_solve(Hodlr& H, double* T, int nrhs) { hodlr_solve_batch(*H.ctx, {&H}, {T}, {nrhs}); }

void hodlr_matvec(Hodlr& H, const double* x, double* y, int nrhs) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    if (H.factorized && !H.p32()) {
        // W (W^T x) (HODLR/src/baseoverloads.jl:170-177)
        copy2d(c, {{x, (int)n, y, (int)n, (int)n, nrhs, 1.0}});
        hodlr_factor_apply(H, KMH_WT, y, nrhs);
        hodlr_factor_apply(H, KMH_W, y, nrhs);
        return;
    }
    operator_matvec(H, x, y, nrhs);
}

// y = K x from the HODLR matrix's FP64 pieces: the leaves (their Cholesky
// factors L L^T once factorized) and the off-diagonal U V^T (the Nystrom
// panels, or the ACA panels, FP32 in the mixed form), FP64 arithmetic
static void operator_matvec(Hodlr& H, const double* x, double* y, int nrhs) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    if (H.factorized || H.factor_started) {
        copy2d(c, {{x, (int)n, y, (int)n, (int)n, nrhs, 1.0}});
        leaf_mul(H, true, y, nrhs);
        leaf_mul(H, false, y, nrhs);
    } else {
        ensure_leaves(c, {&H});
        std::vector<GemmProb> g;
        for (int i = 0; i < H.nleaves(); ++i) {
            const int m = (int)H.leaf_size(i);
            const int64_t s = H.leaf_off[i];
            g.push_back({m, nrhs, m, H.leaf_ptr(i), m, x + s, (int)n, y + s, (int)n, 1.0, 0.0, 0});
        }
        gemm(c, false, false, g);
    }
    if (H.L == 0) return;
    size_t ssz = 0;
    for (int l = 0; l < H.L; ++l)
        for (int r : H.ranks[l]) ssz += (size_t)2 * r * nrhs;
    double* S = static_cast<double*>(c.scratch(ssz * sizeof(double)));
    std::vector<GemmProb> tn;
    std::vector<std::vector<GemmProb>> up(H.L);
    size_t off = 0;
    for (int l = 0; l < H.L; ++l) {
        for (size_t k = 0; k < H.blocks[l].size(); ++k) {
            const auto& b = H.blocks[l][k];
            const int r = H.ranks[l][k];
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
            const double* U = H.Uptr(l, (int)k);
            const double* V = H.Vptr(l, (int)k);
            double* s1 = S + off;
            double* s2 = s1 + (size_t)r * nrhs;
            off += (size_t)2 * r * nrhs;
            const uint8_t e = H.uv32();
            GemmProb t1{r, nrhs, h2, U, (int)n, x + b.a0, (int)n, s1, r, 1.0, 0.0, 0};
            GemmProb t2{r, nrhs, h1, V, (int)n, x + b.c0, (int)n, s2, r, 1.0, 0.0, 0};
            t1.ea = t2.ea = e;
            tn.push_back(t1);
            tn.push_back(t2);
            // HODLR/src/baseoverloads.jl:166-167
            GemmProb u1{h1, nrhs, r, V, (int)n, s1, r, y + b.c0, (int)n, 1.0, 1.0, 0};
            GemmProb u2{h2, nrhs, r, U, (int)n, s2, r, y + b.a0, (int)n, 1.0, 1.0, 0};
            u1.ea = u2.ea = e;
            up[l].push_back(u1);
            up[l].push_back(u2);
        }
    }
    gemm(c, true, false, tn);
    for (int l = 0; l < H.L; ++l) gemm(c, false, false, up[l]);
    c.scratch_free(S);
}

// copy an h x r column block (ld n on the device) to a dense host array
static void fetch_cols(Ctx& c, const double* src, int64_t ld, int64_t h, int r, double* dst, bool f32 = false) {
    if (h <= 0 || r <= 0) return;
    if (!f32) {
        KMH_HIP(hipMemcpy2DAsync(dst, h * sizeof(double), src, ld * sizeof(double), h * sizeof(double), r,
                                 hipMemcpyDeviceToHost, c.q()));
        return;
    }
    std::vector<float> tmp((size_t)h * r);
    KMH_HIP(hipMemcpy2DAsync(tmp.data(), h * sizeof(float), src, ld * sizeof(float), h * sizeof(float), r,
                             hipMemcpyDeviceToHost, c.q()));
    c.sync();
    for (size_t i = 0; i < tmp.size(); ++i) dst[i] = (double)tmp[i];
}

void hodlr_full_host(Hodlr& H, double* out) {
    KMH_CHECK(!H.factorized, KMH_E_STATE, "full(HK) needs the unfactorized U and V.");
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    ensure_leaves(c, {&H});
    std::vector<double> lv(H.leaves.bytes / sizeof(double));
    KMH_HIP(hipMemcpyAsync(lv.data(), H.leaves.p, H.leaves.bytes, hipMemcpyDeviceToHost, c.q()));
    c.sync();
    for (int i = 0; i < H.nleaves(); ++i) {
        const int64_t s = H.leaf_off[i], m = H.leaf_size(i);
        const double* Lm = lv.data() + H.leaf_mat_off[i];
        for (int64_t j = 0; j < m; ++j)
            for (int64_t q = 0; q < m; ++q) out[(s + q) + (s + j) * n] = Lm[q + j * m];
    }
    for (int l = 0; l < H.L; ++l) {
        for (size_t k = 0; k < H.blocks[l].size(); ++k) {
            const auto& b = H.blocks[l][k];
            const int r = H.ranks[l][k];
            const int64_t h2 = b.b1 - b.a0, h1 = b.d1 - b.c0;
            std::vector<double> U((size_t)h2 * r), V((size_t)h1 * r);
            fetch_cols(c, H.Uptr(l, (int)k), n, h2, r, U.data(), H.uv32());
            fetch_cols(c, H.Vptr(l, (int)k), n, h1, r, V.data(), H.uv32());
            c.sync();
            for (int64_t j = 0; j < h1; ++j)
                for (int64_t i = 0; i < h2; ++i) {
                    double s = 0.0;
                    for (int t = 0; t < r; ++t) s += U[i + (size_t)t * h2] * V[j + (size_t)t * h1];
                    out[(b.a0 + i) + (b.c0 + j) * n] = s;
                    out[(b.c0 + j) + (b.a0 + i) * n] = s;
                }
        }
    }
}

void hodlr_block_uv_host(Hodlr& H, int level, int block, double* U, double* V) {
    KMH_CHECK(!H.factorized, KMH_E_STATE, "U and V are abandoned after factorization.");
    KMH_CHECK(level >= 1 && level <= H.L, KMH_E_BAD_ARG, "level out of range");
    KMH_CHECK(block >= 0 && block < (int)H.blocks[level - 1].size(), KMH_E_BAD_ARG, "block out of range");
    Ctx& c = *H.ctx;
    const auto& b = H.blocks[level - 1][block];
    const int r = H.ranks[level - 1][block];
    fetch_cols(c, H.Uptr(level - 1, block), H.n, b.b1 - b.a0, r, U, H.uv32());
    fetch_cols(c, H.Vptr(level - 1, block), H.n, b.d1 - b.c0, r, V, H.uv32());
    c.sync();
}

}  // namespace kmh
