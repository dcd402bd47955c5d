// HODLR build / symmetric factorization / solve / logdet / matvec on one
// device, composed from the batched primitives of la.h and assemble.h.
//
// Data layout in HBM (column-major):
//   pts    x[0..n) then y[0..n)
//   leaves leaf i (m_i x m_i, ld m_i) at leaf_mat_off[i]; Cholesky in place
//   P      n x Rtot, ld n.  Level l owns columns coff[l] .. coff[l]+rpad[l];
//          block (a0,b1,c0,d1) keeps U in rows a0..b1 and V in rows c0..d1
//          of those columns (ranks below rpad are zero-padded).  Every row
//          belongs to exactly one child per level, so each level's panel is
//          one n x rpad matrix.  The factorization overwrites U and V by the
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

static void check_kernel(const kmh_kernel& k, bool allow_deriv) {
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

static void check_info(Ctx& c, const char* what) {
    int info = 0;
    KMH_HIP(hipMemcpyAsync(&info, c.d_info, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (info) {
        KMH_HIP(hipMemsetAsync(c.d_info, 0, sizeof(int), c.stream));
        throw Error(KMH_E_NOT_PD, std::string("PosDefException: matrix is not positive definite; ") + what);
    }
}

static void build_nystrom(Hodlr& H) {
    Ctx& c = *H.ctx;
    const int r = H.opts.maxrank;
    KMH_CHECK(r >= 2, KMH_E_BAD_ARG, "Please supply a fixed off-diagonal rank that is >=2.");
    int64_t minleaf = H.n;
    for (int i = 0; i < H.nleaves(); ++i) minleaf = std::min(minleaf, H.leaf_size(i) - 1);
    KMH_CHECK(r < minleaf, KMH_E_BAD_ARG, "Your nystrom rank is too big. Reduce the HODLR level or nystrom rank.");
    const int64_t n = H.n;
    // landmarks: Int64.(round.(LinRange(1, n, r)))  (constructor.jl:31)
    H.nyind.resize(r);
    for (int j = 0; j < r; ++j) {
        const double t = (double)j / (double)(r - 1);
        H.nyind[j] = (int64_t)std::nearbyint((1.0 - t) * 1.0 + t * (double)n) - 1;
    }
    H.nys_r = r;
    const double* px = H.pts.as<double>();
    const double* py = px + n;
    H.lm.alloc(&c, 2 * (size_t)r * sizeof(double));
    double* lx = H.lm.as<double>();
    double* ly = lx + r;
    {
        std::vector<Copy2D> g;
        for (int j = 0; j < r; ++j) {
            g.push_back({px + H.nyind[j], 1, lx + j, 1, 1, 1, 1.0});
            g.push_back({py + H.nyind[j], 1, ly + j, 1, 1, 1, 1.0});
        }
        copy2d(c, g);
    }
    H.kpp.alloc(&c, (size_t)r * r * sizeof(double));
    H.kxp.alloc(&c, (size_t)n * r * sizeof(double));
    eval_blocks(c, H.kern, {{lx, ly, lx, ly, r, r, H.kpp.as<double>(), r, 0, 0},
                            {px, py, lx, ly, (int)n, r, H.kxp.as<double>(), (int)n, 0, 0}});
    potrf(c, {{H.kpp.as<double>(), r, r}});
    check_info(c, "Cholesky of the Nystrom landmark matrix K(P,P) failed.");
    // Y^T = K(P,P)^{-1} K(P, x)  (src/factorizations.jl:92)
    double* yt = static_cast<double*>(c.scratch((size_t)r * n * sizeof(double)));
    copy2d(c, {{H.kxp.as<double>(), (int)n, yt, r, (int)n, r, 1.0}}, true);
    trsm_left(c, false, {{H.kpp.as<double>(), r, r, yt, (int)n, r}});
    trsm_left(c, true, {{H.kpp.as<double>(), r, r, yt, (int)n, r}});
    H.ynys.alloc(&c, (size_t)n * r * sizeof(double));
    copy2d(c, {{yt, r, H.ynys.as<double>(), (int)n, r, (int)n, 1.0}}, true);
    H.rpad.assign(H.L, r);
    H.coff.assign(H.L + 1, 0);
    for (int l = 0; l < H.L; ++l) H.coff[l + 1] = H.coff[l] + r;
    H.Rtot = H.coff[H.L];
    H.ranks.assign(H.L, {});
    for (int l = 0; l < H.L; ++l) H.ranks[l].assign(H.blocks[l].size(), r);
    c.scratch_free(yt);
}

static void build_aca(Hodlr& H) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    const double* px = H.pts.as<double>();
    const double* py = px + n;
    const int L = H.L;
    std::vector<int> cap(L), maxsz_lv(L, 1);
    for (int l = 0; l < L; ++l) {
        for (const auto& b : H.blocks[l]) {
            const int h = (int)std::min(b.b1 - b.a0, b.d1 - b.c0);
            const int ms = H.opts.maxrank == 0 ? h : std::min(H.opts.maxrank, h);
            maxsz_lv[l] = std::max(maxsz_lv[l], ms);
        }
        cap[l] = std::min(maxsz_lv[l], 256);
    }
    std::vector<DevBuf> tmp(L);
    int nblk = 0;
    for (int l = 0; l < L; ++l) nblk += (int)H.blocks[l].size();
    DevBuf out;
    out.alloc(&c, 2 * sizeof(int) * std::max(1, nblk));
    int* d_rank = out.as<int>();
    int* d_stat = d_rank + nblk;
    std::vector<int> h_out(2 * std::max(1, nblk));
    std::vector<char> redo(L, 1);
    H.ranks.assign(L, {});
    for (int attempt = 0;; ++attempt) {
        std::vector<AcaDesc> ds;
        int bi = 0;
        for (int l = 0; l < L; ++l) {
            if (!redo[l]) { bi += (int)H.blocks[l].size(); continue; }
            tmp[l].alloc(&c, (size_t)n * cap[l] * sizeof(double));
            double* T = tmp[l].as<double>();
            for (const auto& b : H.blocks[l]) {
                const int nro = (int)(b.b1 - b.a0), ncl = (int)(b.d1 - b.c0);
                const int h = std::min(nro, ncl);
                const int ms = H.opts.maxrank == 0 ? h : std::min(H.opts.maxrank, h);
                ds.push_back({(int)b.a0, nro, (int)b.c0, ncl, T + b.a0, T + b.c0, (int)n, ms, cap[l], 0,
                              d_rank + bi, d_stat + bi});
                ++bi;
            }
        }
        aca_launch(c, H.kern, px, py, H.opts.tol, ds);
        KMH_HIP(hipMemcpyAsync(h_out.data(), d_rank, 2 * sizeof(int) * nblk, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        bool again = false;
        bi = 0;
        for (int l = 0; l < L; ++l) {
            bool lv_redo = false;
            for (size_t k = 0; k < H.blocks[l].size(); ++k, ++bi) {
                if (!redo[l]) continue;
                const int st = h_out[nblk + bi];
                KMH_CHECK(st != 1, KMH_E_ACA, "Incompatible maxrank or matrix dimensions.");
                if (st == 2) lv_redo = true;
            }
            if (redo[l] && lv_redo) {
                KMH_CHECK(cap[l] < maxsz_lv[l], KMH_E_DEVICE, "ACA capacity logic error");
                cap[l] = std::min(maxsz_lv[l], cap[l] * 2);
                KMH_CHECK((size_t)3 * cap[l] * 8 < 96 * 1024, KMH_E_BAD_ARG,
                          "ACA rank exceeds the supported maximum of 4096");
                again = true;
            }
            redo[l] = lv_redo ? 1 : 0;
        }
        if (!again) break;
    }
    // compact into P
    H.rpad.assign(L, 1);
    int bi = 0;
    for (int l = 0; l < L; ++l)
        for (size_t k = 0; k < H.blocks[l].size(); ++k, ++bi) {
            H.ranks[l].push_back(h_out[bi]);
            H.rpad[l] = std::max(H.rpad[l], h_out[bi]);
        }
    H.coff.assign(L + 1, 0);
    for (int l = 0; l < L; ++l) H.coff[l + 1] = H.coff[l] + H.rpad[l];
    H.Rtot = H.coff[L];
    H.P.alloc(&c, (size_t)n * H.Rtot * sizeof(double));
    KMH_HIP(hipMemsetAsync(H.P.p, 0, H.P.bytes, c.stream));
    std::vector<Copy2D> cp;
    double* P = H.P.as<double>();
    for (int l = 0; l < L; ++l) {
        double* Pl = P + (size_t)H.coff[l] * n;
        const double* T = tmp[l].as<double>();
        for (size_t k = 0; k < H.blocks[l].size(); ++k) {
            const auto& b = H.blocks[l][k];
            const int r = H.ranks[l][k];
            cp.push_back({T + b.a0, (int)n, Pl + b.a0, (int)n, (int)(b.b1 - b.a0), r, 1.0});
            cp.push_back({T + b.c0, (int)n, Pl + b.c0, (int)n, (int)(b.d1 - b.c0), r, 1.0});
        }
    }
    copy2d(c, cp);
    c.sync();
}

Hodlr* hodlr_build(Ctx& c, const double* pts, bool dev, int64_t n, int d, const kmh_kernel& k,
                   const kmh_build_opts& o) {
    KMH_CHECK(d == 2, KMH_E_BAD_ARG, "points must be 2-dimensional");
    KMH_CHECK(n >= 1 && n < (int64_t(1) << 31), KMH_E_BAD_ARG, "invalid number of points");
    KMH_CHECK(pts != nullptr, KMH_E_BAD_ARG, "null points");
    check_kernel(k, false);
    KMH_CHECK(o.method == KMH_ACA || o.method == KMH_NYSTROM, KMH_E_BAD_ARG, "unknown compression method");
    KMH_CHECK(o.level_kind == KMH_FIXED_LEVEL || o.level_kind == KMH_LOG_LEVEL, KMH_E_BAD_ARG, "unknown level kind");
    KMH_CHECK(o.precision == KMH_FP64, KMH_E_BAD_ARG, "unsupported precision mode");
    KMH_CHECK(o.maxrank >= 0, KMH_E_BAD_ARG, "maxrank must be >= 0");
    c.set_device();
    std::unique_ptr<Hodlr> H(new Hodlr());
    H->ctx = &c;
    H->n = n;
    H->kern = k;
    H->opts = o;
    tree_indices(*H);
    H->pts.alloc(&c, 2 * (size_t)n * sizeof(double));
    double* px = H->pts.as<double>();
    if (!dev) {
        std::vector<double> soa(2 * (size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            soa[i] = pts[2 * i];
            soa[n + i] = pts[2 * i + 1];
        }
        KMH_HIP(hipMemcpyAsync(px, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
        c.sync();
    } else {
        copy2d(c, {{pts, 2, px, (int)n, 2, (int)n, 1.0}}, true);
    }
    const double* py = px + n;
    // leaves (constructor.jl:39)
    const int nl = H->nleaves();
    H->leaf_mat_off.assign(nl + 1, 0);
    for (int i = 0; i < nl; ++i) H->leaf_mat_off[i + 1] = H->leaf_mat_off[i] + H->leaf_size(i) * H->leaf_size(i);
    H->leaves.alloc(&c, (size_t)H->leaf_mat_off[nl] * sizeof(double));
    {
        std::vector<EvalDesc> ev;
        for (int i = 0; i < nl; ++i) {
            const int64_t s = H->leaf_off[i];
            const int m = (int)H->leaf_size(i);
            ev.push_back({px + s, py + s, px + s, py + s, m, m, H->leaves.as<double>() + H->leaf_mat_off[i], m, 0, 0});
        }
        eval_blocks(c, k, ev);
    }
    if (H->L > 0) {
        if (o.method == KMH_NYSTROM) build_nystrom(*H);
        else build_aca(*H);
    } else {
        H->coff.assign(1, 0);
        H->Rtot = 0;
    }
    c.sync();
    return H.release();
}

// ---------------------------------------------------------------------------
// factor pieces
// ---------------------------------------------------------------------------
enum { W_INV = 0, W_INVT = 1, W_MUL = 2, W_MULT = 3 };

static double* leaf_ptr(Hodlr& H, int i) { return H.leaves.as<double>() + H.leaf_mat_off[i]; }

static void leaf_trsm(Hodlr& H, bool trans, double* T, int64_t ldT, int ncols) {
    std::vector<TrsmProb> pr;
    for (int i = 0; i < H.nleaves(); ++i) {
        const int m = (int)H.leaf_size(i);
        pr.push_back({leaf_ptr(H, i), m, m, T + H.leaf_off[i], ncols, (int)ldT});
    }
    trsm_left(*H.ctx, trans, pr);
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
        g.push_back({m, ncols, m, leaf_ptr(H, i), m, tmp + s, (int)n, T + s, (int)n, 1.0, 0.0, 0});
    }
    gemm(c, trans, false, g);
    c.scratch_free(tmp);
}

// Apply the level-lv block factors to T (n x ncols, ld ldT) in place.
static void woodbury(Hodlr& H, int lv, int mode, double* T, int64_t ldT, int ncols) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    const auto& bl = H.blocks[lv];
    size_t ssz = 0;
    for (size_t k = 0; k < bl.size(); ++k) ssz += (size_t)2 * H.ranks[lv][k] * ncols;
    double* S = static_cast<double*>(c.scratch(2 * ssz * sizeof(double)));
    double* S2 = S + ssz;
    const double* P = H.P.as<double>() + (size_t)H.coff[lv] * n;
    const bool inv = (mode == W_INV || mode == W_INVT);
    const bool tr = (mode == W_INVT || mode == W_MULT);
    const double* Mb = inv ? H.G.as<double>() : H.X.as<double>();
    std::vector<GemmProb> tn, op, up;
    size_t off = 0;
    for (size_t k = 0; k < bl.size(); ++k) {
        const auto& b = bl[k];
        const int r = H.ranks[lv][k];
        const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
        const double* Q1 = P + b.c0;
        const double* Q2 = P + b.a0;
        double* s = S + off;
        double* s2 = S2 + off;
        off += (size_t)2 * r * ncols;
        tn.push_back({r, ncols, h1, Q1, (int)n, T + b.c0, (int)ldT, s, 2 * r, 1.0, 0.0, 0});
        tn.push_back({r, ncols, h2, Q2, (int)n, T + b.a0, (int)ldT, s + r, 2 * r, 1.0, 0.0, 0});
        op.push_back({2 * r, ncols, 2 * r, Mb + H.coff_blk[lv][k], 2 * r, s, 2 * r, s2, 2 * r, 1.0, 0.0, 0});
        const double sg = inv ? -1.0 : 1.0;
        up.push_back({h1, ncols, r, Q1, (int)n, s2, 2 * r, T + b.c0, (int)ldT, sg, 1.0, 0});
        up.push_back({h2, ncols, r, Q2, (int)n, s2 + r, 2 * r, T + b.a0, (int)ldT, sg, 1.0, 0});
    }
    gemm(c, true, false, tn);
    gemm(c, tr, false, op);
    gemm(c, false, false, up);
    c.scratch_free(S);
}

void hodlr_factorize(Hodlr& H) {
    KMH_CHECK(!H.factorized, KMH_E_STATE, "The matrix is already factorized.");
    Ctx& c = *H.ctx;
    c.set_device();
    const int64_t n = H.n;
    const int L = H.L, nl = H.nleaves();
    KMH_HIP(hipMemsetAsync(c.d_info, 0, sizeof(int), c.stream));
    // leaf Cholesky (symfact.jl:17)
    std::vector<SqMat> lm;
    for (int i = 0; i < nl; ++i) lm.push_back({leaf_ptr(H, i), (int)H.leaf_size(i), (int)H.leaf_size(i)});
    potrf(c, lm);
    int nblk = 0;
    for (int l = 0; l < L; ++l) nblk += (int)H.blocks[l].size();
    double* parts = static_cast<double*>(c.scratch((size_t)(nl + nblk + 1) * sizeof(double)));
    log_diag_sums(c, lm, parts);
    // Nystrom: every block's U and V are row ranges of A = K(x,P) and
    // B = K(x,P)K(P,P)^{-1}, and every block factor W_l is block-diagonal over
    // whole nodes, so the transformed panels of all higher levels stay row
    // ranges of W_l^{-1}...W_1^{-1} L^{-1} [A B]: the cascade updates these
    // 2r columns instead of every higher level's panel (same W).
    double* AB = nullptr;
    const int r = H.nys_r;
    if (H.nystrom() && L > 0) {
        H.P.alloc(&c, (size_t)n * H.Rtot * sizeof(double));
        AB = static_cast<double*>(c.scratch((size_t)n * 2 * r * sizeof(double)));
        copy2d(c, {{H.kxp.as<double>(), (int)n, AB, (int)n, (int)n, r, 1.0},
                   {H.ynys.as<double>(), (int)n, AB + (size_t)n * r, (int)n, (int)n, r, 1.0}});
        leaf_trsm(H, false, AB, n, 2 * r);
    } else if (L > 0) {
        // leaf inverse on every level's U and V (symfact.jl:21)
        leaf_trsm(H, false, H.P.as<double>(), n, (int)H.Rtot);
    }
    // per-level factors
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
    double* X = H.X.as<double>();
    double* G = H.G.as<double>();
    int pbase = nl;
    for (int lv = 0; lv < L; ++lv) {
        const auto& bl = H.blocks[lv];
        double* P = H.P.as<double>() + (size_t)H.coff[lv] * n;
        if (AB) {
            // this level's panels: V = B rows (first child), U = A rows (second child)
            std::vector<Copy2D> cp;
            for (const auto& b : bl) {
                cp.push_back({AB + (size_t)n * r + b.c0, (int)n, P + b.c0, (int)n, (int)(b.d1 - b.c0), r, 1.0});
                cp.push_back({AB + b.a0, (int)n, P + b.a0, (int)n, (int)(b.b1 - b.a0), r, 1.0});
            }
            copy2d(c, cp);
        }
        size_t rsz = 0;
        for (size_t k = 0; k < bl.size(); ++k) rsz += (size_t)3 * H.ranks[lv][k] * H.ranks[lv][k];
        double* Rb = static_cast<double*>(c.scratch(rsz * sizeof(double)));
        std::vector<QrPanel> qp;
        std::vector<GemmProb> bm;
        std::vector<DiagOp> ident_x, ident_g, minus1;
        std::vector<Copy2D> lowleft;
        std::vector<SqMat> cm;
        std::vector<TrsmProb> ginv;
        std::vector<Axpby2D> neg;
        size_t off = 0;
        for (size_t k = 0; k < bl.size(); ++k) {
            const auto& b = bl[k];
            const int r = H.ranks[lv][k];
            double* R1 = Rb + off;
            double* R2 = R1 + (size_t)r * r;
            double* Bm = R2 + (size_t)r * r;
            off += (size_t)3 * r * r;
            // lrsymfact(V, U): V (first child) first (symfact.jl:31)
            qp.push_back({P + b.c0, (int)(b.d1 - b.c0), r, (int)n, R1, r});
            qp.push_back({P + b.a0, (int)(b.b1 - b.a0), r, (int)n, R2, r});
            bm.push_back({r, r, r, R1, r, R2, r, Bm, r, 1.0, 0.0, 0});
            double* Xb = X + H.coff_blk[lv][k];
            double* Gb = G + H.coff_blk[lv][k];
            ident_x.push_back({Xb, 2 * r, 2 * r, 1.0});
            ident_g.push_back({Gb, 2 * r, 2 * r, 1.0});
            minus1.push_back({Xb, 2 * r, 2 * r, -1.0});
            lowleft.push_back({Bm, r, Xb + r, 2 * r, r, r, 1.0});
            cm.push_back({Xb, 2 * r, 2 * r});
            ginv.push_back({Xb, 2 * r, 2 * r, Gb, 2 * r, 2 * r});
            neg.push_back({Gb, 2 * r, Gb, 2 * r, 2 * r, 2 * r, 0.0, -1.0});
        }
        qr_panels(c, qp);
        gemm(c, false, true, bm);                    // R1 R2^T
        set_identity(c, ident_x);
        copy2d(c, lowleft, true);                    // C = [I, *; (R1R2^T)^T, I] (lower part)
        potrf(c, cm);                                // chol(C)
        log_diag_sums(c, cm, parts + pbase);
        pbase += (int)bl.size();
        set_identity(c, ident_g);
        trsm_left(c, false, ginv);                   // G = chol(C)^{-1}
        axpby2d(c, neg);
        for (auto& d : ident_g) d.v = 1.0;
        add_diag(c, ident_g);                        // G = I - chol(C)^{-1}
        add_diag(c, minus1);                         // X = chol(C) - I
        c.scratch_free(Rb);
        // Woodbury update of every higher level (symfact.jl:34-37)
        if (lv + 1 < L) {
            if (AB) woodbury(H, lv, W_INV, AB, n, 2 * r);
            else woodbury(H, lv, W_INV, H.P.as<double>() + (size_t)H.coff[lv + 1] * n, n,
                          (int)(H.Rtot - H.coff[lv + 1]));
        }
    }
    if (AB) c.scratch_free(AB);
    double* dsum = parts + nl + nblk;
    sum_vector(c, parts, nl + nblk, dsum);
    double hs = 0.0;
    KMH_HIP(hipMemcpyAsync(&hs, dsum, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    c.scratch_free(parts);
    check_info(c, "Cholesky factorization failed.");
    H.logdet = 2.0 * hs;
    H.factorized = true;
}

void hodlr_factor_apply(Hodlr& H, int op, double* T, int nrhs) {
    KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
    const int64_t n = H.n;
    switch (op) {
        case KMH_WINV:
            leaf_trsm(H, false, T, n, nrhs);
            for (int l = 0; l < H.L; ++l) woodbury(H, l, W_INV, T, n, nrhs);
            break;
        case KMH_WTINV:
            for (int l = H.L - 1; l >= 0; --l) woodbury(H, l, W_INVT, T, n, nrhs);
            leaf_trsm(H, true, T, n, nrhs);
            break;
        case KMH_W:
            for (int l = H.L - 1; l >= 0; --l) woodbury(H, l, W_MUL, T, n, nrhs);
            leaf_mul(H, false, T, nrhs);
            break;
        case KMH_WT:
            leaf_mul(H, true, T, nrhs);
            for (int l = 0; l < H.L; ++l) woodbury(H, l, W_MULT, T, n, nrhs);
            break;
        default:
            throw Error(KMH_E_BAD_ARG, "unknown factor operation");
    }
}

void hodlr_solve(Hodlr& H, double* T, int nrhs) {
    KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
    hodlr_factor_apply(H, KMH_WINV, T, nrhs);
    hodlr_factor_apply(H, KMH_WTINV, T, nrhs);
}

void hodlr_matvec(Hodlr& H, const double* x, double* y, int nrhs) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    if (H.factorized) {
        copy2d(c, {{x, (int)n, y, (int)n, (int)n, nrhs, 1.0}});
        hodlr_factor_apply(H, KMH_WT, y, nrhs);
        hodlr_factor_apply(H, KMH_W, y, nrhs);
        return;
    }
    std::vector<GemmProb> g;
    for (int i = 0; i < H.nleaves(); ++i) {
        const int m = (int)H.leaf_size(i);
        const int64_t s = H.leaf_off[i];
        g.push_back({m, nrhs, m, leaf_ptr(H, i), m, x + s, (int)n, y + s, (int)n, 1.0, 0.0, 0});
    }
    gemm(c, false, false, g);
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
            tn.push_back({r, nrhs, h2, U, (int)n, x + b.a0, (int)n, s1, r, 1.0, 0.0, 0});
            tn.push_back({r, nrhs, h1, V, (int)n, x + b.c0, (int)n, s2, r, 1.0, 0.0, 0});
            // HODLR/src/baseoverloads.jl:166-167
            up[l].push_back({h1, nrhs, r, V, (int)n, s1, r, y + b.c0, (int)n, 1.0, 1.0, 0});
            up[l].push_back({h2, nrhs, r, U, (int)n, s2, r, y + b.a0, (int)n, 1.0, 1.0, 0});
        }
    }
    gemm(c, true, false, tn);
    for (int l = 0; l < H.L; ++l) gemm(c, false, false, up[l]);
    c.scratch_free(S);
}

// copy an h x r column block (ld n on the device) to a dense host array
static void fetch_cols(Ctx& c, const double* src, int64_t ld, int64_t h, int r, double* dst) {
    if (h <= 0 || r <= 0) return;
    KMH_HIP(hipMemcpy2DAsync(dst, h * sizeof(double), src, ld * sizeof(double), h * sizeof(double), r,
                             hipMemcpyDeviceToHost, c.stream));
}

void hodlr_full_host(Hodlr& H, double* out) {
    KMH_CHECK(!H.factorized, KMH_E_STATE, "full(HK) needs the unfactorized U and V.");
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    std::vector<double> lv(H.leaves.bytes / sizeof(double));
    KMH_HIP(hipMemcpyAsync(lv.data(), H.leaves.p, H.leaves.bytes, hipMemcpyDeviceToHost, c.stream));
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
            fetch_cols(c, H.Uptr(l, (int)k), n, h2, r, U.data());
            fetch_cols(c, H.Vptr(l, (int)k), n, h1, r, V.data());
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
    fetch_cols(c, H.Uptr(level - 1, block), H.n, b.b1 - b.a0, r, U);
    fetch_cols(c, H.Vptr(level - 1, block), H.n, b.d1 - b.c0, r, V);
    c.sync();
}

}  // namespace kmh
