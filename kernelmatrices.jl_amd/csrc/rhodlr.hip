// RKernelHODLR: the recursive HODLR of a (possibly non-symmetric) kernel
// matrix K(x1, x2), for matvecs only (constructor.jl:130-161,
// structstypes.jl:22-27, baseoverloads.jl:285-305).
//
// The tree halves rows and columns separately at floor(len / 2) (the first
// half is the smaller one) down to the level; every internal node holds ACA
// factors of its K12 (first-half rows x second-half columns) and K21 blocks,
// the last level's diagonal blocks are dense.  At one depth the K12 and K21
// blocks of all nodes cover every row once and every column once, so the
// depth's U factors form one n1-row slab and its V factors one n2-row slab
// (stored as one (n1 + n2)-row buffer: U rows first, then V rows).  The ACA
// of all blocks of all depths runs as one step-parallel AcaRun.
#include <algorithm>
#include <memory>

#include "assemble.h"
#include "hodlr.h"
#include "la.h"

namespace kmh {

struct RNode {
    int64_t r0, r1, c0, c1;
};

struct RHodlr {
    Ctx* ctx = nullptr;
    int64_t n1 = 0, n2 = 0;
    int L = 0;
    kmh_kernel kern{};
    std::vector<std::vector<RNode>> nodes;        // depth 0..L; depth L: dense leaves
    std::vector<int64_t> leaf_off;                // packed leaf storage offsets
    std::vector<std::vector<int>> ranks;          // depth < L: 2 per node (K12, K21)
    std::vector<int> rpad;                        // slab width per depth
    DevBuf pts;                                   // x1, y1 (n1 each), x2, y2 (n2 each)
    DevBuf leaves;
    std::vector<DevBuf> UV;                       // per depth: (n1 + n2) x rpad, ld n1 + n2
};

static int floor_log2_i(int64_t n) {
    int l = -1;
    while (n) { n >>= 1; ++l; }
    return l;
}

// constructor.jl:133-141: LogLevel(k) -> FixedLevel(floor(log2(min size)) - k);
// the level is clamped to [0, floor(log2(min size))] so every node has rows
// and columns.
static int r_levels(int64_t n1, int64_t n2, const kmh_build_opts& o) {
    const int fl = floor_log2_i(std::min(n1, n2));
    const int lv = o.level_kind == KMH_FIXED_LEVEL ? o.level : fl - o.level;
    return std::min(std::max(0, lv), fl);
}

static std::unique_ptr<RHodlr> r_build(Ctx& c, const double* p1, int64_t n1, const double* p2, int64_t n2, int d,
                                       const kmh_kernel& k, const kmh_build_opts& o) {
    KMH_CHECK(p1 && p2 && n1 >= 1 && n2 >= 1 && d == 2, KMH_E_BAD_ARG, "invalid points (n >= 1, d = 2)");
    KMH_CHECK(o.maxrank >= 0 && o.tol >= 0.0, KMH_E_BAD_ARG, "invalid tolerance or maxrank");
    check_kernel(k, false);
    c.set_device();
    auto R = std::make_unique<RHodlr>();
    RHodlr& H = *R;
    H.ctx = &c;
    H.n1 = n1;
    H.n2 = n2;
    H.kern = k;
    H.L = r_levels(n1, n2, o);
    H.nodes.assign(H.L + 1, {});
    H.nodes[0].push_back({0, n1, 0, n2});
    for (int dd = 0; dd < H.L; ++dd)
        for (const RNode& q : H.nodes[dd]) {
            const int64_t rm = q.r0 + (q.r1 - q.r0) / 2, cm = q.c0 + (q.c1 - q.c0) / 2;
            H.nodes[dd + 1].push_back({q.r0, rm, q.c0, cm});
            H.nodes[dd + 1].push_back({rm, q.r1, cm, q.c1});
        }
    // points, SoA
    std::vector<double> soa(2 * (size_t)(n1 + n2));
    for (int64_t i = 0; i < n1; ++i) {
        soa[i] = p1[2 * i];
        soa[n1 + i] = p1[2 * i + 1];
    }
    for (int64_t j = 0; j < n2; ++j) {
        soa[2 * n1 + j] = p2[2 * j];
        soa[2 * n1 + n2 + j] = p2[2 * j + 1];
    }
    H.pts.alloc(&c, soa.size() * sizeof(double));
    KMH_HIP(hipMemcpyAsync(H.pts.p, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c.q()));
    const double* x1 = H.pts.as<double>();
    const double* y1 = x1 + n1;
    const double* x2 = y1 + n1;
    const double* y2 = x2 + n2;
    const KDev kd = to_kdev(k);
    // dense leaves (the last level's diagonal blocks; the whole matrix at level 0)
    const auto& lv = H.nodes[H.L];
    H.leaf_off.assign(lv.size() + 1, 0);
    for (size_t i = 0; i < lv.size(); ++i)
        H.leaf_off[i + 1] = H.leaf_off[i] + (lv[i].r1 - lv[i].r0) * (lv[i].c1 - lv[i].c0);
    H.leaves.alloc(&c, (size_t)H.leaf_off.back() * sizeof(double));
    std::vector<EvalDesc> ev;
    for (size_t i = 0; i < lv.size(); ++i) {
        const RNode& q = lv[i];
        ev.push_back({x1 + q.r0, y1 + q.r0, x2 + q.c0, y2 + q.c0, (int)(q.r1 - q.r0), (int)(q.c1 - q.c0),
                      H.leaves.as<double>() + H.leaf_off[i], (int)(q.r1 - q.r0), 0, 0, kd});
    }
    eval_blocks(c, ev);
    if (H.L == 0) {
        c.sync();
        return R;
    }
    // ACA of K12 and K21 of every internal node (src/factorizations.jl:9-84)
    const int64_t ld = n1 + n2;
    struct Dep { int cap; int maxsz; DevBuf T; };
    std::vector<Dep> deps(H.L);
    int nblk = 0;
    for (int dd = 0; dd < H.L; ++dd) {
        int ms_max = 1;
        for (const RNode& q : H.nodes[dd]) {
            const int64_t rm = q.r0 + (q.r1 - q.r0) / 2, cm = q.c0 + (q.c1 - q.c0) / 2;
            const int64_t h12 = std::min(rm - q.r0, q.c1 - cm), h21 = std::min(q.r1 - rm, cm - q.c0);
            for (int64_t h : {h12, h21})
                ms_max = std::max<int>(ms_max, (int)(o.maxrank == 0 ? h : std::min<int64_t>(o.maxrank, h)));
            nblk += 2;
        }
        deps[dd].maxsz = ms_max;
        deps[dd].cap = std::min(ms_max, 256);
    }
    DevBuf out;
    out.alloc(&c, 2 * sizeof(int) * nblk);
    int* d_rank = out.as<int>();
    int* d_stat = d_rank + nblk;
    std::vector<AcaDesc> ds;
    std::vector<int> b0(H.L + 1, 0);
    for (int dd = 0; dd < H.L; ++dd) {
        deps[dd].T.alloc(&c, (size_t)ld * deps[dd].cap * sizeof(double));
        double* T = deps[dd].T.as<double>();
        b0[dd] = (int)ds.size();
        for (const RNode& q : H.nodes[dd]) {
            const int64_t rm = q.r0 + (q.r1 - q.r0) / 2, cm = q.c0 + (q.c1 - q.c0) / 2;
            // K12: rows r0..rm, cols cm..c1 ; K21: rows rm..r1, cols c0..cm
            const int64_t blk[2][4] = {{q.r0, rm, cm, q.c1}, {rm, q.r1, q.c0, cm}};
            for (const auto& b : blk) {
                const int nro = (int)(b[1] - b[0]), ncl = (int)(b[3] - b[2]);
                const int h = std::min(nro, ncl);
                const int ms = o.maxrank == 0 ? h : std::min(o.maxrank, h);
                const int bi = (int)ds.size();
                ds.push_back({x1, y1, x2, y2, (int)b[0], nro, (int)b[2], ncl, T + b[0], T + n1 + b[2], (int)ld, ms,
                              deps[dd].cap, 0, d_rank + bi, d_stat + bi, o.tol, kd});
            }
        }
    }
    b0[H.L] = (int)ds.size();
    std::vector<int> h_out(2 * nblk);
    {
        AcaRun run(c, ds);
        for (;;) {
            run.run();
            KMH_HIP(hipMemcpyAsync(h_out.data(), d_rank, 2 * sizeof(int) * nblk, hipMemcpyDeviceToHost, c.q()));
            c.sync();
            bool again = false;
            for (int dd = 0; dd < H.L; ++dd) {
                bool wider = false;
                for (int b = b0[dd]; b < b0[dd + 1]; ++b) {
                    KMH_CHECK(h_out[nblk + b] != 1, KMH_E_ACA, "Incompatible maxrank or matrix dimensions.");
                    if (h_out[nblk + b] == 2) wider = true;
                }
                if (!wider) continue;
                Dep& dp = deps[dd];
                KMH_CHECK(dp.cap < dp.maxsz, KMH_E_DEVICE, "ACA capacity logic error");
                const int cap = std::min(dp.maxsz, dp.cap * 2);
                KMH_CHECK(cap <= 4096, KMH_E_BAD_ARG, "ACA rank exceeds the supported maximum of 4096");
                DevBuf wide;
                wide.alloc(&c, (size_t)ld * cap * sizeof(double));
                KMH_HIP(hipMemcpyAsync(wide.p, dp.T.p, (size_t)ld * dp.cap * sizeof(double), hipMemcpyDeviceToDevice,
                                       c.q()));
                dp.T = std::move(wide);
                dp.cap = cap;
                double* T = dp.T.as<double>();
                for (int b = b0[dd]; b < b0[dd + 1]; ++b)
                    if (h_out[nblk + b] == 2) run.widen(b, T + ds[b].a, T + n1 + ds[b].c, cap);
                again = true;
            }
            if (!again) break;
        }
    }
    // ranks and compaction into (n1 + n2) x rpad slabs
    H.ranks.assign(H.L, {});
    H.rpad.assign(H.L, 1);
    H.UV.resize(H.L);
    std::vector<Copy2D> cp;
    for (int dd = 0; dd < H.L; ++dd) {
        for (int b = b0[dd]; b < b0[dd + 1]; ++b) {
            H.ranks[dd].push_back(h_out[b]);
            H.rpad[dd] = std::max(H.rpad[dd], h_out[b]);
        }
        H.UV[dd].alloc(&c, (size_t)ld * H.rpad[dd] * sizeof(double));
        double* Dst = H.UV[dd].as<double>();
        const double* T = deps[dd].T.as<double>();
        for (int b = b0[dd]; b < b0[dd + 1]; ++b) {
            const AcaDesc& a = ds[b];
            const int r = h_out[b];
            cp.push_back({T + a.a, (int)ld, Dst + a.a, (int)ld, a.nro, r, 1.0});
            cp.push_back({T + n1 + a.c, (int)ld, Dst + n1 + a.c, (int)ld, a.ncl, r, 1.0});
        }
    }
    copy2d(c, cp);
    c.sync();
    return R;
}

// y = R x (baseoverloads.jl:299-303), x n2 x k and y n1 x k on the device
// (ld n2 and n1): dense leaves, then V^T x for every block of every depth,
// then U t depth by depth (the row ranges of one depth are disjoint).
static void r_matvec(RHodlr& H, const double* x, double* y, int k) {
    Ctx& c = *H.ctx;
    const int64_t n1 = H.n1, n2 = H.n2, ld = n1 + n2;
    std::vector<GemmProb> lf;
    const auto& lv = H.nodes[H.L];
    for (size_t i = 0; i < lv.size(); ++i) {
        const RNode& q = lv[i];
        const int m = (int)(q.r1 - q.r0), kk = (int)(q.c1 - q.c0);
        lf.push_back({m, k, kk, H.leaves.as<double>() + H.leaf_off[i], m, x + q.c0, (int)n2, y + q.r0, (int)n1, 1.0,
                      0.0, 0});
    }
    gemm(c, false, false, lf);
    if (H.L == 0) return;
    size_t tot = 0;
    for (int dd = 0; dd < H.L; ++dd)
        for (int r : H.ranks[dd]) tot += (size_t)r * k;
    if (tot == 0) return;
    double* Tm = static_cast<double*>(c.scratch(tot * sizeof(double)));
    std::vector<GemmProb> tn;
    std::vector<std::vector<GemmProb>> up(H.L);
    size_t off = 0;
    for (int dd = 0; dd < H.L; ++dd) {
        const double* UV = H.UV[dd].as<double>();
        int b = 0;
        for (const RNode& q : H.nodes[dd]) {
            const int64_t rm = q.r0 + (q.r1 - q.r0) / 2, cm = q.c0 + (q.c1 - q.c0) / 2;
            const int64_t blk[2][4] = {{q.r0, rm, cm, q.c1}, {rm, q.r1, q.c0, cm}};
            for (const auto& bb : blk) {
                const int r = H.ranks[dd][b++];
                if (r == 0) continue;
                double* t = Tm + off;
                off += (size_t)r * k;
                tn.push_back({r, k, (int)(bb[3] - bb[2]), UV + n1 + bb[2], (int)ld, x + bb[2], (int)n2, t, r, 1.0, 0.0,
                              0});
                up[dd].push_back({(int)(bb[1] - bb[0]), k, r, UV + bb[0], (int)ld, t, r, y + bb[0], (int)n1, 1.0, 1.0,
                                  0});
            }
        }
    }
    gemm(c, true, false, tn);
    for (int dd = 0; dd < H.L; ++dd) gemm(c, false, false, up[dd]);
    c.scratch_free(Tm);
}

}  // namespace kmh

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct kmh_rhodlr {
    std::unique_ptr<kmh::RHodlr> h;
    kmh_ctx* owner = nullptr;
    explicit kmh_rhodlr(kmh_ctx* c) : owner(c) { kmh_ctx_hold(c); }
    ~kmh_rhodlr() {
        h.reset();
        kmh_ctx_release(owner);
    }
};

static kmh::RHodlr& R_of(kmh_rhodlr* h) {
    KMH_CHECK(h && h->h, KMH_E_BAD_ARG, "null RKernelHODLR handle");
    h->h->ctx->set_device();
    return *h->h;
}

extern "C" {

int kmh_rhodlr_build(kmh_ctx* ctx, const double* pts1, int64_t n1, const double* pts2, int64_t n2, int32_t d,
                     const kmh_kernel* kernel, const kmh_build_opts* opts, kmh_rhodlr** out) {
    KMH_API_TRY({
        KMH_CHECK(ctx && kernel && opts && out, KMH_E_BAD_ARG, "null argument");
        std::unique_ptr<kmh_rhodlr> h(new kmh_rhodlr(ctx));
        h->h = kmh::r_build(ctx->c, pts1, n1, pts2, n2, d, *kernel, *opts);
        *out = h.release();
    })
}

int kmh_rhodlr_levels(kmh_rhodlr* h, int32_t* levels, int64_t* n1, int64_t* n2) {
    KMH_API_TRY({
        kmh::RHodlr& H = R_of(h);
        if (levels) *levels = H.L;
        if (n1) *n1 = H.n1;
        if (n2) *n2 = H.n2;
    })
}

int kmh_rhodlr_ranks(kmh_rhodlr* h, int32_t* ranks) {
    KMH_API_TRY({
        kmh::RHodlr& H = R_of(h);
        KMH_CHECK(ranks, KMH_E_BAD_ARG, "null output");
        int q = 0;
        for (const auto& rd : H.ranks)
            for (int r : rd) ranks[q++] = r;
    })
}

int kmh_rhodlr_matvec_dev(kmh_rhodlr* h, const double* x, double* y, int64_t nrhs) {
    KMH_API_TRY({
        kmh::RHodlr& H = R_of(h);
        KMH_CHECK(x && y && nrhs >= 1 && nrhs < (int64_t(1) << 20), KMH_E_BAD_ARG, "invalid matvec arguments");
        kmh::r_matvec(H, x, y, (int)nrhs);
    })
}

int kmh_rhodlr_matvec(kmh_rhodlr* h, const double* x, double* y, int64_t nrhs) {
    KMH_API_TRY({
        kmh::RHodlr& H = R_of(h);
        KMH_CHECK(x && y && nrhs >= 1 && nrhs < (int64_t(1) << 20), KMH_E_BAD_ARG, "invalid matvec arguments");
        kmh::Ctx& c = *H.ctx;
        const size_t bx = (size_t)H.n2 * nrhs * sizeof(double), by = (size_t)H.n1 * nrhs * sizeof(double);
        double* dx = static_cast<double*>(c.scratch(bx));
        double* dy = static_cast<double*>(c.scratch(by));
        KMH_HIP(hipMemcpyAsync(dx, x, bx, hipMemcpyHostToDevice, c.q()));
        kmh::r_matvec(H, dx, dy, (int)nrhs);
        KMH_HIP(hipMemcpyAsync(y, dy, by, hipMemcpyDeviceToHost, c.q()));
        c.sync();
        c.scratch_free(dy);
        c.scratch_free(dx);
    })
}

int kmh_rhodlr_destroy(kmh_rhodlr* h) {
    KMH_API_TRY({
        if (!h) return KMH_OK;
        if (h->h) h->h->ctx->set_device();
        delete h;
    })
}

}  // extern "C"
