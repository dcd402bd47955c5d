// Thin QR of batched tall panels on gfx950 (see la.h): Householder QR of
// <=256 x 32 tiles held in registers (a row per thread, butterfly
// reductions across lanes, one barrier per column), combined by a TSQR tree
// of fan-in 8; columns beyond 32 by block Gram-Schmidt with one
// reorthogonalisation pass whose projections run on the MFMA GEMM.
#include <algorithm>
#include <cmath>

#include "la.h"

namespace kmh {

// ---------------------------------------------------------------------------
// Householder QR of small tiles (TSQR building block)
// ---------------------------------------------------------------------------
constexpr int QT = 256;   // max rows of a tile
constexpr int QC = 32;    // max columns of a tile (wider panels: block Gram-Schmidt)
constexpr int QF = 8;     // TSQR fan-in: QF stacked QC x QC factors fill one tile

// A tile is nstack blocks of nbs rows stacked vertically; block i starts at
// src + i * slot (column-major, leading dimension ld).  A leaf tile is one
// block of its panel; an inner tree node stacks its children's R factors.
struct SqrDesc {
    const double* src; int ld; int nstack; int nbs; long long slot;
    double* qout; int ldq;
    double* rout; int ldr;
    int nc, pad;
};

constexpr int ilog2(int p) { return p <= 1 ? 0 : 1 + ilog2(p / 2); }

// Per-wave partial sums over the wave's 64 rows of aj * a[k], k < NC, into
// part[k * 4 + wave].  The lanes write their products to the wave's own LDS
// block pw (64 x (NC+1), padded against bank conflicts); lane l then sums
// column l % NC over a segment of 64 / (64 / NC) rows, and the segments meet
// by xor-shuffles.  LDS operations of one wave complete in issue order, so
// the wave reads its own writes without a block barrier.  Fixed order of
// summation (deterministic).
template <int NC>
__device__ __forceinline__ void tile_reduce(const double (&a)[NC], double aj, double* pw, double* part, int lane,
                                            int wave) {
    constexpr int SEGS = 64 / NC, RPS = 64 / SEGS;
#pragma unroll
    for (int k = 0; k < NC; ++k) pw[lane * (NC + 1) + k] = aj * a[k];
    __builtin_amdgcn_wave_barrier();
    const int k = lane % NC, seg = lane / NC;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < RPS; ++i) s += pw[(seg * RPS + i) * (NC + 1) + k];
#pragma unroll
    for (int m = NC; m < 64; m <<= 1) s += __shfl_xor(s, m);
    if (seg == 0) part[k * 4 + wave] = s;
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double tile_colsum(const double* part, int k) {
    const double* p = part + 4 * k;
    return (p[0] + p[1]) + (p[2] + p[3]);
}

// value of v held by `lane` (a compile-time constant after unrolling)
__device__ __forceinline__ double qr_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Register-resident Householder QR (LAPACK dgeqr2 + dorg2r arithmetic):
// thread t holds row t of the tile.  The column loop runs at run time with a
// fixed body: after step j the registers shift left by one, so the pivot
// column is always a[0] and register indices stay static.  Step j needs
// S_k = sum_{rows > j} a(row, j) a(row, k); the four wave partials meet in
// LDS (one barrier per column), lane k of every wave completes S_k and
// v^T a_k, and readlane hands them to the other lanes.  The Householder
// vectors and R stay in LDS, so no global memory operation is outstanding
// at a barrier; the backward accumulation of Q shifts right one column per
// step, and Q and R are written once at the end.
template <int NC>
__global__ __launch_bounds__(256) void k_tile_qr(const SqrDesc* __restrict__ ds) {
    __shared__ double pw_all[4][64 * (NC + 1)];
    __shared__ double Vs[NC][QT];
    __shared__ double Rs[NC][NC + 1];
    __shared__ __align__(16) double part[2][NC * 4];
    __shared__ double rowj[2][NC];
    __shared__ double taus[NC];
    const SqrDesc d = ds[blockIdx.x];
    const int nc = d.nc, mt = d.nstack * d.nbs;
    const int r = threadIdx.x, lane = r & 63, wave = r >> 6;
    double* pw = pw_all[wave];
    double a[NC];
    {
        const double* src = (r < mt) ? d.src + (size_t)(r / d.nbs) * d.slot + (r % d.nbs) : d.src;
#pragma unroll
        for (int k = 0; k < NC; ++k) a[k] = (r < mt && k < nc) ? src[(size_t)k * d.ld] : 0.0;
    }
#pragma unroll 1
    for (int j = 0; j < nc; ++j) {
        const int buf = j & 1;
        if (r == j) {
#pragma unroll
            for (int k = 0; k < NC; ++k) rowj[buf][k] = a[k];
        }
        tile_reduce<NC>(a, r > j ? a[0] : 0.0, pw, part[buf], lane, wave);
        __syncthreads();
        // lane l finishes column l % NC; the other lanes read it by readlane
        const double sk = tile_colsum(part[buf], lane % NC);
        const double rk = rowj[buf][lane % NC];
        const double alpha = qr_bcast(rk, 0);
        const double x2 = qr_bcast(sk, 0);
        double tau, beta, scal;
        if (x2 == 0.0) {
            tau = 0.0; beta = alpha; scal = 0.0;
        } else {
            beta = -copysign(sqrt(alpha * alpha + x2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        const double vr = (r > j) ? scal * a[0] : (r == j ? 1.0 : 0.0);
        if (tau != 0.0) {
            const double wk = rk + scal * sk;   // v^T a_k for this lane's column
            const double tv = tau * vr;
#pragma unroll
            for (int k = 1; k < NC; ++k) a[k] -= tv * qr_bcast(wk, k);
        }
        Vs[j][r] = vr;
        if (r == j) {
            // row j of R: zeros left of the diagonal, beta, then the updated row
            for (int cc = 0; cc < j; ++cc) Rs[j][cc] = 0.0;
            Rs[j][j] = beta;
#pragma unroll
            for (int k = 1; k < NC; ++k)
                if (j + k < nc) Rs[j][j + k] = a[k];
            taus[j] = tau;
        }
#pragma unroll
        for (int k = 0; k < NC - 1; ++k) a[k] = a[k + 1];
        a[NC - 1] = 0.0;
    }
    __syncthreads();   // V, R, taus visible; the last step's LDS reads are done
    for (int e = r; e < nc * nc; e += 256) {
        const int i = e % nc, cc = e / nc;
        d.rout[i + (size_t)cc * d.ldr] = Rs[i][cc];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) a[k] = 0.0;
#pragma unroll 1
    for (int j = nc - 1; j >= 0; --j) {
        const int buf = j & 1;
        const double vr = Vs[j][r];
#pragma unroll
        for (int k = NC - 1; k > 0; --k) a[k] = a[k - 1];
        a[0] = 0.0;
        const double tau = taus[j];
        tile_reduce<NC>(a, vr, pw, part[buf], lane, wave);
        __syncthreads();
        if (tau != 0.0) {
            const double sk = tile_colsum(part[buf], lane % NC);
            const double tv = tau * vr;
#pragma unroll
            for (int k = 1; k < NC; ++k) a[k] -= tv * qr_bcast(sk, k);
        }
        a[0] = (r > j) ? -tau * vr : (r == j ? 1.0 - tau : 0.0);
    }
    if (r < mt) {
#pragma unroll
        for (int k = 0; k < NC; ++k)
            if (k < nc) d.qout[r + (size_t)k * d.ldq] = a[k];
    }
}

// One launch per column class (8, 16, 32) of the descriptors.
static void launch_tile_qr(Ctx& c, const std::vector<SqrDesc>& ds) {
    std::vector<SqrDesc> cls[3];
    for (const auto& d : ds) {
        KMH_CHECK(d.nc <= QC && d.nstack * d.nbs <= QT, KMH_E_BAD_ARG, "tile QR beyond its size limits");
        cls[d.nc <= 8 ? 0 : d.nc <= 16 ? 1 : 2].push_back(d);
    }
    for (int k = 0; k < 3; ++k) {
        if (cls[k].empty()) continue;
        SqrDesc* dd = c.upload_vec(cls[k]);
        const dim3 g((unsigned)cls[k].size());
        if (k == 0) KMH_LAUNCH(c, k_tile_qr<8>, g, dim3(256), dd);
        else if (k == 1) KMH_LAUNCH(c, k_tile_qr<16>, g, dim3(256), dd);
        else KMH_LAUNCH(c, k_tile_qr<32>, g, dim3(256), dd);
    }
}

struct TsqrPanel {
    double* A; int m; int nb; int lda;
    double* R; int ldr;
};

// Thin QR of m x nb panels (nb <= QC, m >= nb) by TSQR: leaf tiles of <= 256
// rows, then levels of inner nodes each stacking up to QF children's R
// factors; the leaf Q factors are finally multiplied by the product of the
// inner Q blocks on the path to the root.
static void tsqr(Ctx& c, const std::vector<TsqrPanel>& ps) {
    if (ps.empty()) return;
    struct Tile { int panel; int row0; int rows; int node; };
    std::vector<Tile> tiles;
    std::vector<std::vector<int>> cur(ps.size());
    int nnodes = 0;
    for (size_t p = 0; p < ps.size(); ++p) {
        KMH_CHECK(ps[p].nb <= QC && ps[p].m >= ps[p].nb, KMH_E_BAD_ARG, "tsqr: invalid panel shape");
        const int nt = std::max<int>(1, (int)ceil_div(ps[p].m, QT));
        const int base = ps[p].m / nt, rem = ps[p].m % nt;
        int r0 = 0;
        for (int t = 0; t < nt; ++t) {
            const int rows = base + (t < rem ? 1 : 0);
            tiles.push_back({(int)p, r0, rows, nnodes});
            cur[p].push_back(nnodes++);
            r0 += rows;
        }
    }
    int ninternal = 0;
    for (size_t p = 0; p < ps.size(); ++p) {
        int k = (int)cur[p].size();
        while (k > 1) { const int up = (int)ceil_div(k, QF); ninternal += up; k = up; }
    }
    const int total_nodes = nnodes + ninternal;
    const size_t RS = (size_t)QC * QC;          // R slot, ld QC
    const size_t QS = (size_t)QF * QC * QC;     // Q slot of an inner node, (QF QC) x QC, ld QF QC
    double* RT = static_cast<double*>(c.scratch(std::max<size_t>(1, total_nodes) * RS * sizeof(double)));
    double* QTb = static_cast<double*>(c.scratch(std::max(1, ninternal) * QS * sizeof(double)));
    double* CS = static_cast<double*>(c.scratch(std::max<size_t>(1, total_nodes) * RS * sizeof(double)));
    // leaf tile QRs (Q in place)
    {
        std::vector<SqrDesc> ds;
        for (const auto& t : tiles) {
            const auto& P = ps[t.panel];
            const bool single = cur[t.panel].size() == 1;
            double* a = P.A + t.row0;
            ds.push_back({a, P.lda, 1, t.rows, 0, a, P.lda,
                          single ? P.R : RT + (size_t)t.node * RS, single ? P.ldr : QC, P.nb, 0});
        }
        launch_tile_qr(c, ds);
    }
    // inner levels
    struct Node { int panel; int node; int first; int count; int level; };
    std::vector<Node> inner;
    int next_internal = 0, level = 0;
    std::vector<int> root(ps.size(), -1);
    for (size_t p = 0; p < ps.size(); ++p)
        if (cur[p].size() == 1) root[p] = cur[p][0];
    for (;;) {
        std::vector<SqrDesc> ds;
        for (size_t p = 0; p < ps.size(); ++p) {
            auto& L = cur[p];
            if (L.size() <= 1) continue;
            const int nb = ps[p].nb;
            const bool final_level = L.size() <= (size_t)QF;
            std::vector<int> nxt;
            for (size_t i = 0; i < L.size(); i += QF) {
                const int cnt = (int)std::min<size_t>(QF, L.size() - i);
                const int node = nnodes + next_internal;
                const int qi = next_internal++;
                inner.push_back({(int)p, node, L[i], cnt, level});
                if (cnt == 1) {
                    // a lone child passes through: its R is the node's R
                    ds.push_back({RT + (size_t)L[i] * RS, QC, 1, nb, 0, QTb + (size_t)qi * QS, QF * QC,
                                  final_level ? ps[p].R : RT + (size_t)node * RS, final_level ? ps[p].ldr : QC, nb, 0});
                } else {
                    ds.push_back({RT + (size_t)L[i] * RS, QC, cnt, nb, (long long)RS, QTb + (size_t)qi * QS, QF * QC,
                                  final_level ? ps[p].R : RT + (size_t)node * RS, final_level ? ps[p].ldr : QC, nb, 0});
                }
                nxt.push_back(node);
                if (final_level) root[p] = node;
            }
            L = nxt;
        }
        if (ds.empty()) break;
        launch_tile_qr(c, ds);
        ++level;
    }
    // down pass: coefficient of each node (pointer + ld); the root's is I
    std::vector<const double*> cptr(total_nodes, nullptr);
    std::vector<int> cld(total_nodes, 0);
    std::vector<char> is_ident(total_nodes, 0);
    for (size_t p = 0; p < ps.size(); ++p) is_ident[root[p]] = 1;
    for (int lv = level - 1; lv >= 0; --lv) {
        std::vector<GemmProb> gp;
        for (const auto& nd : inner) {
            if (nd.level != lv) continue;
            const int nb = ps[nd.panel].nb;
            const double* Q = QTb + (size_t)(nd.node - nnodes) * QS;
            for (int h = 0; h < nd.count; ++h) {
                const int child = nd.first + h;
                if (is_ident[nd.node]) {
                    cptr[child] = Q + (size_t)h * nb;
                    cld[child] = QF * QC;
                } else {
                    double* dst = CS + (size_t)child * RS;
                    gp.push_back({nb, nb, nb, Q + (size_t)h * nb, QF * QC, cptr[nd.node], cld[nd.node], dst, QC,
                                  1.0, 0.0, 0});
                    cptr[child] = dst;
                    cld[child] = QC;
                }
            }
        }
        gemm(c, false, false, gp);
    }
    // tiles: A_tile <- Q_tile * C_tile (in place; n = nb <= QC keeps each
    // output row block inside one workgroup, which reads it before writing)
    std::vector<GemmProb> gp;
    for (const auto& t : tiles) {
        if (is_ident[t.node]) continue;
        const auto& P = ps[t.panel];
        double* a = P.A + t.row0;
        gp.push_back({t.rows, P.nb, P.nb, a, P.lda, cptr[t.node], cld[t.node], a, P.lda, 1.0, 0.0, 0});
    }
    gemm(c, false, false, gp);
    c.scratch_free(CS);
    c.scratch_free(QTb);
    c.scratch_free(RT);
}

void qr_panels(Ctx& c, const std::vector<QrPanel>& ps) {
    if (ps.empty()) return;
    std::vector<DiagOp> z;
    int rmax = 0;
    for (const auto& p : ps) {
        KMH_CHECK(p.m >= p.r, KMH_E_BAD_ARG, "qr_panels: panel with more columns than rows");
        if (p.r > 0) z.push_back({p.R, p.r, p.ldr, 0.0});
        rmax = std::max(rmax, p.r);
    }
    set_identity(c, z);
    for (int c0 = 0; c0 < rmax; c0 += QC) {
        std::vector<TsqrPanel> tp;
        std::vector<GemmProb> p1, u1, p2, u2;
        std::vector<Axpby2D> acc;
        size_t tsz = 0;
        for (const auto& p : ps)
            if (p.r > c0 && c0 > 0) tsz += (size_t)c0 * std::min(QC, p.r - c0);
        double* T2 = tsz ? static_cast<double*>(c.scratch(tsz * sizeof(double))) : nullptr;
        size_t toff = 0;
        for (const auto& p : ps) {
            if (p.r <= c0) continue;
            const int nb = std::min(QC, p.r - c0);
            double* Aj = p.A + (size_t)c0 * p.lda;
            double* Rj = p.R + (size_t)c0 * p.ldr;
            if (c0 > 0) {
                double* t2 = T2 + toff;
                toff += (size_t)c0 * nb;
                p1.push_back({c0, nb, p.m, p.A, p.lda, Aj, p.lda, Rj, p.ldr, 1.0, 0.0, 0});
                u1.push_back({p.m, nb, c0, p.A, p.lda, Rj, p.ldr, Aj, p.lda, -1.0, 1.0, 0});
                p2.push_back({c0, nb, p.m, p.A, p.lda, Aj, p.lda, t2, c0, 1.0, 0.0, 0});
                u2.push_back({p.m, nb, c0, p.A, p.lda, t2, c0, Aj, p.lda, -1.0, 1.0, 0});
                acc.push_back({t2, c0, Rj, p.ldr, c0, nb, 1.0, 1.0});
            }
            tp.push_back({Aj, p.m, nb, p.lda, Rj + c0, p.ldr});
        }
        if (!p1.empty()) {
            gemm(c, true, false, p1);
            gemm(c, false, false, u1);
            gemm(c, true, false, p2);
            gemm(c, false, false, u2);
            axpby2d(c, acc);
        }
        tsqr(c, tp);
        if (T2) c.scratch_free(T2);
    }
}

}  // namespace kmh
