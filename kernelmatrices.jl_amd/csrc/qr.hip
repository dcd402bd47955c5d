// Thin QR of batched tall panels on gfx950 (see la.h): Householder QR of
// <=128 x 24 tiles, one wave per tile with the rows in registers, combined by
// a TSQR tree of fan-in 4; columns beyond 24 by block Gram-Schmidt with one
// reorthogonalisation pass whose projections run on the MFMA GEMM.  (The
// tile cost grows with the square of its width: three 24-column blocks of a
// 72-column panel cost less than 32 + 32 + 8.)
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "la.h"

namespace kmh {

// ---------------------------------------------------------------------------
// Householder QR of small tiles (TSQR building block)
// ---------------------------------------------------------------------------
constexpr int QT = 128;   // max rows of a tile: one wave, two rows per lane
constexpr int QC = 24;    // max columns of a tile (wider panels: block Gram-Schmidt;
                          // the Nystrom rank 72 splits into three blocks)
constexpr int QF = 4;     // TSQR fan-in: QF stacked QC x QC factors fill one tile

// A tile is nstack blocks of nbs rows stacked vertically; block i starts at
// src + i * slot (column-major, leading dimension ld).  A leaf tile is one
// block of its panel; an inner tree node stacks its children's R factors.
struct SqrDesc {
    const double* src; int ld; int nstack; int nbs; long long slot;
    double* qout; int ldq;
    double* rout; int ldr;
    int nc, pad;
};

// value of v held by lane `lane` (uniform lane index)
__device__ __forceinline__ double qr_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// One round of column sums over the wave's 128 rows of p(row, k) =
// c(row) a(row, k) for the H columns cb..cb+H-1: lane l ends with the sum of
// column cb + l % H.  Lane l holds rows l and l + 64 (c0, a0 and c1, a1).
// The lanes write their per-lane products (two rows already added) to LDS,
// lane l sums a segment of column l % H, and the segments meet by
// xor-shuffles.  LDS operations of one wave complete in issue order, so no
// barrier is needed.  Fixed summation order.  LDS: 64 x (H + 1) doubles.
template <int NC, int H, int CB>
__device__ __forceinline__ double colsum_round(double c0, const double (&a0)[NC], double c1, const double (&a1)[NC],
                                               double* pw, int lane) {
    constexpr int SEGS = 64 / H, RPS = 64 / SEGS;
#pragma unroll
    for (int k = 0; k < H; ++k) pw[lane * (H + 1) + k] = c0 * a0[CB + k] + c1 * a1[CB + k];
    __builtin_amdgcn_wave_barrier();
    const int k = lane % H, seg = lane / H;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < RPS; i += 2) {
        s0 += pw[(seg * RPS + i) * (H + 1) + k];
        s1 += pw[(seg * RPS + i + 1) * (H + 1) + k];
    }
    double s = s0 + s1;
#pragma unroll
    for (int m = H; m < 64; m <<= 1) s += __shfl_xor(s, m);
    __builtin_amdgcn_wave_barrier();
    return s;
}

// Column sums of the first W columns (W a multiple of 8): lane l < W ends
// with S_l (rounds of 16 columns, then one of 8 if W is odd in eighths).
template <int NC, int W, int CB = 0>
__device__ __forceinline__ double wave_colsum(double c0, const double (&a0)[NC], double c1, const double (&a1)[NC],
                                              double* pw, int lane, double res = 0.0) {
    if constexpr (CB + 16 <= W) {
        const double s = colsum_round<NC, 16, CB>(c0, a0, c1, a1, pw, lane);
        if (lane >= CB && lane < CB + 16) res = s;
        return wave_colsum<NC, W, CB + 16>(c0, a0, c1, a1, pw, lane, res);
    } else if constexpr (CB + 8 <= W) {
        const double s = colsum_round<NC, 8, CB>(c0, a0, c1, a1, pw, lane);
        if (lane >= CB && lane < CB + 8) res = s;
        return res;
    } else {
        return res;
    }
}

template <int N> using ic = std::integral_constant<int, N>;

// Householder QR (LAPACK dgeqr2 + dorg2r arithmetic) of a tile of <= 128
// rows and nc <= NC columns by one wave: lane l holds rows l and l + 64 in
// registers.  The column loop has a fixed body: after step j the registers
// shift left by one, so the pivot column is always a[0] and all register
// indices are static.  The steps run in phases of 8 whose body covers only
// the W columns that can still be non-zero (W = NC, NC - 8, ...).  Pivot-row
// values come from lane j by readlane, column sums from the wave-local
// reduction, and the products w_k of a step reach every lane through LDS.
// The Householder vector of step j is parked in column j of the Q output
// (each lane its own rows) for the backward accumulation of Q, which shifts
// right one column per step (phases of growing W).  Lanes k < nc - j
// write row j of R straight to the R output.  Three waves per SIMD: the tile
// (2 x NC doubles per lane) stays in registers.
template <int NC>
__global__ __launch_bounds__(64, NC <= 24 ? 4 : 3) void k_tile_qr(const SqrDesc* __restrict__ ds) {
    __shared__ double pw[64 * 17];
    __shared__ double rowj[NC];
    __shared__ double taus[NC];
    __shared__ double wsh[NC];   // w_k of the current step, read back as broadcasts
    const SqrDesc d = ds[blockIdx.x];
    const int nc = d.nc, mt = d.nstack * d.nbs;
    const int lane = threadIdx.x;
    const int r0 = lane, r1 = lane + 64;
    double a0[NC], a1[NC];
    {
        // branch-free loads: out-of-range rows and columns read a valid
        // element and are replaced by zero
        const bool in0 = r0 < mt, in1 = r1 < mt;
        const double* s0 = d.src + (in0 ? (size_t)(r0 / d.nbs) * d.slot + (r0 % d.nbs) : 0);
        const double* s1 = d.src + (in1 ? (size_t)(r1 / d.nbs) * d.slot + (r1 % d.nbs) : 0);
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const size_t off = (size_t)(k < nc ? k : 0) * d.ld;
            const double x0 = s0[off], x1 = s1[off];
            a0[k] = (in0 && k < nc) ? x0 : 0.0;
            a1[k] = (in1 && k < nc) ? x1 : 0.0;
        }
    }
    double* q0 = d.qout + (r0 < mt ? r0 : 0);
    double* q1 = d.qout + (r1 < mt ? r1 : 0);

    // factorization steps j0 <= j < j1 with W live columns
    auto fwd = [&](auto wc, int j0, int j1) {
        constexpr int W = decltype(wc)::value;
#pragma unroll 1
        for (int j = j0; j < j1; ++j) {
            // S_k = sum over rows > j of a(row, 0) a(row, k)   (rows of a1 are >= 64 > j)
            const double h0 = r0 > j ? a0[0] : 0.0;
            if (r0 == j) {
#pragma unroll
                for (int k = 0; k < W; ++k) rowj[k] = a0[k];
            }
            const double sk = wave_colsum<NC, W>(h0, a0, a1[0], a1, pw, lane);   // its LDS traffic orders rowj too
            const double rk = rowj[lane < W ? lane : 0];
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
            const double v0 = (r0 > j) ? scal * a0[0] : (r0 == j ? 1.0 : 0.0);
            const double v1 = scal * a1[0];
            const double wl = rk + scal * sk;   // v^T a_k for this lane's column
            if (tau != 0.0) {
                if (lane < W) wsh[lane] = wl;
                __builtin_amdgcn_wave_barrier();
                const double t0 = tau * v0, t1 = tau * v1;
#pragma unroll
                for (int k = 1; k < W; ++k) {
                    const double wk = wsh[k];
                    a0[k] = __fma_rn(-t0, wk, a0[k]);
                    a1[k] = __fma_rn(-t1, wk, a1[k]);
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (r0 < mt) q0[(size_t)j * d.ldq] = v0;
            if (r1 < mt) q1[(size_t)j * d.ldq] = v1;
            // row j of R: lane k holds column j + k (the pivot row's updated
            // entry, the same arithmetic as the pivot lane's update above)
            if (lane < nc - j) {
                const double rv = lane == 0 ? beta : (tau != 0.0 ? __fma_rn(-tau, wl, rk) : rk);
                d.rout[j + (size_t)(j + lane) * d.ldr] = rv;
            }
            if (r0 == j) taus[j] = tau;
#pragma unroll
            for (int k = 0; k < W - 1; ++k) {
                a0[k] = a0[k + 1];
                a1[k] = a1[k + 1];
            }
            a0[W - 1] = 0.0;
            a1[W - 1] = 0.0;
        }
    };
    fwd(ic<NC>{}, 0, min(nc, 8));
    if constexpr (NC > 8) fwd(ic<NC - 8>{}, 8, min(nc, 16));
    if constexpr (NC > 16) fwd(ic<NC - 16>{}, 16, min(nc, 24));
    if constexpr (NC > 24) fwd(ic<NC - 24>{}, 24, min(nc, 32));
    __builtin_amdgcn_wave_barrier();
    for (int e = lane; e < nc * nc; e += 64) {   // zero strict lower triangle of R
        const int i = e % nc, cc = e / nc;
        if (i > cc) d.rout[i + (size_t)cc * d.ldr] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        a0[k] = 0.0;
        a1[k] = 0.0;
    }
    double n0 = (nc > 0 && r0 < mt) ? q0[(size_t)(nc - 1) * d.ldq] : 0.0;
    double n1 = (nc > 0 && r1 < mt) ? q1[(size_t)(nc - 1) * d.ldq] : 0.0;
    // accumulation steps jhi >= j >= jlo with W live columns
    auto bwd = [&](auto wc, int jhi, int jlo) {
        constexpr int W = decltype(wc)::value;
#pragma unroll 1
        for (int j = jhi; j >= jlo; --j) {
            const double v0 = n0, v1 = n1;
            if (j > 0) {
                n0 = (r0 < mt) ? q0[(size_t)(j - 1) * d.ldq] : 0.0;
                n1 = (r1 < mt) ? q1[(size_t)(j - 1) * d.ldq] : 0.0;
            }
#pragma unroll
            for (int k = W - 1; k > 0; --k) {
                a0[k] = a0[k - 1];
                a1[k] = a1[k - 1];
            }
            a0[0] = 0.0;
            a1[0] = 0.0;
            const double tau = taus[j];
            const double dk = wave_colsum<NC, W>(v0, a0, v1, a1, pw, lane);
            if (tau != 0.0) {
                if (lane < W) wsh[lane] = dk;
                __builtin_amdgcn_wave_barrier();
                const double t0 = tau * v0, t1 = tau * v1;
#pragma unroll
                for (int k = 1; k < W; ++k) {
                    const double w = wsh[k];
                    a0[k] -= t0 * w;
                    a1[k] -= t1 * w;
                }
                __builtin_amdgcn_wave_barrier();
            }
            a0[0] = (r0 > j) ? -tau * v0 : (r0 == j ? 1.0 - tau : 0.0);
            a1[0] = -tau * v1;
        }
    };
    // step j touches columns j..nc-1: W >= NC - j
    bwd(ic<8>{}, nc - 1, NC - 8);
    if constexpr (NC > 8) bwd(ic<16>{}, min(nc - 1, NC - 9), NC - 16);
    if constexpr (NC > 16) bwd(ic<24>{}, min(nc - 1, NC - 17), NC - 24);
    if constexpr (NC > 24) bwd(ic<32>{}, min(nc - 1, NC - 25), NC - 32);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        if (k < nc) {
            if (r0 < mt) q0[(size_t)k * d.ldq] = a0[k];
            if (r1 < mt) q1[(size_t)k * d.ldq] = a1[k];
        }
    }
}

// One launch per column class (8, 16, QC = 24) of the descriptors.
static void launch_tile_qr(Ctx& c, const std::vector<SqrDesc>& ds) {
    std::vector<SqrDesc> cls[3];
    for (const auto& d : ds) {
        KMH_CHECK(d.nc <= QC && d.nstack * d.nbs <= QT && d.nstack * d.nbs >= d.nc, KMH_E_BAD_ARG,
                  "tile QR beyond its size limits");
        cls[d.nc <= 8 ? 0 : d.nc <= 16 ? 1 : 2].push_back(d);   // NC = 8, 16, 24
    }
    for (int k = 0; k < 3; ++k) {
        if (cls[k].empty()) continue;
        SqrDesc* dd = c.upload_vec(cls[k]);
        const dim3 g((unsigned)cls[k].size());
        if (k == 0) KMH_LAUNCH(c, k_tile_qr<8>, g, dim3(64), dd);
        else if (k == 1) KMH_LAUNCH(c, k_tile_qr<16>, g, dim3(64), dd);
        else {
            // algorithmic bytes: the tile read, Q and R written
            double by = 0.0;
            if (c.prof)
                for (const auto& d : cls[k]) {
                    const double mt = (double)d.nstack * d.nbs;
                    by += 8.0 * (2.0 * mt * d.nc + (double)d.nc * d.nc);
                }
            KMH_LAUNCH_TIMED(c, KMH_PROF_TILE_QR, by, -1.0, k_tile_qr<QC>, g, dim3(64), dd);
        }
    }
}

struct TsqrPanel {
    double* A; int m; int nb; int lda;
    double* R; int ldr;
    const double* src;   // input panel (null: A itself), leading dimension lds
    int lds;
};

// Thin QR of m x nb panels (nb <= QC, m >= nb) by TSQR: leaf tiles of <= 128
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
            const double* in = P.src ? P.src + t.row0 : a;
            ds.push_back({in, P.src ? P.lds : P.lda, 1, t.rows, 0, a, P.lda,
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
        std::vector<Copy2D> pre;
        for (const auto& p : ps) {
            if (p.r <= c0) continue;
            const int nb = std::min(QC, p.r - c0);
            double* Aj = p.A + (size_t)c0 * p.lda;
            double* Rj = p.R + (size_t)c0 * p.ldr;
            // this column block's input: the source panel on first touch
            const double* sj = p.src ? p.src + (size_t)c0 * p.lds : nullptr;
            if (sj && c0 > 64) {   // beyond the out-of-place update's k range: copy first
                pre.push_back({sj, p.lds, Aj, p.lda, p.m, nb, 1.0});
                sj = nullptr;
            }
            if (c0 > 0) {
                double* t2 = T2 + toff;
                toff += (size_t)c0 * nb;
                p1.push_back({c0, nb, p.m, p.A, p.lda, sj ? sj : Aj, sj ? p.lds : p.lda, Rj, p.ldr, 1.0, 0.0, 0});
                GemmProb g1{p.m, nb, c0, p.A, p.lda, Rj, p.ldr, Aj, p.lda, -1.0, 1.0, 0};
                g1.Cin = sj;
                u1.push_back(g1);
                p2.push_back({c0, nb, p.m, p.A, p.lda, Aj, p.lda, t2, c0, 1.0, 0.0, 0});
                u2.push_back({p.m, nb, c0, p.A, p.lda, t2, c0, Aj, p.lda, -1.0, 1.0, 0});
                acc.push_back({t2, c0, Rj, p.ldr, c0, nb, 1.0, 1.0});
                tp.push_back({Aj, p.m, nb, p.lda, Rj + c0, p.ldr, nullptr, 0});
            } else {
                tp.push_back({Aj, p.m, nb, p.lda, Rj + c0, p.ldr, sj, p.lds});
            }
        }
        copy2d(c, pre);
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
