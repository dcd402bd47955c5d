// Thin QR of batched tall panels on gfx950 (see la.h): Householder QR of
// <=128 x 24 tiles, one wave per tile with the rows in registers, combined by
// a TSQR tree of fan-in 4, with Q kept in compact-WY form (vectors V and
// coupling Z) until one fused pass per tree level forms Q times the
// coefficient block from above; columns beyond 24 by block Gram-Schmidt with
// one reorthogonalisation pass whose projections run on the MFMA GEMM.  (The
// tile cost grows with the square of its width: three 24-column blocks of a
// 72-column panel cost less than 32 + 32 + 8.)
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "la.h"
#include "wavesum.h"

namespace kmh {

// ---------------------------------------------------------------------------
// Householder QR of small tiles (TSQR building block)
// ---------------------------------------------------------------------------
constexpr int QT = 128;   // tiles of at most QT rows: one wave, two rows per lane
constexpr int QTL = 256;  // max rows of a tile: one wave, four rows per lane
constexpr int QC = 24;    // max columns of a tile (wider panels: block Gram-Schmidt;
                          // the Nystrom rank 72 splits into three blocks)
constexpr int QF = 8;     // TSQR fan-in: QF stacked QC x QC factors fill one tile

// A tile is nstack blocks of nbs rows stacked vertically; block i starts at
// src + i * slot (column-major, leading dimension ld).  A leaf tile is one
// block of its panel; an inner tree node stacks its children's R factors.
struct SqrDesc {
    const double* src; int ld; int nstack; int nbs; long long slot;
    double* qout; int ldq;
    double* rout; int ldr;
    double* zout;     // compact-WY coupling Z, ld QC
    int nc;
    int e32;          // src and qout stored in FP32 (R and Z stay FP64)
    // optional projection removed on load (single-block leaf tiles):
    // tile <- tile - P T, P the tile's rows of kp prior columns (storage as
    // src, ld ldp), T (kp x nc, FP64, ld kp)
    const double* pq = nullptr; int ldp = 0; int kp = 0;
    const double* pt = nullptr;
};

// value of v held by lane `lane` (uniform lane index)
__device__ __forceinline__ double qr_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// One round of column sums over the wave's rows of p(row, k) = c(row)
// a(row, k) for the H columns cb..cb+H-1: lane l ends with the sum of
// column cb + l % H.  Lane l holds rows l + 64 i, i < RPL (c[i], a[i]).
template <int RPL, int NC, int H, int CB>
__device__ __forceinline__ double colsum_round(const double (&c)[RPL], const double (&a)[RPL][NC], double* pw,
                                               int lane) {
#pragma unroll
    for (int k = 0; k < H; ++k) {
        double v = c[0] * a[0][CB + k];
#pragma unroll
        for (int i = 1; i < RPL; ++i) v = __fma_rn(c[i], a[i][CB + k], v);
        pw[lane * (H + 1) + k] = v;
    }
    return lds_colsum_read<H>(pw, lane);
}

// Column sums of the first W columns (W a multiple of 8): lane l < W ends
// with S_l (rounds of 16 columns, then one of 8 if W is odd in eighths).
template <int RPL, int NC, int W, int CB = 0>
__device__ __forceinline__ double wave_colsum(const double (&c)[RPL], const double (&a)[RPL][NC], double* pw, int lane,
                                              double res = 0.0) {
    if constexpr (CB + 16 <= W) {
        const double s = colsum_round<RPL, NC, 16, CB>(c, a, pw, lane);
        if (lane >= CB && lane < CB + 16) res = s;
        return wave_colsum<RPL, NC, W, CB + 16>(c, a, pw, lane, res);
    } else if constexpr (CB + 8 <= W) {
        const double s = colsum_round<RPL, NC, 8, CB>(c, a, pw, lane);
        if (lane >= CB && lane < CB + 8) res = s;
        return res;
    } else {
        return res;
    }
}

template <int N> using ic = std::integral_constant<int, N>;

// Householder QR (LAPACK dgeqr2 arithmetic) of a tile of <= 64 RPL rows and
// nc <= NC columns by one wave: lane l holds rows l + 64 i (i < RPL) in the
// register slots a[i][NC].  The outputs are R, the Householder vectors
// V (explicit: zero above row j, 1 at row j, below it scal times the pivot
// column) and the compact-WY coupling Z = diag(tau) + striu(V^T V) (ld QC),
// from which k_wy_apply forms Q C: with T the triangular factor of
// Q = H_0 ... H_{nc-1} = I - V T V^T, T^-1 = diag(1/tau) + striu(V^T V)
// (Schreiber-Van Loan's recurrence for T, inverted column by column).
//
// The steps run in phases of 8 whose body covers only the W = NC - 8p slots
// [0, W) that can still hold live columns.  Before step j = 8p + s:
//   [0, W - s)   columns j .. (the pivot in slot 0; zero beyond nc)
//   [W - s, W)   v_{8p} .. v_{j-1}, this phase's vectors
//   [W, NC)      the vectors of earlier phases, v_{8p'+q} in NC - 8(p'+1) + q
// and after the step the slots of the window shift left by one, v_j
// entering slot W - 1.  Each step reduces the window against the pivot
// column.  For a column slot k, w_k = r_k + scal s_k (pivot-row entry plus
// scal times the sum over the rows below the pivot) is v^T a_k, the update
// coefficient; for a vector slot the same expression is v_i^T v_j, the Z
// entry of two vectors of one phase, and the update leaves the slot alone.
// The Z entries of vectors of different phases come after the last step
// from the same wave-wide reduction over pair products of the slots.
// Pivot-row values come from lane j through LDS (rowj), column sums from the
// wave-local reduction, and w reaches every lane through LDS (wsh).  Lanes
// k < nc - j write row j of R straight to the R output.  Pivot rows are
// rows < 24 < 64, so only slot row 0 can lie at or above the pivot.
//   RPL = 2 (tiles <= 128 rows): the 24-column tile runs three waves per
//   SIMD; its step loop spills at the 128 registers of four.
//   RPL = 4 (tiles <= 256 rows): two waves per SIMD; half the LDS
//   reduction work per row, and a 256-row panel is one tile.
template <int NC, class E, int RPL>
__global__ __launch_bounds__(64, RPL == 4 ? 2 : (NC <= 16 ? 4 : 3)) void k_tile_qr(const SqrDesc* __restrict__ ds) {
    static_assert(NC % 8 == 0 && NC <= QC, "tile width");
    static_assert(RPL == 2 || RPL == 4, "rows per lane");
    __shared__ double pw[64 * 17];
    __shared__ double rowj[NC];
    __shared__ double wsh[NC];   // w_k of the current step, read back as broadcasts
    const SqrDesc d = ds[blockIdx.x];
    const int nc = d.nc, mt = d.nstack * d.nbs;
    const int lane = threadIdx.x;
    double a[RPL][NC];
    {
        // Branch-free loads: out-of-range rows and columns read a valid
        // element of the tile, and a bit mask (not a select, which the
        // compiler turns into a branch around a load and a wait for it per
        // column) zeroes them.  Global address space: the loads count on
        // the vector-memory counter only, so all of them stay in flight.
        typedef const __attribute__((address_space(1))) E* GSrc;
        GSrc sp[RPL];
        bool in[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int ri = lane + 64 * i;
            in[i] = ri < mt;
            sp[i] = (GSrc)d.src + (in[i] ? (size_t)(ri / d.nbs) * d.slot + (ri % d.nbs) : 0);
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const size_t off = (size_t)(k < nc ? k : 0) * d.ld;
#pragma unroll
            for (int i = 0; i < RPL; ++i) {
                const unsigned long long m = (in[i] && k < nc) ? ~0ull : 0ull;
                a[i][k] = __longlong_as_double(__double_as_longlong((double)sp[i][off]) & m);
            }
        }
    }
    if (d.kp > 0) {
        // the block Gram-Schmidt reorthogonalisation update X -= P T, one
        // prior column l at a time: P's rows of column l and lane k's T(l, k)
        // loaded PD columns ahead (a ring of PD register slots, so that the
        // loads' latency is covered by PD columns of products), T(l, k)
        // reaching every lane by a lane broadcast (two SGPRs live at a time)
        typedef const __attribute__((address_space(1))) E* GSrc;
        const GSrc pl = (GSrc)d.pq + lane;   // row lane + 64 i at pl[64 i]
        const Glob<const double> tg = glob(d.pt);
        auto load_p = [&](int l, double (&q)[RPL]) {   // P's rows of prior column l (clamped to a valid one)
            const size_t lc = (size_t)(l < d.kp ? l : 0) * d.ldp;
#pragma unroll
            for (int i = 0; i < RPL; ++i) {
                const bool ok = lane + 64 * i < mt;
                q[i] = keep_or_zero((double)pl[(ok ? 64 * i : 0) + lc], ok);
            }
        };
        auto load_t = [&](int l) {   // T(l, lane), zero for lanes >= nc
            const bool ok = lane < nc && l < d.kp;
            return keep_or_zero(tg[ok ? (size_t)l + (size_t)lane * d.kp : 0], ok);
        };
        constexpr int PD = 2;
        double qb[PD][RPL], tb[PD];
#pragma unroll
        for (int s = 0; s < PD; ++s) {
            load_p(s, qb[s]);
            tb[s] = load_t(s);
        }
        for (int l0 = 0; l0 < d.kp; l0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
                const int l = l0 + s;
                double q[RPL];
#pragma unroll
                for (int i = 0; i < RPL; ++i) q[i] = qb[s][i];
                const double tl = tb[s];
                load_p(l + PD, qb[s]);   // slot s refilled with column l + PD
                tb[s] = load_t(l + PD);
                if (l < d.kp) {
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        const double tk = qr_bcast(tl, k);
#pragma unroll
                        for (int i = 0; i < RPL; ++i) a[i][k] = __fma_rn(-q[i], tk, a[i][k]);
                    }
                }
            }
        }
    }
    // global-space views of the outputs (generic stores would make every
    // LDS wait of the steps wait for them too)
    Glob<E> qo[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
        const int ri = lane + 64 * i;
        qo[i] = glob(reinterpret_cast<E*>(d.qout)) + (ri < mt ? ri : 0);
    }
    const Glob<double> rout = glob(d.rout), zout = glob(d.zout);

    // the eight factorization steps j0 <= j < j0 + 8 of a phase with window
    // W, unrolled, so the slot shifts are register renaming.  Every step
    // runs (a short last phase's steps past nc see zero columns: tau = 0,
    // the identity) and only the stores stop at nc, so the slots end where
    // a full phase leaves them.
    auto fwd = [&](auto wc, int j0) {
        constexpr int W = decltype(wc)::value;
#pragma unroll
        for (int s8 = 0; s8 < 8; ++s8) {
            const int j = j0 + s8;
            const bool live = j < nc;   // uniform
            const int vb = W - s8;   // first vector slot of this phase
            // S_k = sum over rows > j of a(row, 0) a(row, k)
            double h[RPL];
            h[0] = lane > j ? a[0][0] : 0.0;
#pragma unroll
            for (int i = 1; i < RPL; ++i) h[i] = a[i][0];
            if (lane == j) {
#pragma unroll
                for (int k = 0; k < W; ++k) rowj[k] = a[0][k];
            }
            const double sk = wave_colsum<RPL, NC, W>(h, a, pw, lane);   // its LDS traffic orders rowj too
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
            double v[RPL];
            v[0] = (lane > j) ? scal * a[0][0] : (lane == j ? 1.0 : 0.0);
#pragma unroll
            for (int i = 1; i < RPL; ++i) v[i] = scal * a[i][0];
            const double wl = rk + scal * sk;   // v^T a_k, or v_i^T v_j in a vector slot
            if (tau != 0.0) {
                if (lane < W) wsh[lane] = lane < vb ? wl : 0.0;
                __builtin_amdgcn_wave_barrier();
                double tv[RPL];
#pragma unroll
                for (int i = 0; i < RPL; ++i) tv[i] = tau * v[i];
#pragma unroll
                for (int k = 1; k < W; ++k) {
                    const double wk = wsh[k];
#pragma unroll
                    for (int i = 0; i < RPL; ++i) a[i][k] = __fma_rn(-tv[i], wk, a[i][k]);
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (live) {
#pragma unroll
                for (int i = 0; i < RPL; ++i)
                    if (lane + 64 * i < mt) qo[i][(size_t)j * d.ldq] = (E)v[i];
                // row j of R: lane k holds column j + k (the pivot row's updated
                // entry, the same arithmetic as the pivot lane's update above)
                if (lane < nc - j) {
                    const double rv = lane == 0 ? beta : (tau != 0.0 ? __fma_rn(-tau, wl, rk) : rk);
                    rout[j + (size_t)(j + lane) * d.ldr] = rv;
                }
                // column j of Z: tau on the diagonal, v_i^T v_j of this phase's vectors above it
                const Glob<double> zc = zout + (size_t)j * QC;
                if (lane == 0) zc[j] = tau;
                else if (lane >= vb && lane < W) zc[j0 + lane - vb] = wl;
            }
#pragma unroll
            for (int k = 0; k < W - 1; ++k) {
#pragma unroll
                for (int i = 0; i < RPL; ++i) a[i][k] = a[i][k + 1];
            }
#pragma unroll
            for (int i = 0; i < RPL; ++i) a[i][W - 1] = v[i];
        }
    };
    // Z entries v_i^T v_j, i in phase PI < PJ the phase of j: 16 pairs
    // (i = 8 PI + Q0 + k / 8, j = 8 PJ + k % 8) per reduction round
    auto cross = [&](auto pic, auto pjc, auto qc) {
        constexpr int PI = decltype(pic)::value, PJ = decltype(pjc)::value, Q0 = decltype(qc)::value;
        double pr[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int si = NC - 8 * (PI + 1) + Q0 + k / 8, sj = NC - 8 * (PJ + 1) + k % 8;
            double v = a[0][si] * a[0][sj];
#pragma unroll
            for (int i = 1; i < RPL; ++i) v = __fma_rn(a[i][si], a[i][sj], v);
            pr[k] = v;
        }
        const double g = lds_colsum<16>(pr, pw, lane);
        const int i = 8 * PI + Q0 + (lane & 15) / 8, j = 8 * PJ + (lane & 7);
        if (lane < 16 && j < nc) zout[i + (size_t)j * QC] = g;
    };
    fwd(ic<NC>{}, 0);
    if constexpr (NC > 8)
        if (nc > 8) fwd(ic<NC - 8>{}, 8);
    if constexpr (NC > 16)
        if (nc > 16) fwd(ic<NC - 16>{}, 16);
    if constexpr (NC > 8) {
        if (nc > 8) {
            cross(ic<0>{}, ic<1>{}, ic<0>{});
            cross(ic<0>{}, ic<1>{}, ic<2>{});
            cross(ic<0>{}, ic<1>{}, ic<4>{});
            cross(ic<0>{}, ic<1>{}, ic<6>{});
        }
        if constexpr (NC > 16)
            if (nc > 16) {
                cross(ic<0>{}, ic<2>{}, ic<0>{});
                cross(ic<0>{}, ic<2>{}, ic<2>{});
                cross(ic<0>{}, ic<2>{}, ic<4>{});
                cross(ic<0>{}, ic<2>{}, ic<6>{});
                cross(ic<1>{}, ic<2>{}, ic<0>{});
                cross(ic<1>{}, ic<2>{}, ic<2>{});
                cross(ic<1>{}, ic<2>{}, ic<4>{});
                cross(ic<1>{}, ic<2>{}, ic<6>{});
            }
    }
    __builtin_amdgcn_wave_barrier();
    for (int e = lane; e < nc * nc; e += 64) {   // zero strict lower triangle of R
        const int i = e % nc, cc = e / nc;
        if (i > cc) rout[i + (size_t)cc * d.ldr] = 0.0;
    }
}

// One launch per column class (8, 16, QC = 24), row class (tiles of at
// most QT rows: two rows per lane; more: four) and storage type.
template <class E, int RPL>
static void launch_tile_qr_t(Ctx& c, std::vector<SqrDesc> (&cls)[3]) {
    for (int k = 0; k < 3; ++k) {
        if (cls[k].empty()) continue;
        SqrDesc* dd = c.upload_vec(cls[k]);
        const dim3 g((unsigned)cls[k].size());
        if (k == 0) KMH_LAUNCH(c, (k_tile_qr<8, E, RPL>), g, dim3(64), dd);
        else if (k == 1) KMH_LAUNCH(c, (k_tile_qr<16, E, RPL>), g, dim3(64), dd);
        else {
            // algorithmic bytes: the tile read, V and R written, and the
            // prior columns and coefficients of a fused projection read
            double by = 0.0;
            if (c.prof)
                for (const auto& d : cls[k]) {
                    const double mt = (double)d.nstack * d.nbs;
                    by += 2.0 * sizeof(E) * mt * d.nc + 8.0 * (double)d.nc * d.nc;
                    by += sizeof(E) * mt * d.kp + 8.0 * (double)d.kp * d.nc;
                }
            KMH_LAUNCH_TIMED(c, KMH_PROF_TILE_QR, by, -1.0, (k_tile_qr<QC, E, RPL>), g, dim3(64), dd);
        }
    }
}

static void launch_tile_qr(Ctx& c, const std::vector<SqrDesc>& ds) {
    std::vector<SqrDesc> cls[2][2][3];   // [FP32][more than QT rows][NC = 8, 16, 24]
    for (const auto& d : ds) {
        const int mt = d.nstack * d.nbs;
        KMH_CHECK(d.nc <= QC && mt <= QTL && mt >= d.nc, KMH_E_BAD_ARG, "tile QR beyond its size limits");
        cls[d.e32 ? 1 : 0][mt > QT ? 1 : 0][d.nc <= 8 ? 0 : d.nc <= 16 ? 1 : 2].push_back(d);
    }
    launch_tile_qr_t<double, 2>(c, cls[0][0]);
    launch_tile_qr_t<double, 4>(c, cls[0][1]);
    launch_tile_qr_t<float, 2>(c, cls[1][0]);
    launch_tile_qr_t<float, 4>(c, cls[1][1]);
}

// out = Q C = [C; 0] - V Y with Y = T V1^T C (V1 the top nb rows of V) for
// the Householder vectors V (mt x nb) and coupling Z of k_tile_qr; C null
// means C = I (the explicit Q).  out may be V itself.
struct WyDesc {
    const double* V; int ldv; int mt;
    const double* Z;          // ld QC
    const double* C; int ldc; int nb;
    double* out; int ldo;
    int e32;                  // V and out stored in FP32 (Z and C stay FP64)
};

typedef double qd4 __attribute__((ext_vector_type(4)));

// One 128-thread workgroup per tile.  V1, C and Z go to LDS; X = V1^T C
// (MFMA); Y solves T^-1 Y = X, i.e. y_i = tau_i (x_i - sum_{k>i}
// z_ik y_k), one column per thread with the right-hand side in registers
// (bottom row first; a zero tau gives a zero row, H_i = I).  Then each wave
// forms 16-row blocks of out^T = Cpad^T - Y^T V^T on v_mfma_f64_16x16x4f64
// (first operand -Y^T from LDS, second V^T streamed from memory, the
// accumulator starting from Cpad^T), so that a store instruction writes
// 16-row column segments of out.  A wave reads a row block of V before it
// writes that block, and V1 is in LDS before any block is written, so out
// may alias V.  (Up to QTL = 256 rows: sixteen 16-row blocks, eight per wave.)
template <class E>
__global__ __launch_bounds__(128, 4) void k_wy_apply(const WyDesc* __restrict__ ds) {
    constexpr int S = QC + 1;
    constexpr int KS = QC / 4;   // k steps of 4
    __shared__ double Vs[QC * S];
    __shared__ double Cs[QC * S];
    __shared__ double Zs[QC * S];
    __shared__ double Ys[QC * QC];
    const WyDesc d = ds[blockIdx.x];
    const int nb = d.nb, t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
    // wave w takes row blocks w + 4 p and w + 4 p + 2 in pass p, whose V^T
    // fragments V(row, 4 ks + lk) load at the start of the pass (the first
    // pass's right after the substitution: held through it, they made the
    // substitution's 24 unknowns spill at the 128 registers of four waves
    // per SIMD, and the kernel ran 10 % slower)
    double vf[2][KS];
    auto load_rows = [&](int pass) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = (wave + 4 * pass + 2 * h) * 16 + li;
            const bool rok = row < d.mt;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int k = 4 * ks + lk;
                const double x = (double)__builtin_nontemporal_load(reinterpret_cast<const E*>(d.V) + (rok ? row : 0) +
                                                                    (size_t)(k < nb ? k : 0) * d.ldv);
                vf[h][ks] = (rok && k < nb) ? x : 0.0;
            }
        }
    };
    {
        // V1, C and Z, zero outside nb x nb: all loads in flight before the LDS writes
        constexpr int NE = (QC * QC + 127) / 128;
        double lv[NE], lc[NE], lz[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = t + 128 * q, i = e % QC, k = e / QC;
            const bool in = e < QC * QC && i < nb && k < nb;
            const size_t ii = in ? i : 0, kk = in ? k : 0;
            lv[q] = in ? (double)reinterpret_cast<const E*>(d.V)[ii + kk * d.ldv] : 0.0;
            lc[q] = in ? (d.C ? d.C[ii + kk * d.ldc] : (i == k ? 1.0 : 0.0)) : 0.0;
            lz[q] = (in && i <= k) ? d.Z[ii + kk * QC] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = t + 128 * q, i = e % QC, k = e / QC;
            if (e < QC * QC) {
                Vs[i * S + k] = lv[q];
                Cs[i * S + k] = lc[q];
                Zs[i * S + k] = lz[q];
            }
        }
    }
    __syncthreads();
    {
        // X = V1^T C on v_mfma_f64_16x16x4f64: wave w takes the 16-row block
        // w of X (rows a = 16 w + x), both 16-column blocks; V1^T and C are
        // zero outside nb x nb, and rows / columns beyond QC are not stored
        qd4 xa[2] = {qd4{0.0, 0.0, 0.0, 0.0}, qd4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + lk, a = 16 * wave + li;
            const double va = a < QC ? Vs[k * S + a] : 0.0;     // V1^T(a, k)
#pragma unroll
            for (int yb = 0; yb < 2; ++yb) {
                const int b = 16 * yb + li;
                const double cb = b < QC ? Cs[k * S + b] : 0.0;   // C(k, b)
                xa[yb] = __builtin_amdgcn_mfma_f64_16x16x4f64(va, cb, xa[yb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int yb = 0; yb < 2; ++yb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = 16 * wave + 4 * r + lk, b = 16 * yb + li;
                if (a < QC && b < QC) Ys[a * QC + b] = xa[yb][r];
            }
    }
    __syncthreads();
    if (t < nb) {
        double x[QC];
#pragma unroll
        for (int i = 0; i < QC; ++i) x[i] = Ys[i * QC + t];
#pragma unroll
        for (int i = QC - 1; i >= 0; --i) {
            if (i < nb) {
                const double y = Zs[i * S + i] * x[i];
                x[i] = y;
#pragma unroll
                for (int r = 0; r < i; ++r) x[r] = __fma_rn(-Zs[r * S + i], y, x[r]);
            }
        }
#pragma unroll
        for (int i = 0; i < QC; ++i)
            if (i < nb) Ys[i * QC + t] = x[i];
    }
    __syncthreads();
    load_rows(0);
    double yf[2][KS];   // -Y(4 ks + lk, 16 cb + li)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + lk, col = 16 * cb + li;
            yf[cb][ks] = (k < nb && col < nb) ? -Ys[k * QC + col] : 0.0;
        }
#pragma unroll 1
    for (int pass = 0; pass < QTL / 64; ++pass) {
        if ((wave + 4 * pass) * 16 >= d.mt) break;
        if (pass > 0) load_rows(pass);
        qd4 acc[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = (wave + 4 * pass + 2 * h) * 16 + li;
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * cb + 4 * r + lk;
                    acc[h][cb][r] = (row < nb && col < nb) ? Cs[row * S + col] : 0.0;
                }
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            if (4 * ks < nb) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    acc[h][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(yf[0][ks], vf[h][ks], acc[h][0], 0, 0, 0);
                    acc[h][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(yf[1][ks], vf[h][ks], acc[h][1], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = (wave + 4 * pass + 2 * h) * 16 + li;
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * cb + 4 * r + lk;
                    if (row < d.mt && col < nb)
                        __builtin_nontemporal_store((E)acc[h][cb][r], reinterpret_cast<E*>(d.out) + row + (size_t)col * d.ldo);
                }
        }
    }
}

static void wy_apply(Ctx& c, const std::vector<WyDesc>& ds) {
    if (ds.empty()) return;
    std::vector<WyDesc> by[2];
    for (const auto& d : ds) {
        KMH_CHECK(d.nb <= QC && d.mt <= QTL && d.mt >= d.nb, KMH_E_BAD_ARG, "WY apply beyond its size limits");
        by[d.e32 ? 1 : 0].push_back(d);
    }
    if (!by[0].empty()) {
        WyDesc* dd = c.upload_vec(by[0]);
        KMH_LAUNCH(c, k_wy_apply<double>, dim3((unsigned)by[0].size()), dim3(128), dd);
    }
    if (!by[1].empty()) {
        WyDesc* dd = c.upload_vec(by[1]);
        KMH_LAUNCH(c, k_wy_apply<float>, dim3((unsigned)by[1].size()), dim3(128), dd);
    }
}

struct TsqrPanel {
    double* A; int m; int nb; int lda;
    double* R; int ldr;
    const double* src;   // input panel (null: A itself), leading dimension lds
    int lds;
    int e32;             // A and src stored in FP32
    // optional projection removed from the input on load: input - P T with
    // P (m x kp, same storage, ld ldp) and T (kp x nb, FP64, ld kp)
    const double* pq = nullptr; int ldp = 0; int kp = 0;
    const double* pt = nullptr;
};

// Thin QR of m x nb panels (nb <= QC, m >= nb) by TSQR: leaf tiles of <= 128
// rows, then levels of inner nodes each stacking up to QF children's R
// factors.  Every tile QR leaves its Householder vectors and coupling Z;
// the down pass then forms each node's Q times its parent's coefficient
// block (the root's is I) level by level, in place, and finally each leaf
// tile's Q times its coefficient block, in place in the panel.
// p advanced by e elements of its storage type (FP32 when f32)
static double* adv_e(double* p, size_t e, bool f32) { return f32 ? f64p(f32p(p) + e) : p + e; }
static const double* adv_e(const double* p, size_t e, bool f32) { return f32 ? f64p(f32p(p) + e) : p + e; }

static void tsqr(Ctx& c, const std::vector<TsqrPanel>& ps) {
    if (ps.empty()) return;
    struct Tile { int panel; int row0; int rows; int node; };
    std::vector<Tile> tiles;
    std::vector<std::vector<int>> cur(ps.size());
    int nnodes = 0;
    for (size_t p = 0; p < ps.size(); ++p) {
        KMH_CHECK(ps[p].nb <= QC && ps[p].m >= ps[p].nb, KMH_E_BAD_ARG, "tsqr: invalid panel shape");
        const int nt = std::max<int>(1, (int)ceil_div(ps[p].m, QTL));
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
    const size_t RS = (size_t)QC * QC;          // R and Z slots, ld QC
    const size_t QS = (size_t)QF * QC * QC;     // V slot of an inner node, (QF QC) x QC, ld QF QC
    double* RT = static_cast<double*>(c.scratch(std::max<size_t>(1, total_nodes) * RS * sizeof(double)));
    double* QTb = static_cast<double*>(c.scratch(std::max(1, ninternal) * QS * sizeof(double)));
    double* ZT = static_cast<double*>(c.scratch(std::max<size_t>(1, total_nodes) * RS * sizeof(double)));
    // leaf tile QRs (V in place)
    {
        std::vector<SqrDesc> ds;
        for (const auto& t : tiles) {
            const auto& P = ps[t.panel];
            const bool single = cur[t.panel].size() == 1;
            double* a = adv_e(P.A, t.row0, P.e32);
            const double* in = P.src ? adv_e(P.src, t.row0, P.e32) : a;
            SqrDesc sd{in, P.src ? P.lds : P.lda, 1, t.rows, 0, a, P.lda,
                       single ? P.R : RT + (size_t)t.node * RS, single ? P.ldr : QC,
                       ZT + (size_t)t.node * RS, P.nb, P.e32};
            if (P.kp > 0) {
                sd.pq = adv_e(P.pq, t.row0, P.e32);
                sd.ldp = P.ldp;
                sd.kp = P.kp;
                sd.pt = P.pt;
            }
            ds.push_back(sd);
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
                // children's R factors stacked (a lone child passes through)
                ds.push_back({RT + (size_t)L[i] * RS, QC, cnt, nb, cnt == 1 ? 0 : (long long)RS,
                              QTb + (size_t)qi * QS, QF * QC,
                              final_level ? ps[p].R : RT + (size_t)node * RS, final_level ? ps[p].ldr : QC,
                              ZT + (size_t)node * RS, nb, 0});
                nxt.push_back(node);
                if (final_level) root[p] = node;
            }
            L = nxt;
        }
        if (ds.empty()) break;
        launch_tile_qr(c, ds);
        ++level;
    }
    // down pass: each inner node's Q times its coefficient block (the
    // root's is I), in place in its V slot; child h's coefficient block is
    // rows h nb .. h nb + nb - 1 of that product
    std::vector<const double*> cptr(total_nodes, nullptr);
    std::vector<int> cld(total_nodes, 0);
    for (int lv = level - 1; lv >= 0; --lv) {
        std::vector<WyDesc> wd;
        for (const auto& nd : inner) {
            if (nd.level != lv) continue;
            const int nb = ps[nd.panel].nb;
            double* Q = QTb + (size_t)(nd.node - nnodes) * QS;
            wd.push_back({Q, QF * QC, nd.count * nb, ZT + (size_t)nd.node * RS, cptr[nd.node], cld[nd.node], nb,
                          Q, QF * QC, 0});
            for (int h = 0; h < nd.count; ++h) {
                cptr[nd.first + h] = Q + (size_t)h * nb;
                cld[nd.first + h] = QF * QC;
            }
        }
        wy_apply(c, wd);
    }
    // leaf tiles: A_tile <- Q_tile C_tile in place
    {
        std::vector<WyDesc> wd;
        for (const auto& t : tiles) {
            const auto& P = ps[t.panel];
            double* a = adv_e(P.A, t.row0, P.e32);
            wd.push_back({a, P.lda, t.rows, ZT + (size_t)t.node * RS, cptr[t.node], cld[t.node], P.nb, a, P.lda, P.e32});
        }
        wy_apply(c, wd);
    }
    c.scratch_free(ZT);
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
        std::vector<GemmProb> p1, u1, p2;
        std::vector<CgsProb> cg;
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
            const uint8_t e = p.e32;
            double* Aj = adv_e(p.A, (size_t)c0 * p.lda, e);
            double* Rj = p.R + (size_t)c0 * p.ldr;
            // this column block's input: the source panel on first touch
            const double* sj = p.src ? adv_e(p.src, (size_t)c0 * p.lds, e) : nullptr;
            // FP64 panels: the update X = S - Q T1 and the projection Q^T X in one pass
            const bool fused = !e && c0 <= CGS_KQ_MAX && nb <= CGS_NB_MAX;
            if (sj && c0 > 64 && !fused) {   // beyond the out-of-place update's k range: copy first
                Copy2D cp{sj, p.lds, Aj, p.lda, p.m, nb, 1.0};
                cp.es = cp.ed = e;
                pre.push_back(cp);
                sj = nullptr;
            }
            if (c0 > 0) {
                double* t2 = T2 + toff;
                toff += (size_t)c0 * nb;
                // projections Q^T X (FP64 out), updates X -= Q R (X in the panel's type)
                GemmProb g0{c0, nb, p.m, p.A, p.lda, sj ? sj : Aj, sj ? p.lds : p.lda, Rj, p.ldr, 1.0, 0.0, 0};
                g0.ea = g0.eb = e;
                p1.push_back(g0);
                if (fused) {
                    cg.push_back({p.A, p.lda, p.m, c0, sj ? sj : Aj, sj ? p.lds : p.lda, Rj, p.ldr, Aj, p.lda, nb, t2, c0});
                } else {
                    GemmProb g1{p.m, nb, c0, p.A, p.lda, Rj, p.ldr, Aj, p.lda, -1.0, 1.0, 0};
                    g1.Cin = sj;
                    g1.ea = g1.ec = e;
                    u1.push_back(g1);
                    GemmProb g2{c0, nb, p.m, p.A, p.lda, Aj, p.lda, t2, c0, 1.0, 0.0, 0};
                    g2.ea = g2.eb = e;
                    p2.push_back(g2);
                }
                acc.push_back({t2, c0, Rj, p.ldr, c0, nb, 1.0, 1.0});
                // the second pass's update X -= Q T2 happens as the leaf
                // tiles load X (k_tile_qr)
                TsqrPanel tpj{Aj, p.m, nb, p.lda, Rj + c0, p.ldr, nullptr, 0, e};
                tpj.pq = p.A;
                tpj.ldp = p.lda;
                tpj.kp = c0;
                tpj.pt = t2;
                tp.push_back(tpj);
            } else {
                tp.push_back({Aj, p.m, nb, p.lda, Rj + c0, p.ldr, sj, p.lds, e});
            }
        }
        copy2d(c, pre);
        if (!p1.empty()) {
            gemm(c, true, false, p1);
            cgs_pass(c, cg);
            gemm(c, false, false, u1);
            gemm(c, true, false, p2);
            axpby2d(c, acc);
        }
        tsqr(c, tp);
        if (T2) c.scratch_free(T2);
    }
}

}  // namespace kmh
