// Batched dense FP64 linear algebra for gfx950 (see la.h).
//
// Design notes
//  * GEMM: 64x64 output tiles, 256 threads = 4 waves, each wave a 32x32
//    sub-tile made of 2x2 V_MFMA_F64_16X16X4F64 blocks; operands staged in
//    LDS in 16-deep k-chunks.  Problems of a launch are uneven; a block finds
//    its problem by binary search over per-problem tile offsets.  Long-k,
//    few-tile problems are split along k into partial products that a second
//    launch reduces in a fixed order (deterministic, no atomics).
//  * POTRF / TRSM: right-looking blocked algorithms with 64-wide diagonal
//    blocks, trailing updates by GEMM.  A diagonal Cholesky block is one
//    wave with a row per lane in registers; diagonal solves split the 64
//    rows into four 16-row groups, one per wave.
//  * QR of tall panels: qr.hip.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "la.h"

namespace kmh {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// grouped GEMM
// ---------------------------------------------------------------------------
struct GemmDesc {
    const double* A;
    const double* B;
    double* C;
    int m, n, k, lda, ldb, ldc;
    double alpha, beta;
    int tile0, tn, lower, pad;
    const double* Cin;   // beta term source (k_gemm_nn_rows only; others read C)
};

constexpr int GT = 64;
constexpr int GK = 32;

// Lane layout of V_MFMA_F64_16X16X4F64 on gfx950 (measured with one-hot
// operands): A(i,k) is held by lane i + 16k, B(k,j) by lane j + 16k, and
// accumulator register r of lane j + 16g holds D(4r + g, j).  (The FP32
// 16x16x4 form instead holds D(4g + r, j).)
__device__ __forceinline__ int mfma_d_row(int g, int r) { return 4 * r + g; }

// Streamed operands of the skinny products are touched once per launch:
// non-temporal loads and stores keep them from displacing reused data.
__device__ __forceinline__ double ld_stream(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(double v, double* p) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ double ld_stream(const float* p) { return (double)__builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(double v, float* p) { __builtin_nontemporal_store((float)v, p); }

// Storage types of a GEMM launch: code TC = ea | eb << 1 | ec << 2, a set
// bit meaning FP32 in memory (see GemmProb).
template <int F> struct ElemOf { using T = double; };
template <> struct ElemOf<1> { using T = float; };
template <int TC> using EltA = typename ElemOf<TC & 1>::T;
template <int TC> using EltB = typename ElemOf<(TC >> 1) & 1>::T;
template <int TC> using EltC = typename ElemOf<(TC >> 2) & 1>::T;
template <int N> using icst = std::integral_constant<int, N>;

// call f(icst<TC>) for a supported storage code
template <class F> void with_types(int tc, F&& f) {
    switch (tc) {
        case 0: f(icst<0>{}); break;
        case 1: f(icst<1>{}); break;
        case 2: f(icst<2>{}); break;
        case 3: f(icst<3>{}); break;
        case 5: f(icst<5>{}); break;
        case 6: f(icst<6>{}); break;
        default: throw Error(KMH_E_BAD_ARG, "gemm: unsupported combination of FP32 operands");
    }
}
static int type_code(const GemmProb& p) { return (p.ea ? 1 : 0) | (p.eb ? 2 : 0) | (p.ec ? 4 : 0); }
// p advanced by e elements of its storage type
static const double* adv(const double* p, size_t e, bool f32) { return f32 ? f64p(f32p(p) + e) : p + e; }
static double* adv(double* p, size_t e, bool f32) { return f32 ? f64p(f32p(p) + e) : p + e; }

// Operand staging: each thread moves GT*GK/256 elements of A and of B per
// k-chunk.  Element e of thread tid maps to (row-or-col index, k index) so
// that consecutive threads read consecutive addresses of the column-major
// operand.
template <bool T>
__device__ __forceinline__ void gemm_stage_idx(int tid, int q, int& i, int& kk) {
    if (!T) { i = tid & 63; kk = (tid >> 6) + 4 * q; }          // i contiguous in memory
    else    { kk = tid & (GK - 1); i = tid / GK + (256 / GK) * q; }   // k contiguous in memory
}

template <bool TA, bool TB, int TC>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc* __restrict__ ds, const int* __restrict__ map) {
    constexpr int NQ = GT * GK / 256;
    __shared__ double As[GK][GT + 1];
    __shared__ double Bs[GK][GT + 1];
    const int t = blockIdx.x;
    const GemmDesc d = ds[map[t]];
    const EltA<TC>* __restrict__ Ap = reinterpret_cast<const EltA<TC>*>(d.A);
    const EltB<TC>* __restrict__ Bp = reinterpret_cast<const EltB<TC>*>(d.B);
    EltC<TC>* Cp = reinterpret_cast<EltC<TC>*>(d.C);
    const int lt = t - d.tile0;
    const int row0 = (lt / d.tn) * GT, col0 = (lt % d.tn) * GT;
    if (d.lower && col0 > row0 + GT - 1) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    // prefetch registers for one k-chunk of A and B
    double ra[NQ], rb[NQ];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            int i, kk;
            gemm_stage_idx<TA>(tid, q, i, kk);
            const int gi = row0 + i, gk = k0 + kk;
            double v = 0.0;
            if (gi < d.m && gk < d.k) v = TA ? (double)Ap[gk + (size_t)gi * d.lda] : (double)Ap[gi + (size_t)gk * d.lda];
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            int j, kk;
            gemm_stage_idx<!TB>(tid, q, j, kk);
            const int gj = col0 + j, gk = k0 + kk;
            double v = 0.0;
            if (gj < d.n && gk < d.k) v = TB ? (double)Bp[gj + (size_t)gk * d.ldb] : (double)Bp[gk + (size_t)gj * d.ldb];
            rb[q] = v;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < d.k; k0 += GK) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            int i, kk;
            gemm_stage_idx<TA>(tid, q, i, kk);
            As[kk][i] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            int j, kk;
            gemm_stage_idx<!TB>(tid, q, j, kk);
            Bs[kk][j] = rb[q];
        }
        __syncthreads();
        if (k0 + GK < d.k) fetch(k0 + GK);   // next chunk in flight during the MFMAs
#pragma unroll
        for (int ks = 0; ks < GK / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            const double a0 = As[kk][wr + (lane & 15)];
            const double a1 = As[kk][wr + 16 + (lane & 15)];
            const double b0 = Bs[kk][wc + (lane & 15)];
            const double b1 = Bs[kk][wc + 16 + (lane & 15)];
            // operands swapped: each accumulator holds a 16x16 block of C^T,
            // so a store instruction covers whole column segments of C
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue: all loads of C before any store, so they are in flight together
    double cv[2][2][4];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = row0 + wr + mb * 16 + (lane & 15);
                const int col = col0 + wc + nb * 16 + mfma_d_row(lane >> 4, j);
                cv[mb][nb][j] = (d.beta != 0.0 && row < d.m && col < d.n) ? (double)Cp[row + (size_t)col * d.ldc] : 0.0;
            }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = row0 + wr + mb * 16 + (lane & 15);
                const int col = col0 + wc + nb * 16 + mfma_d_row(lane >> 4, j);
                if (row < d.m && col < d.n) {
                    double v = d.alpha * acc[mb][nb][j];
                    if (d.beta != 0.0) v += d.beta * cv[mb][nb][j];
                    Cp[row + (size_t)col * d.ldc] = (EltC<TC>)v;
                }
            }
}

struct RedDesc {
    double* C;
    const double* W;
    int ldc, m, n, s;
    double alpha, beta;
    long long elem0;   // first workgroup of the problem in the reduction launch
};

// elem0 is the problem's first workgroup; each workgroup sums 256 elements
// (partials FP64; C of storage type EC)
template <class EC>
__global__ __launch_bounds__(256) void k_splitk_reduce(const RedDesc* __restrict__ ds, const int* __restrict__ map) {
    const RedDesc d = ds[map[blockIdx.x]];
    const long long loc = ((long long)blockIdx.x - d.elem0) * 256 + threadIdx.x;
    if (loc >= (long long)d.m * d.n) return;
    const int i = (int)(loc % d.m), j = (int)(loc / d.m);
    const size_t mn = (size_t)d.m * d.n;
    double s = 0.0;
    for (int q = 0; q < d.s; ++q) s += d.W[q * mn + loc];
    EC* cp = reinterpret_cast<EC*>(d.C) + i + (size_t)j * d.ldc;
    double v = d.alpha * s;
    if (d.beta != 0.0) v += d.beta * (double)(*cp);
    *cp = (EC)v;
}

// Tall TN products C = alpha A^T B + beta C with m, n <= 80 and long k (the
// projections Q^T X of the panel QR and of the Woodbury cascade): one
// workgroup holds the whole (16-padded) output for a slab of TT_KS rows of k,
// so A and B are read once; slabs of one problem are summed by the split-k
// reduction in slab order.  The 16x16 output blocks are dealt round-robin to
// the four waves.  Two staging shapes: KC = 16 rows of k by up to 80 columns
// per step, and for outputs of at most 32 x 32 KC = 64 rows by 32 columns,
// so narrow products keep four times the loads in flight.  Both issue the
// MFMAs in the same k order, so the shape class never changes a result.
constexpr int TT_MAX = 80;
constexpr int TT_KS = 256;
constexpr int TT_NARROW = 32;

template <int KC, int MW, int TC>
__global__ __launch_bounds__(256, 2) void k_gemm_tn_tall(const GemmDesc* __restrict__ ds) {
    constexpr int NG = 256 / KC;                      // column groups of the staging
    constexpr int CPT = (MW + NG - 1) / NG;           // columns per thread and operand
    constexpr int NBLK = ((MW + 15) / 16) * ((MW + 15) / 16);
    constexpr int T = (NBLK + 3) / 4;                 // output blocks per wave
    __shared__ double As[KC][MW + 1];
    __shared__ double Bs[KC][MW + 1];
    const GemmDesc d = ds[blockIdx.x];
    const EltA<TC>* __restrict__ Ap = reinterpret_cast<const EltA<TC>*>(d.A);
    const EltB<TC>* __restrict__ Bp = reinterpret_cast<const EltB<TC>*>(d.B);
    EltC<TC>* Cp = reinterpret_cast<EltC<TC>*>(d.C);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int NB = (d.n + 15) >> 4, nblk = ((d.m + 15) >> 4) * NB;
    d4 acc[T];
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    // staging: thread tid moves (k = tid % KC, columns tid / KC + NG q)
    const int skk = tid % KC, scol = tid / KC;
    double ra[CPT], rb[CPT];
    auto fetch = [&](int k0) {
        const int gk = k0 + skk;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int col = scol + NG * q;
            ra[q] = (col < d.m && gk < d.k) ? ld_stream(Ap + gk + (size_t)col * d.lda) : 0.0;
            rb[q] = (col < d.n && gk < d.k) ? ld_stream(Bp + gk + (size_t)col * d.ldb) : 0.0;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < d.k; k0 += KC) {
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int col = scol + NG * q;
            if (col < MW) {
                As[skk][col] = ra[q];
                Bs[skk][col] = rb[q];
            }
        }
        __syncthreads();
        if (k0 + KC < d.k) fetch(k0 + KC);
#pragma unroll
        for (int ks = 0; ks < KC / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int b = wave + 4 * t;
                if (b < nblk) {
                    const int mb = b / NB, nb = b - mb * NB;
                    // operands swapped: the accumulator holds a block of C^T
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Bs[kk][nb * 16 + (lane & 15)],
                                                                  As[kk][mb * 16 + (lane & 15)], acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    double cv[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int b = wave + 4 * t;
        const int mb = b / NB, nb = b - mb * NB;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = mb * 16 + (lane & 15), col = nb * 16 + mfma_d_row(lane >> 4, r);
            cv[t][r] = (b < nblk && d.beta != 0.0 && row < d.m && col < d.n) ? (double)Cp[row + (size_t)col * d.ldc] : 0.0;
        }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int b = wave + 4 * t;
        const int mb = b / NB, nb = b - mb * NB;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = mb * 16 + (lane & 15), col = nb * 16 + mfma_d_row(lane >> 4, r);
            if (b < nblk && row < d.m && col < d.n) {
                double v = d.alpha * acc[t][r];
                if (d.beta != 0.0) v += d.beta * cv[t][r];
                Cp[row + (size_t)col * d.ldc] = (EltC<TC>)v;
            }
        }
    }
}

// Wide NN products C = alpha A B + beta C with n <= 160 (panel updates
// X -= Q R, the Woodbury apply): a workgroup owns 64
// rows and every output column, so each row of A is read once; wave w takes
// rows 16w..16w+15.  A is loaded straight from memory (16 consecutive rows
// per k index), B through LDS in 16-deep chunks.  A wave reads all of its
// own rows of A before it writes the same rows of C, so C may alias A (an
// in-place product A <- A B with n = k).
constexpr int WN_MAX = 160;
constexpr int WN_KC = 16;

template <int TC>
__global__ __launch_bounds__(256, 2) void k_gemm_nn_wide(const GemmDesc* __restrict__ ds,
                                                        const int* __restrict__ map) {
    __shared__ double Bs[WN_KC][WN_MAX + 1];
    const int t = blockIdx.x;
    const GemmDesc d = ds[map[t]];
    const EltB<TC>* __restrict__ Bp = reinterpret_cast<const EltB<TC>*>(d.B);
    EltC<TC>* Cp = reinterpret_cast<EltC<TC>*>(d.C);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int NB = (d.n + 15) >> 4;
    const int arow = (t - d.tile0) * 64 + wave * 16 + (lane & 15);
    const bool arow_ok = arow < d.m;
    const EltA<TC>* ap = reinterpret_cast<const EltA<TC>*>(d.A) + (arow_ok ? arow : 0);
    d4 acc[10];
#pragma unroll
    for (int b = 0; b < 10; ++b) acc[b] = d4{0.0, 0.0, 0.0, 0.0};
    const int skk = tid & 15, scol = tid >> 4;
    double ra[4], rb[10];
    auto fetch = [&](int k0) {
        const int gk = k0 + skk;
#pragma unroll
        for (int q = 0; q < 10; ++q) {
            const int col = scol + 16 * q;
            rb[q] = (col < d.n && gk < d.k) ? (double)Bp[gk + (size_t)col * d.ldb] : 0.0;
        }
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int ka = k0 + ks * 4 + (lane >> 4);
            ra[ks] = (arow_ok && ka < d.k) ? ld_stream(ap + (size_t)ka * d.lda) : 0.0;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < d.k; k0 += WN_KC) {
#pragma unroll
        for (int q = 0; q < 10; ++q) Bs[skk][scol + 16 * q] = rb[q];
        double a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a[ks] = ra[ks];
        __syncthreads();
        if (k0 + WN_KC < d.k) fetch(k0 + WN_KC);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
#pragma unroll
            for (int b = 0; b < 10; ++b)
                if (b < NB)   // operands swapped: the accumulator holds a block of C^T
                    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(Bs[kk][b * 16 + (lane & 15)], a[ks], acc[b], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue in two halves: the loads of C of a half before its stores
#pragma unroll
    for (int h = 0; h < 10; h += 5) {
        double cv[5][4];
#pragma unroll
        for (int b = 0; b < 5; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = (h + b) * 16 + mfma_d_row(lane >> 4, r);
                cv[b][r] = (h + b < NB && d.beta != 0.0 && arow_ok && col < d.n) ? ld_stream(Cp + arow + (size_t)col * d.ldc)
                                                                                   : 0.0;
            }
#pragma unroll
        for (int b = 0; b < 5; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = (h + b) * 16 + mfma_d_row(lane >> 4, r);
                if (h + b < NB && arow_ok && col < d.n) {
                    double v = d.alpha * acc[h + b][r];
                    if (d.beta != 0.0) v += d.beta * cv[b][r];
                    st_stream(v, Cp + arow + (size_t)col * d.ldc);
                }
            }
    }
}

// Narrow NN products C = alpha A B + beta C with n <= 32 and k <= 2048 / NMAX
// (NMAX = 8, 16, 32: the panel QR updates X -= Q R, the one-column
// Woodbury updates of a solve): one thread per row
// of C, B in LDS (broadcast reads), the row of A streamed once with
// coalesced loads and n accumulators in registers.  ROWS threads (128 or
// 256) per workgroup, 128 when every problem of the launch has <= 128 rows.  The multiply-adds run
// in k order for each element.  A thread reads its whole row of A before it
// writes its row of C, so C may alias A.
// k limit of the row kernel per width class: B (k x NMAX) fills 16 KB of LDS
template <int NMAX> constexpr int nr_kmax() { return 2048 / NMAX; }

template <int NMAX, int ROWS, int TC>
__global__ __launch_bounds__(ROWS) void k_gemm_nn_rows(const GemmDesc* __restrict__ ds, const int* __restrict__ map) {
    constexpr int KMAX = nr_kmax<NMAX>();
    __shared__ double Bs[KMAX * NMAX];
    const GemmDesc d = ds[map[blockIdx.x]];
    const EltB<TC>* __restrict__ Bp = reinterpret_cast<const EltB<TC>*>(d.B);
    for (int e = threadIdx.x; e < KMAX * NMAX; e += ROWS) {
        const int kk = e / NMAX, j = e - kk * NMAX;
        Bs[e] = (kk < d.k && j < d.n) ? (double)Bp[kk + (size_t)j * d.ldb] : 0.0;
    }
    __syncthreads();
    const int row = (blockIdx.x - d.tile0) * ROWS + threadIdx.x;
    if (row >= d.m) return;
    double acc[NMAX];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) acc[j] = 0.0;
    const EltA<TC>* ap = reinterpret_cast<const EltA<TC>*>(d.A) + row;
    for (int k0 = 0; k0 < d.k; k0 += 8) {
        double a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = (k0 + q < d.k) ? ld_stream(ap + (size_t)(k0 + q) * d.lda) : 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double* b = Bs + (k0 + q) * NMAX;
#pragma unroll
            for (int j = 0; j < NMAX; ++j) acc[j] = __fma_rn(a[q], b[j], acc[j]);
        }
    }
    EltC<TC>* cp = reinterpret_cast<EltC<TC>*>(d.C) + row;
    if (d.beta != 0.0) {
        const EltC<TC>* ci = reinterpret_cast<const EltC<TC>*>(d.Cin ? d.Cin : d.C) + row;
        double cv[NMAX];
#pragma unroll
        for (int j = 0; j < NMAX; ++j) cv[j] = j < d.n ? ld_stream(ci + (size_t)j * d.ldc) : 0.0;
#pragma unroll
        for (int j = 0; j < NMAX; ++j)
            if (j < d.n) st_stream(d.alpha * acc[j] + d.beta * cv[j], cp + (size_t)j * d.ldc);
    } else {
#pragma unroll
        for (int j = 0; j < NMAX; ++j)
            if (j < d.n) st_stream(d.alpha * acc[j], cp + (size_t)j * d.ldc);
    }
}

// Which kernel takes a problem, and how it is split, depends on the problem's
// own shape only, so a problem gives bitwise the same result whatever else
// shares the launch.
static bool tall_tn(bool ta, bool tb, const GemmProb& p) {
    return ta && !tb && !p.lower && p.m <= TT_MAX && p.n <= TT_MAX;
}
static bool rows_nn(bool ta, bool tb, const GemmProb& p) {
    if (ta || tb || p.lower || p.n > 32) return false;
    return p.k <= (p.n <= 8 ? nr_kmax<8>() : p.n <= 16 ? nr_kmax<16>() : nr_kmax<32>());
}
static bool wide_nn(bool ta, bool tb, const GemmProb& p) {
    return !ta && !tb && !p.lower && p.n <= WN_MAX && (p.k <= 512 || p.m >= 64 * 64);
}

// Algorithmic flops (2mnk) and HBM bytes (A and B read once, C written, and
// read too when beta != 0) of a launch's descriptor list, for the timed spans.
static double desc_flops(const std::vector<GemmDesc>& ds) {
    double f = 0.0;
    for (const auto& d : ds) f += 2.0 * d.m * d.n * (double)d.k;
    return f;
}
static double desc_bytes(const std::vector<GemmDesc>& ds, int tc = 0) {
    const double sa = (tc & 1) ? 4.0 : 8.0, sb = (tc & 2) ? 4.0 : 8.0, sc = (tc & 4) ? 4.0 : 8.0;
    double b = 0.0;
    for (const auto& d : ds)
        b += sa * d.k * d.m + sb * (double)d.k * d.n + sc * (double)d.m * d.n * (d.beta != 0.0 ? 2 : 1);
    return b;
}

static void gemm_generic(Ctx& c, bool ta, bool tb, const std::vector<GemmProb>& probs, int tc) {
    std::vector<int> split(probs.size(), 1), kchunk(probs.size(), 0);
    size_t wsz = 0;
    for (size_t i = 0; i < probs.size(); ++i) {
        const auto& p = probs[i];
        if (p.lower) continue;
        const long long own = ceil_div(p.m, GT) * ceil_div(p.n, GT);
        if (p.k >= 1024 && own < 64) {
            long long s = std::min<long long>(ceil_div(p.k, 512), std::max<long long>(1, 256 / own));
            if (s > 1) {
                int kc = (int)(ceil_div(ceil_div(p.k, s), GK) * GK);
                s = ceil_div(p.k, kc);
                split[i] = (int)s;
                kchunk[i] = kc;
                wsz += (size_t)s * p.m * p.n;
            }
        }
    }
    // an FP32 C cannot share a launch with FP64 split partials: such a call
    // is split in two (the problem lists are independent)
    bool split_any = false, whole_any = false;
    for (size_t i = 0; i < probs.size(); ++i) (split[i] > 1 ? split_any : whole_any) = true;
    if ((tc & 4) && split_any && whole_any) {
        std::vector<GemmProb> a, b;
        for (size_t i = 0; i < probs.size(); ++i) (split[i] > 1 ? a : b).push_back(probs[i]);
        gemm_generic(c, ta, tb, a, tc);
        gemm_generic(c, ta, tb, b, tc);
        return;
    }
    double* W = wsz ? static_cast<double*>(c.scratch(wsz * sizeof(double))) : nullptr;
    std::vector<GemmDesc> ds;
    std::vector<int> map;
    int tiles = 0;
    std::vector<RedDesc> rs;
    std::vector<int> rmap;
    size_t woff = 0;
    long long rblk = 0;
    for (size_t i = 0; i < probs.size(); ++i) {
        const auto& p = probs[i];
        const int tn = (int)ceil_div(p.n, GT);
        const int nt = (int)ceil_div(p.m, GT) * tn;
        if (split[i] == 1) {
            map_append(map, (int)ds.size(), nt);
            ds.push_back({p.A, p.B, p.C, p.m, p.n, p.k, p.lda, p.ldb, p.ldc, p.alpha, p.beta,
                          tiles, tn, p.lower, 0});
            tiles += nt;
        } else {
            const int s = split[i];
            const int kc = kchunk[i];
            for (int q = 0; q < s; ++q) {
                const int k0 = q * kc;
                const int kk = std::min(kc, p.k - k0);
                const double* A = adv(p.A, ta ? (size_t)k0 : (size_t)k0 * p.lda, tc & 1);
                const double* B = adv(p.B, tb ? (size_t)k0 * p.ldb : (size_t)k0, tc & 2);
                map_append(map, (int)ds.size(), nt);
                ds.push_back({A, B, W + woff + (size_t)q * p.m * p.n, p.m, p.n, kk, p.lda, p.ldb,
                              p.m, 1.0, 0.0, tiles, tn, 0, 0});
                tiles += nt;
            }
            const long long nb = ceil_div((long long)p.m * p.n, 256);
            map_append(rmap, (int)rs.size(), nb);
            rs.push_back({p.C, W + woff, p.ldc, p.m, p.n, s, p.alpha, p.beta, rblk});
            woff += (size_t)s * p.m * p.n;
            rblk += nb;
        }
    }
    if (!ds.empty()) {
        GemmDesc* dd = c.upload_vec(ds);
        const int* dm = c.upload_vec(map);
        const dim3 g(tiles);
        const double fl = c.prof ? desc_flops(ds) : 0.0;
        const int T = KMH_PROF_GEMM;
        // split partials are FP64: their launches carry an FP64 C
        const int tcw = split_any ? (tc & 3) : tc;
        with_types(tcw, [&](auto k) {
            constexpr int TC = decltype(k)::value;
            if (!ta && !tb) KMH_LAUNCH_TIMED(c, T, fl, fl, (k_gemm<false, false, TC>), g, dim3(256), dd, dm);
            else if (!ta && tb) KMH_LAUNCH_TIMED(c, T, fl, fl, (k_gemm<false, true, TC>), g, dim3(256), dd, dm);
            else if (ta && !tb) KMH_LAUNCH_TIMED(c, T, fl, fl, (k_gemm<true, false, TC>), g, dim3(256), dd, dm);
            else KMH_LAUNCH_TIMED(c, T, fl, fl, (k_gemm<true, true, TC>), g, dim3(256), dd, dm);
        });
    }
    if (!rs.empty()) {
        RedDesc* dr = c.upload_vec(rs);
        const int* drm = c.upload_vec(rmap);
        if (tc & 4)
            KMH_LAUNCH_TIMED(c, KMH_PROF_GEMM, 0.0, 0.0, k_splitk_reduce<float>, dim3((unsigned)rblk), dim3(256), dr, drm);
        else
            KMH_LAUNCH_TIMED(c, KMH_PROF_GEMM, 0.0, 0.0, k_splitk_reduce<double>, dim3((unsigned)rblk), dim3(256), dr,
                             drm);
    }
    if (W) c.scratch_free(W);
}

static void gemm_tall_tn(Ctx& c, const std::vector<GemmProb>& probs, int tc) {
    size_t wsz = 0;
    for (const auto& p : probs) {
        const int s = (int)ceil_div(p.k, TT_KS);
        if (s > 1) wsz += (size_t)s * p.m * p.n;
    }
    double* W = wsz ? static_cast<double*>(c.scratch(wsz * sizeof(double))) : nullptr;
    std::vector<GemmDesc> ds;
    std::vector<RedDesc> rs;
    std::vector<int> rmap;
    size_t woff = 0;
    long long rblk = 0;
    for (const auto& p : probs) {
        const int s = (int)ceil_div(p.k, TT_KS);
        if (s <= 1) {
            ds.push_back({p.A, p.B, p.C, p.m, p.n, p.k, p.lda, p.ldb, p.ldc, p.alpha, p.beta, 0, 0, 0, 0});
            continue;
        }
        for (int q = 0; q < s; ++q) {
            const int k0 = q * TT_KS;
            ds.push_back({adv(p.A, k0, tc & 1), adv(p.B, k0, tc & 2), W + woff + (size_t)q * p.m * p.n, p.m, p.n,
                          std::min(TT_KS, p.k - k0), p.lda, p.ldb, p.m, 1.0, 0.0, 0, 0, 0, 0});
        }
        const long long nb = ceil_div((long long)p.m * p.n, 256);
        map_append(rmap, (int)rs.size(), nb);
        rs.push_back({p.C, W + woff, p.ldc, p.m, p.n, s, p.alpha, p.beta, rblk});
        woff += (size_t)s * p.m * p.n;
        rblk += nb;
    }
    // one launch per staging shape (by the problem's own m and n) and per C
    // type (split slabs write FP64 partials)
    std::vector<GemmDesc> dn[2], dw[2];
    for (const auto& d : ds) {
        const int part = (d.C >= W && d.C < W + wsz) ? 1 : 0;
        (d.m <= TT_NARROW && d.n <= TT_NARROW ? dn[part] : dw[part]).push_back(d);
    }
    for (int part = 0; part < 2; ++part) {
        const int tcl = part ? (tc & 3) : tc;
        with_types(tcl, [&](auto k) {
            constexpr int TC = decltype(k)::value;
            if (!dn[part].empty()) {
                GemmDesc* dd = c.upload_vec(dn[part]);
                KMH_LAUNCH_TIMED(c, KMH_PROF_TN_TALL, c.prof ? desc_bytes(dn[part], TC) : 0.0,
                                 c.prof ? desc_flops(dn[part]) : 0.0, (k_gemm_tn_tall<64, TT_NARROW, TC>),
                                 dim3((unsigned)dn[part].size()), dim3(256), dd);
            }
            if (!dw[part].empty()) {
                GemmDesc* dd = c.upload_vec(dw[part]);
                KMH_LAUNCH_TIMED(c, KMH_PROF_TN_TALL, c.prof ? desc_bytes(dw[part], TC) : 0.0,
                                 c.prof ? desc_flops(dw[part]) : 0.0, (k_gemm_tn_tall<16, TT_MAX, TC>),
                                 dim3((unsigned)dw[part].size()), dim3(256), dd);
            }
        });
    }
    if (!rs.empty()) {
        RedDesc* dr = c.upload_vec(rs);
        const int* drm = c.upload_vec(rmap);
        if (tc & 4)
            KMH_LAUNCH_TIMED(c, KMH_PROF_GEMM, 0.0, 0.0, k_splitk_reduce<float>, dim3((unsigned)rblk), dim3(256), dr, drm);
        else
            KMH_LAUNCH_TIMED(c, KMH_PROF_GEMM, 0.0, 0.0, k_splitk_reduce<double>, dim3((unsigned)rblk), dim3(256), dr,
                             drm);
    }
    if (W) c.scratch_free(W);
}

// ---------------------------------------------------------------------------
// Fused Gram-Schmidt pass (la.h): one workgroup per slab of TT_KS rows, in
// chunks of CG_KC rows staged in LDS.  Per chunk, each wave forms one 16 x 16
// tile of X = S - Q T1 on MFMA (T1 held in LDS for the whole slab), writes it
// out and back to LDS, and the waves accumulate the slab's Q^T X on MFMA; the
// slabs' partial products are summed by the split-k reduction.  Q and S are
// read once and X written once, where the unfused pass reads Q and X again
// for the projection.
constexpr int CG_KC = 32;

struct CgsDesc {
    const double* Q; const double* S; const double* T1;
    double* X; double* W;
    int ldq, lds, ldt1, ldx;
    int rows, kq, nb, pad;
};

template <int KQ>
__global__ __launch_bounds__(256, 2) void k_cgs_pass(const CgsDesc* __restrict__ ds) {
    constexpr int NB = CGS_NB_MAX;
    constexpr int NQ = KQ / 16;                          // 16-row blocks of Q^T X
    constexpr int NT = (NQ * (NB / 16) + 3) / 4;         // its tiles per wave
    constexpr int QPT = (CG_KC * KQ) / 256;              // staged Q values per thread
    constexpr int SPT = (CG_KC * NB) / 256;              // staged S values per thread
    __shared__ double Qs[CG_KC][KQ + 1];
    __shared__ double Xs[CG_KC][NB + 1];
    __shared__ double Ts[KQ][NB + 1];
    const CgsDesc d = ds[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, lg = lane >> 4;
    const Glob<const double> Qg = glob(d.Q), Sg = glob(d.S);
    const Glob<double> Xg = glob(d.X);
    for (int e = tid; e < KQ * NB; e += 256) {
        const int i = e % KQ, j = e / KQ;
        Ts[i][j] = (i < d.kq && j < d.nb) ? d.T1[i + (size_t)j * d.ldt1] : 0.0;
    }
    // staging: thread tid moves row tid % CG_KC of columns tid / CG_KC + 8q
    const int sk = tid % CG_KC, sc = tid / CG_KC;
    double rq[QPT], rs[SPT];
    auto fetch = [&](int r0) {
        const int row = r0 + sk;
        const bool rok = row < d.rows;
#pragma unroll
        for (int q = 0; q < QPT; ++q) {
            const int col = sc + 8 * q;
            const bool ok = rok && col < d.kq;
            rq[q] = keep_or_zero(Qg[ok ? row + (size_t)col * d.ldq : 0], ok);
        }
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int col = sc + 8 * q;
            const bool ok = rok && col < d.nb;
            rs[q] = keep_or_zero(Sg[ok ? row + (size_t)col * d.lds : 0], ok);
        }
    };
    d4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    const int xrb = wave & 1, xcb = wave >> 1;   // this wave's tile of X in a chunk
    for (int r0 = 0; r0 < d.rows; r0 += CG_KC) {
#pragma unroll
        for (int q = 0; q < QPT; ++q) Qs[sk][sc + 8 * q] = rq[q];
#pragma unroll
        for (int q = 0; q < SPT; ++q) Xs[sk][sc + 8 * q] = rs[q];
        __syncthreads();
        if (r0 + CG_KC < d.rows) fetch(r0 + CG_KC);
        // X = S - Q T1, tile (xrb, xcb)
        d4 xa = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < KQ / 4; ++s4) {
            const int k = 4 * s4 + lg;
            xa = __builtin_amdgcn_mfma_f64_16x16x4f64(Qs[16 * xrb + li][k], Ts[k][16 * xcb + li], xa, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * xrb + mfma_d_row(lg, r), col = 16 * xcb + li;
            const double x = Xs[row][col] - xa[r];
            Xs[row][col] = x;
            if (r0 + row < d.rows && col < d.nb) Xg[r0 + row + (size_t)col * d.ldx] = x;
        }
        __syncthreads();
        // Q^T X of the chunk: block (mb, nb) of the output, A = Q^T from Qs
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int b = wave + 4 * t;
            if (b < NQ * (NB / 16)) {
                const int mb = b / (NB / 16), nb = b - mb * (NB / 16);
#pragma unroll
                for (int s4 = 0; s4 < CG_KC / 4; ++s4) {
                    const int kk = 4 * s4 + lg;
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Qs[kk][16 * mb + li], Xs[kk][16 * nb + li], acc[t], 0, 0,
                                                                  0);
                }
            }
        }
        __syncthreads();
    }
    // the slab's partial Q^T X (kq x nb, ld kq)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int b = wave + 4 * t;
        if (b < NQ * (NB / 16)) {
            const int mb = b / (NB / 16), nb = b - mb * (NB / 16);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * mb + mfma_d_row(lg, r), j = 16 * nb + li;
                if (i < d.kq && j < d.nb) d.W[i + (size_t)j * d.kq] = acc[t][r];
            }
        }
    }
}

void cgs_pass(Ctx& c, const std::vector<CgsProb>& ps) {
    if (ps.empty()) return;
    size_t wsz = 0;
    for (const auto& p : ps) {
        KMH_CHECK(p.kq > 0 && p.kq <= CGS_KQ_MAX && p.nb > 0 && p.nb <= CGS_NB_MAX, KMH_E_BAD_ARG,
                  "cgs_pass: block beyond its limits");
        wsz += (size_t)ceil_div(p.m, TT_KS) * p.kq * p.nb;
    }
    double* W = static_cast<double*>(c.scratch(wsz * sizeof(double)));
    std::vector<CgsDesc> d32, d48;
    std::vector<RedDesc> rs;
    std::vector<int> rmap;
    size_t woff = 0;
    long long rblk = 0;
    double by32 = 0.0, by48 = 0.0;   // algorithmic bytes: Q and S read, X written
    for (const auto& p : ps) {
        const int s = (int)ceil_div(p.m, TT_KS);
        for (int q = 0; q < s; ++q) {
            const int r0 = q * TT_KS;
            CgsDesc dd{p.Q + r0, p.S + r0, p.T1, p.X + r0, W + woff + (size_t)q * p.kq * p.nb,
                       p.ldq, p.lds, p.ldt1, p.ldx, std::min(TT_KS, p.m - r0), p.kq, p.nb, 0};
            (p.kq <= 32 ? d32 : d48).push_back(dd);
        }
        const long long nb = ceil_div((long long)p.kq * p.nb, 256);
        map_append(rmap, (int)rs.size(), nb);
        rs.push_back({p.T2, W + woff, p.ldt2, p.kq, p.nb, s, 1.0, 0.0, rblk});
        woff += (size_t)s * p.kq * p.nb;
        rblk += nb;
        (p.kq <= 32 ? by32 : by48) += 8.0 * p.m * (p.kq + 2.0 * p.nb);
    }
    if (!d32.empty()) {
        CgsDesc* dd = c.upload_vec(d32);
        KMH_LAUNCH_TIMED(c, KMH_PROF_TN_TALL, by32, 0.0, k_cgs_pass<32>, dim3((unsigned)d32.size()), dim3(256), dd);
    }
    if (!d48.empty()) {
        CgsDesc* dd = c.upload_vec(d48);
        KMH_LAUNCH_TIMED(c, KMH_PROF_TN_TALL, by48, 0.0, k_cgs_pass<48>, dim3((unsigned)d48.size()), dim3(256), dd);
    }
    RedDesc* dr = c.upload_vec(rs);
    const int* drm = c.upload_vec(rmap);
    KMH_LAUNCH_TIMED(c, KMH_PROF_GEMM, 0.0, 0.0, k_splitk_reduce<double>, dim3((unsigned)rblk), dim3(256), dr, drm);
    c.scratch_free(W);
}

static void gemm_wide_nn(Ctx& c, const std::vector<GemmProb>& probs, int tc) {
    std::vector<GemmDesc> ds;
    std::vector<int> map;
    int tiles = 0;
    for (const auto& p : probs) {
        const int nt = (int)ceil_div(p.m, 64);
        map_append(map, (int)ds.size(), nt);
        ds.push_back({p.A, p.B, p.C, p.m, p.n, p.k, p.lda, p.ldb, p.ldc, p.alpha, p.beta, tiles, 0, 0, 0});
        tiles += nt;
    }
    GemmDesc* dd = c.upload_vec(ds);
    const int* dm = c.upload_vec(map);
    with_types(tc, [&](auto k) {
        constexpr int TC = decltype(k)::value;
        KMH_LAUNCH_TIMED(c, KMH_PROF_NN_WIDE, c.prof ? desc_bytes(ds, TC) : 0.0, c.prof ? desc_flops(ds) : 0.0,
                         k_gemm_nn_wide<TC>, dim3(tiles), dim3(256), dd, dm);
    });
}

static void gemm_rows_nn(Ctx& c, const std::vector<GemmProb>& probs, int tc) {
    // one launch per width class (8, 16, 32); 128-row workgroups when every
    // problem of the class has at most 128 rows (TSQR tiles), else 256
    std::vector<const GemmProb*> cls[3];
    for (const auto& p : probs) cls[p.n <= 8 ? 0 : p.n <= 16 ? 1 : 2].push_back(&p);
    for (int cl = 0; cl < 3; ++cl) {
        if (cls[cl].empty()) continue;
        int mmax = 0;
        for (const GemmProb* p : cls[cl]) mmax = std::max(mmax, p->m);
        const int rows = mmax <= 128 ? 128 : 256;
        std::vector<GemmDesc> ds;
        std::vector<int> map;
        int blocks = 0;
        for (const GemmProb* pp : cls[cl]) {
            const GemmProb& p = *pp;
            const int nb = (int)ceil_div(p.m, rows);
            map_append(map, (int)ds.size(), nb);
            ds.push_back({p.A, p.B, p.C, p.m, p.n, p.k, p.lda, p.ldb, p.ldc, p.alpha, p.beta, blocks, 0, 0, 0, p.Cin});
            blocks += nb;
        }
        GemmDesc* dd = c.upload_vec(ds);
        const int* dm = c.upload_vec(map);
        const double by = c.prof ? desc_bytes(ds, tc) : 0.0, fl = c.prof ? desc_flops(ds) : 0.0;
        const int T = KMH_PROF_NN_ROWS;
        const dim3 g(blocks), b(rows);
        with_types(tc, [&](auto k) {
            constexpr int TC = decltype(k)::value;
            if (rows == 128) {
                if (cl == 0) KMH_LAUNCH_TIMED(c, T, by, fl, (k_gemm_nn_rows<8, 128, TC>), g, b, dd, dm);
                else if (cl == 1) KMH_LAUNCH_TIMED(c, T, by, fl, (k_gemm_nn_rows<16, 128, TC>), g, b, dd, dm);
                else KMH_LAUNCH_TIMED(c, T, by, fl, (k_gemm_nn_rows<32, 128, TC>), g, b, dd, dm);
            } else {
                if (cl == 0) KMH_LAUNCH_TIMED(c, T, by, fl, (k_gemm_nn_rows<8, 256, TC>), g, b, dd, dm);
                else if (cl == 1) KMH_LAUNCH_TIMED(c, T, by, fl, (k_gemm_nn_rows<16, 256, TC>), g, b, dd, dm);
                else KMH_LAUNCH_TIMED(c, T, by, fl, (k_gemm_nn_rows<32, 256, TC>), g, b, dd, dm);
            }
        });
    }
}

// Kernel classes of a GEMM problem.  The built-in rules pick by shape; a
// tuning table (kmh_gemm_autotune) may name another eligible class per shape
// class: the transposes and the ceil(log2) of m, n and k.  Either way the
// class depends on the problem's own shape only.
enum GemmClass { GC_RULE = 0, GC_GENERIC = 1, GC_TALL = 2, GC_ROWS = 3, GC_WIDE = 4 };

static int ceil_log2(int64_t v) {
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

uint32_t gemm_class_key(bool ta, bool tb, int m, int n, int k) {
    return (uint32_t(ta) << 31) | (uint32_t(tb) << 30) | (uint32_t(ceil_log2(m)) << 20) |
           (uint32_t(ceil_log2(n)) << 10) | uint32_t(ceil_log2(k));
}

bool gemm_class_eligible(int cls, bool ta, bool tb, const GemmProb& p) {
    const bool inplace = !p.Cin || p.Cin == p.C;
    switch (cls) {
        case GC_GENERIC: return inplace;
        case GC_TALL: return inplace && tall_tn(ta, tb, p);
        case GC_ROWS: return rows_nn(ta, tb, p);
        case GC_WIDE: return inplace && !ta && !tb && !p.lower && p.n <= WN_MAX;
        default: return false;
    }
}

static int gemm_route(const Ctx& c, bool ta, bool tb, const GemmProb& p) {
    int cls = c.gemm_force;
    if (cls == GC_RULE && !c.gemm_tune.empty()) {
        auto it = c.gemm_tune.find(gemm_class_key(ta, tb, p.m, p.n, p.k));
        if (it != c.gemm_tune.end()) cls = it->second;
    }
    if (cls != GC_RULE && gemm_class_eligible(cls, ta, tb, p)) return cls;
    if (tall_tn(ta, tb, p)) return GC_TALL;
    if (rows_nn(ta, tb, p)) return GC_ROWS;
    if (wide_nn(ta, tb, p)) return GC_WIDE;
    return GC_GENERIC;
}

void gemm(Ctx& c, bool ta, bool tb, const std::vector<GemmProb>& probs) {
    // one pass per storage-type code present (problem lists are independent)
    int seen = 0;
    for (const auto& p : probs) seen |= 1 << type_code(p);
    for (int tc = 0; tc < 8; ++tc) {
        if (!((seen >> tc) & 1)) continue;
        std::vector<GemmProb> gen, tall, wide, rows;
        for (const auto& p : probs) {
            if (p.m <= 0 || p.n <= 0 || type_code(p) != tc) continue;
            KMH_CHECK(!p.Cin || p.Cin == p.C || rows_nn(ta, tb, p), KMH_E_BAD_ARG,
                      "gemm: an out-of-place beta term needs n <= 32, k <= 2048 / (8, 16 or 32), no transposes");
            switch (gemm_route(c, ta, tb, p)) {
                case GC_TALL: tall.push_back(p); break;
                case GC_ROWS: rows.push_back(p); break;
                case GC_WIDE: wide.push_back(p); break;
                default: gen.push_back(p); break;
            }
        }
        if (!gen.empty()) gemm_generic(c, ta, tb, gen, tc);
        if (!tall.empty()) gemm_tall_tn(c, tall, tc);
        if (!wide.empty()) gemm_wide_nn(c, wide, tc);
        if (!rows.empty()) gemm_rows_nn(c, rows, tc);
    }
}

// ---------------------------------------------------------------------------
// POTRF (lower) and TRSM
// ---------------------------------------------------------------------------
constexpr int PB = 64;

__device__ __forceinline__ double lane_bcast(double v, int lane) {
    // value of v held by `lane` (compile-time constant after unrolling)
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Cholesky of one <= 64 x 64 diagonal block per wave: lane r holds row r of
// the lower triangle in registers; column j is scaled by 1/l_jj (LAPACK
// dpotf2 order) and the trailing rows updated with l_kj broadcast from lane k.
// No LDS and no barriers.
__global__ __launch_bounds__(64) void k_potrf_diag(const SqMat* __restrict__ ms, int* info0) {
    const SqMat m = ms[blockIdx.x];
    int* info = m.info ? m.info : info0;
    const int n = m.n, r = threadIdx.x;
    double a[PB];
#pragma unroll
    for (int k = 0; k < PB; ++k) a[k] = (r < n && k < n && k <= r) ? m.A[r + (size_t)k * m.ld] : 0.0;
    bool bad = false;
#pragma unroll
    for (int j = 0; j < PB; ++j) {
        if (j < n) {
            const double ajj = lane_bcast(a[j], j);
            bad |= !(ajj > 0.0);
            const double ljj = sqrt(ajj);
            const double inv = 1.0 / ljj;
            a[j] = (r == j) ? ljj : ((r > j) ? a[j] * inv : 0.0);
#pragma unroll
            for (int k = j + 1; k < PB; ++k) {
                if (k < n) {
                    const double lkj = lane_bcast(a[j], k);
                    a[k] -= a[j] * lkj;
                }
            }
        }
    }
    if (bad && r == 0) atomicExch(info, 1);
    if (r < n) {
#pragma unroll
        for (int k = 0; k < PB; ++k)
            if (k < n) m.A[r + (size_t)k * m.ld] = (k <= r) ? a[k] : 0.0;
    }
}

// Diagonal-block triangular solves, one 64 x 64 right-hand tile per
// workgroup.  Wave g owns rows 16g..16g+15 of the tile (16 values per lane,
// lane = right-hand column): the owning wave solves its 16 x 16 triangle,
// publishes the result in LDS, and the waves after it (before it for the
// transposed solve) subtract its contribution; one barrier per 16-row phase.
//   MODE 0: L X = B        (B nb x ncols, tiles of 64 columns)
//   MODE 1: L^T X = B      (as 0)
//   MODE 2: X L^T = B      (B ncols x nb: tiles of 64 rows; L X^T = B^T)
struct LdiagDesc {
    const double* L;
    int nb, ldl;
    double* B;
    int ncols, ldb;
    int chunk0, pad;
    double* Z = nullptr;   // MODE 2: the nb x ncols block zeroed beside the solve (ld ldb)
};

// lower triangle of a 64 x 64 block packed by columns: column l holds rows l..63
__device__ __forceinline__ int tri_col(int l) { return (l * (2 * PB + 1 - l)) >> 1; }

template <int MODE, class EB>
__global__ __launch_bounds__(256) void k_trsm_diag(const LdiagDesc* __restrict__ ds, const int* __restrict__ map) {
    __shared__ double Lt[PB * (PB + 1) / 2];   // L's lower triangle, identity beyond nb
    __shared__ double Xp[16][PB + 1];          // the solved rows of the current group, published to the others
    const LdiagDesc d = ds[map[blockIdx.x]];
    const Glob<EB> Bp = glob(reinterpret_cast<EB*>(d.B));
    const Glob<const double> Lg = glob(d.L);
    const int tid = threadIdx.x, c = tid & 63;
    const int g = __builtin_amdgcn_readfirstlane(tid >> 6);   // row group of this wave
    const int c0 = ((int)blockIdx.x - d.chunk0) * PB;
    const int nc = min(PB, d.ncols - c0);
    const int nb = __builtin_amdgcn_readfirstlane(d.nb);
    {
        // the triangle's loads all in flight (clamped addresses, masked
        // values), then the LDS stores
        constexpr int NE = PB * PB / 256;
        double lv[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = tid + 256 * q, i = e & 63, l = e >> 6;
            const bool in = i >= l && i < nb && l < nb;
            lv[q] = keep_or_zero(Lg[in ? i + (size_t)l * d.ldl : 0], in);
        }
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = tid + 256 * q, i = e & 63, l = e >> 6;
            if (i >= l) Lt[tri_col(l) + i - l] = (i < nb && l < nb) ? lv[q] : (i == l ? 1.0 : 0.0);
        }
    }
    if (MODE == 2 && d.Z) {
        // the upper triangle's block mirroring this chunk (potrf zeroes
        // the strict upper triangle this way)
        const Glob<double> Zg = glob(d.Z);
#pragma unroll
        for (int q = 0; q < PB * PB / 256; ++q) {
            const int e = tid + 256 * q, i = e & 63, cc = e >> 6;
            if (i < nb && cc < nc) Zg[i + (size_t)(c0 + cc) * d.ldb] = 0.0;
        }
    }
    // L(i, l) for i >= l (all lanes read the same element)
    auto Lv = [&](int i, int l) { return Lt[tri_col(l) + i - l]; };
    double x[16];
    // MODE 2: right-hand c is row c0 + c of B (ncols x nb); else column c0 + c
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = 16 * g + i;
        const bool ok = c < nc && row < nb;
        // clamped, always valid element, zeroed where out of range
        const int cc = c < nc ? c : 0, rr = row < nb ? row : 0;
        const double v = MODE == 2 ? (double)Bp[(size_t)(c0 + cc) + (size_t)rr * d.ldb]
                                   : (double)Bp[rr + (size_t)(c0 + cc) * d.ldb];
        x[i] = keep_or_zero(v, ok);
    }
    __syncthreads();
    const int ng = (nb + 15) >> 4;   // row groups holding rows of L
    const int rg = min(16, nb - 16 * g);   // rows of this wave's group (<= 0: none)
#pragma unroll 1
    for (int q = 0; q < ng; ++q) {
        const int p = MODE == 1 ? ng - 1 - q : q;   // group solved in this round
        const int rp = min(16, nb - 16 * p);
        if (g == p) {
            const int b = 16 * p;   // diagonal block of group p
            if (MODE != 1) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    if (i < rp) {
                        double s = x[i];
#pragma unroll
                        for (int l = 0; l < i; ++l) s -= Lv(b + i, b + l) * x[l];
                        x[i] = s / Lv(b + i, b + i);
                    }
                }
            } else {
#pragma unroll
                for (int i = 15; i >= 0; --i) {
                    if (i < rp) {
                        double s = x[i];
#pragma unroll
                        for (int l = i + 1; l < 16; ++l)
                            if (l < rp) s -= Lv(b + l, b + i) * x[l];
                        x[i] = s / Lv(b + i, b + i);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) Xp[i][c] = x[i];
        }
        __syncthreads();
        if ((MODE != 1 ? g > p : g < p) && rg > 0) {
            // rows of group g against the solved group p: L(16g + i, 16p + l)
            // (MODE 1: L^T, i.e. L(16p + l, 16g + i))
#pragma unroll
            for (int l = 0; l < 16; ++l) {
                if (l >= rp) break;
                const double xl = Xp[l][c];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const double lv = MODE != 1 ? Lv(16 * g + i, 16 * p + l) : Lv(16 * p + l, 16 * g + i);
                    if (i < rg) x[i] -= lv * xl;
                }
            }
        }
        __syncthreads();   // Xp is rewritten in the next round
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = 16 * g + i;
        if (c < nc && row < nb) {
            if (MODE == 2) Bp[(size_t)(c0 + c) + (size_t)row * d.ldb] = (EB)x[i];
            else Bp[row + (size_t)(c0 + c) * d.ldb] = (EB)x[i];
        }
    }
}

static void launch_trsm_diag(Ctx& c, int mode, const std::vector<LdiagDesc>& ds, int chunks, bool b32 = false) {
    if (ds.empty()) return;
    std::vector<int> map;
    map.reserve(chunks);
    for (size_t i = 0; i < ds.size(); ++i)
        map_append(map, (int)i, (i + 1 < ds.size() ? ds[i + 1].chunk0 : chunks) - ds[i].chunk0);
    LdiagDesc* dd = c.upload_vec(ds);
    const int* dm = c.upload_vec(map);
    KMH_CHECK(!b32 || mode == 0, KMH_E_BAD_ARG, "FP32 right-hand sides only for left lower solves");
    if (b32)
        KMH_LAUNCH(c, (k_trsm_diag<0, float>), dim3(chunks), dim3(256), dd, dm);
    else if (mode == 0)
        KMH_LAUNCH(c, (k_trsm_diag<0, double>), dim3(chunks), dim3(256), dd, dm);
    else if (mode == 1)
        KMH_LAUNCH(c, (k_trsm_diag<1, double>), dim3(chunks), dim3(256), dd, dm);
    else
        KMH_LAUNCH(c, (k_trsm_diag<2, double>), dim3(chunks), dim3(256), dd, dm);
}

void potrf(Ctx& c, const std::vector<SqMat>& mats) {
    if (mats.empty()) return;
    int nmax = 0;
    for (const auto& m : mats) nmax = std::max(nmax, m.n);
    for (int k0 = 0; k0 < nmax; k0 += PB) {
        std::vector<SqMat> diag;
        std::vector<LdiagDesc> pan;
        std::vector<GemmProb> upd;
        int chunks = 0;
        for (const auto& m : mats) {
            if (m.n <= k0) continue;
            const int nb = std::min(PB, m.n - k0);
            double* Akk = m.A + k0 + (size_t)k0 * m.ld;
            diag.push_back({Akk, nb, m.ld, m.info});
            const int rest = m.n - k0 - nb;
            if (rest > 0) {
                pan.push_back({Akk, nb, m.ld, Akk + nb, rest, m.ld, chunks, 0, Akk + (size_t)nb * m.ld});
                chunks += (int)ceil_div(rest, PB);
                upd.push_back({rest, rest, nb, Akk + nb, m.ld, Akk + nb, m.ld,
                               Akk + nb + (size_t)nb * m.ld, m.ld, -1.0, 1.0, 1});
            }
        }
        SqMat* dd = c.upload_vec(diag);
        KMH_LAUNCH(c, k_potrf_diag, dim3((unsigned)diag.size()), dim3(64), dd, c.d_info);
        if (!pan.empty()) {
            launch_trsm_diag(c, 2, pan, chunks);
            gemm(c, false, true, upd);
        }
    }
    // the strict upper triangle is zero: the diagonal blocks' by
    // k_potrf_diag, the rest by the panel solves (LdiagDesc::Z)
}

// ---------------------------------------------------------------------------
// Capacitance factors of lrsymfact (see la.h), one 256-thread workgroup per
// block, r <= CAP_MAX, everything in LDS in 16 x 16 blocks (the matrices are
// padded to nt = ceil(r / 16) blocks with the identity, which the Cholesky
// factor and its inverse keep):
//   1. Bm = R1 R2^T on v_mfma_f64_16x16x4f64 from the upper triangles
//      staged in LDS (output tile (bi, bj) starts at k = 16 max(bi, bj));
//      Bm to LDS and Bm^T straight to X's lower left block;
//   2. S = I - Bm^T Bm, the lower tiles on MFMA from LDS, into a packed
//      lower triangle in LDS;
//   3. blocked right-looking Cholesky: per block column p, wave 0 factors
//      the diagonal block in registers and replaces it by its inverse after
//      writing it out as part of X22 = L22 - I; the blocks below it
//      become S_ip L_pp^-T and the trailing blocks lose L_ip L_jp^T (MFMA);
//   4. blocked inverse, block columns from the last: with the diagonal
//      blocks already inverted, U_k = L_kj Linv_jj replaces L_kj, then
//      Linv_ij = -sum_{j<k<=i} Linv_ik U_k (MFMA);
//   5. G21 = L22^-1 Bm^T as its transpose Bm L22^-T on MFMA (a store writes a
//      16-row column segment), G22 = I - L22^-1, the log-diagonal sum.
constexpr int CAP_LB = CAP_MAX + 1;   // row stride of Bm in LDS

// element (i, j), i >= j, of the packed lower triangle (CAP_MAX rows, by columns)
__device__ __forceinline__ int cap_ix(int i, int j) { return j * CAP_MAX - ((j * (j - 1)) >> 1) + i - j; }

__global__ __launch_bounds__(256, 2) void k_capacitance(const CapDesc* __restrict__ ds) {
    __shared__ double Bs[CAP_MAX * CAP_LB];             // Bm(i, j) at i * CAP_LB + j
    __shared__ double Ls[CAP_MAX * (CAP_MAX + 1) / 2];  // S -> L22 (diagonal blocks: their inverses) -> L22^-1
    const CapDesc d = ds[blockIdx.x];
    const int r = d.r, t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
    if (r <= 0) {
        if (t == 0) *d.part = 0.0;
        return;
    }
    const int nt = (r + 15) >> 4, R = 16 * nt;
    const Glob<const double> R1 = glob(d.R1), R2 = glob(d.R2);
    const Glob<double> X = glob(d.X), G = glob(d.G);
    const d4 zero4 = {0.0, 0.0, 0.0, 0.0};
    // 1. Bm = R1 R2^T: R1 and R2 staged in LDS first (upper triangles
    //    packed by columns, R1 in Ls and R2 in Bs; all loads of a thread in
    //    flight together), the tiles in registers until every wave has read
    //    R2, then Bm to Bs and Bm^T to X's lower left block
    {
        constexpr int NL = (CAP_MAX * CAP_MAX + 255) / 256;
        double v1[NL], v2[NL];
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int e = t + 256 * q, j = e / CAP_MAX, i = e - j * CAP_MAX;
            const bool in = i <= j && j < r;
            const size_t off = in ? (size_t)i + (size_t)j * d.ldr : 0;
            v1[q] = keep_or_zero(R1[off], in);
            v2[q] = keep_or_zero(R2[off], in);
        }
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int e = t + 256 * q, j = e / CAP_MAX, i = e - j * CAP_MAX;
            if (i <= j && j < CAP_MAX) {
                Ls[((j * (j + 1)) >> 1) + i] = v1[q];
                Bs[((j * (j + 1)) >> 1) + i] = v2[q];
            }
        }
    }
    __syncthreads();
    auto rup = [](const double* P, int i, int j) { return i <= j ? P[((j * (j + 1)) >> 1) + i] : 0.0; };
    constexpr int BMT = ((CAP_MAX / 16) * (CAP_MAX / 16) + 3) / 4;   // tiles per wave
    d4 bmt[BMT];
#pragma unroll
    for (int q8 = 0; q8 < BMT; ++q8) {
        const int tile = wave + 4 * q8;
        d4 acc = zero4;
        if (tile < nt * nt) {
            const int bi = tile / nt, bj = tile - bi * nt;
            for (int k0 = 16 * max(bi, bj); k0 < r; k0 += 4) {
                const int k = k0 + lk;   // < CAP_MAX; R1(ra, k) and R2(rb, k) are zero beyond r
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(rup(Ls, 16 * bi + li, k), rup(Bs, 16 * bj + li, k), acc, 0, 0, 0);
            }
        }
        bmt[q8] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int q8 = 0; q8 < BMT; ++q8) {
        const int tile = wave + 4 * q8;
        if (tile < nt * nt) {
            const int bi = tile / nt, bj = tile - bi * nt, rb = 16 * bj + li;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = 16 * bi + mfma_d_row(lk, q);   // D(m, rb) = Bm(m, rb)
                if (m < r && rb < r) {
                    Bs[m * CAP_LB + rb] = bmt[q8][q];
                    X[(size_t)(r + rb) + (size_t)m * d.ld] = bmt[q8][q];   // X21 = Bm^T
                }
            }
        }
    }
    __syncthreads();
    // 2. S = I - Bm^T Bm (lower tiles), identity in the padding
    for (int tile = wave; tile < nt * (nt + 1) / 2; tile += 4) {
        int bi = 0;
        while ((bi + 1) * (bi + 2) / 2 <= tile) ++bi;
        const int bj = tile - bi * (bi + 1) / 2;
        const int ci = 16 * bi + li, cj = 16 * bj + li;   // < CAP_MAX
        d4 acc = zero4;
        for (int k0 = 0; k0 < r; k0 += 4) {
            const int k = k0 + lk;   // < CAP_MAX
            const double a = keep_or_zero(Bs[k * CAP_LB + ci], k < r && ci < r);   // Bm^T(ci, k)
            const double b = keep_or_zero(Bs[k * CAP_LB + cj], k < r && cj < r);   // Bm(k, cj)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int m = 16 * bi + mfma_d_row(lk, q);
            if (m >= cj) Ls[cap_ix(m, cj)] = (m == cj ? 1.0 : 0.0) - ((m < r && cj < r) ? acc[q] : 0.0);
        }
    }
    __syncthreads();
    // 3. blocked Cholesky
    double lsum = 0.0;   // wave 0, lane i < 16: log L22(16p + i, 16p + i) summed over p
    for (int p = 0; p < nt; ++p) {
        const int o = 16 * p;
        if (wave == 0) {
            // the diagonal block in registers, lane i < 16 holding row i
            // (dpotf2 arithmetic, values of other rows by lane broadcasts)
            const int ri = li;   // lanes >= 16 repeat lanes 0..15 and store nothing
            double a[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = k <= ri ? Ls[cap_ix(o + ri, o + k)] : 0.0;
            bool bad = false;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const double ajj = lane_bcast(a[j], j);
                bad |= (o + j < r) && !(ajj > 0.0);
                const double ljj = sqrt(ajj);
                a[j] = ri == j ? ljj : (ri > j ? a[j] / ljj : 0.0);
#pragma unroll
                for (int k = j + 1; k < 16; ++k) a[k] = __fma_rn(-a[j], lane_bcast(a[j], k), a[k]);
            }
            if (bad && lane == 0) atomicExch(d.info, 1);
            if (lane < 16 && o + ri < r) lsum += log(a[ri < 16 ? ri : 0]);
            // X22 = L22 - I on the diagonal block (zero above the diagonal)
            if (lane < 16 && o + ri < r) {
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    if (o + k < r) X[(size_t)(r + o + ri) + (size_t)(r + o + k) * d.ld] = k < ri ? a[k] : (k == ri ? a[k] - 1.0 : 0.0);
            }
            // its inverse, lane c < 16 solving column c: y_i = (delta_ic -
            // sum_{k<i} L(i,k) y_k) / L(i,i)
            double y[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                double sacc = i == ri ? 1.0 : 0.0;
#pragma unroll
                for (int k = 0; k < i; ++k) sacc = __fma_rn(-lane_bcast(a[k], i), y[k], sacc);
                y[i] = sacc / lane_bcast(a[i], i);
            }
            if (lane < 16) {
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    if (i >= ri) Ls[cap_ix(o + i, o + ri)] = y[i];
            }
        }
        __syncthreads();
        // the panel below: L_ip = S_ip L_pp^-T (one block per wave)
        for (int bi = p + 1 + wave; bi < nt; bi += 4) {
            d4 acc = zero4;
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int k = 4 * s4 + lk;
                const double a = Ls[cap_ix(16 * bi + li, o + k)];                      // S_ip(li, k)
                const double b = li >= k ? Ls[cap_ix(o + li, o + k)] : 0.0;            // Linv_pp(li, k)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < 4; ++q) Ls[cap_ix(16 * bi + mfma_d_row(lk, q), o + li)] = acc[q];
        }
        __syncthreads();
        // trailing blocks (bi >= bj > p): S_ij -= L_ip L_jp^T
        const int nrem = nt - 1 - p;
        for (int tile = wave; tile < nrem * (nrem + 1) / 2; tile += 4) {
            int ti = 0;
            while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
            const int bi = p + 1 + ti, bj = p + 1 + tile - ti * (ti + 1) / 2;
            d4 acc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = 16 * bi + mfma_d_row(lk, q), n = 16 * bj + li;
                acc[q] = m >= n ? Ls[cap_ix(m, n)] : 0.0;
            }
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int k = o + 4 * s4 + lk;
                const double a = -Ls[cap_ix(16 * bi + li, k)];
                const double b = Ls[cap_ix(16 * bj + li, k)];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = 16 * bi + mfma_d_row(lk, q), n = 16 * bj + li;
                if (m >= n) Ls[cap_ix(m, n)] = acc[q];
            }
        }
        __syncthreads();
    }
    // X22 below the diagonal blocks (L22) and above them (zero)
    for (int e = t; e < r * r; e += 256) {
        const int c = e / r, i = e - c * r;
        if ((i >> 4) != (c >> 4)) X[(size_t)(r + i) + (size_t)(r + c) * d.ld] = i > c ? Ls[cap_ix(i, c)] : 0.0;
    }
    // 4. blocked inverse, block columns from the last
    for (int bj = nt - 2; bj >= 0; --bj) {
        const int o = 16 * bj;
        d4 acc = zero4;
        const int bk = bj + 1 + wave;   // U_k = L_kj Linv_jj
        if (bk < nt) {
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int k = 4 * s4 + lk;
                const double a = Ls[cap_ix(16 * bk + li, o + k)];               // L_kj(li, k)
                const double b = k >= li ? Ls[cap_ix(o + k, o + li)] : 0.0;     // Linv_jj(k, li)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
        }
        __syncthreads();
        if (bk < nt) {
#pragma unroll
            for (int q = 0; q < 4; ++q) Ls[cap_ix(16 * bk + mfma_d_row(lk, q), o + li)] = acc[q];
        }
        __syncthreads();
        const int bi = bj + 1 + wave;   // Linv_ij = -sum_{j<k<=i} Linv_ik U_k
        acc = zero4;
        if (bi < nt) {
            for (int bk2 = bj + 1; bk2 <= bi; ++bk2) {
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    const int k = 4 * s4 + lk;
                    const int ii = 16 * bi + li, kk = 16 * bk2 + k;
                    const double a = ii >= kk ? -Ls[cap_ix(ii, kk)] : 0.0;          // -Linv_ik(li, k)
                    const double b = Ls[cap_ix(kk, o + li)];                        // U_k(k, li)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
                }
            }
        }
        __syncthreads();
        if (bi < nt) {
#pragma unroll
            for (int q = 0; q < 4; ++q) Ls[cap_ix(16 * bi + mfma_d_row(lk, q), o + li)] = acc[q];
        }
        __syncthreads();
    }
    // 5. G21 = L22^-1 Bm^T as Bm L22^-T; G22 = I - L22^-1
    auto linv = [&](int i, int k) { return i >= k ? Ls[cap_ix(i, k)] : 0.0; };
    for (int tile = wave; tile < nt * nt; tile += 4) {
        const int bi = tile / nt, bj = tile - bi * nt;
        const int mi = 16 * bi + li, nj = 16 * bj + li;   // < CAP_MAX
        d4 acc = zero4;
        const int kend = min(r, 16 * (bj + 1));
        for (int k0 = 0; k0 < kend; k0 += 4) {
            const int k = k0 + lk;   // < CAP_MAX
            const bool ok = k < r;
            const double a = keep_or_zero(Bs[mi * CAP_LB + k], ok && mi < r);   // Bm(mi, k)
            const double b = keep_or_zero(linv(nj, k), ok && nj < r);          // L22^-T(k, nj)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int m = 16 * bi + mfma_d_row(lk, q);   // D(m, nj) = G21(nj, m)
            if (m < r && nj < r) G[(size_t)(r + nj) + (size_t)m * d.ld] = acc[q];
        }
    }
    for (int e = t; e < r * r; e += 256) {
        const int c = e / r, i = e - c * r;
        G[(size_t)(r + i) + (size_t)(r + c) * d.ld] = (i == c ? 1.0 : 0.0) - linv(i, c);
    }
    if (wave == 0) {
        double tot = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) tot += lane_bcast(lsum, i);
        if (lane == 0) *d.part = tot;
    }
}

void capacitance(Ctx& c, const std::vector<CapDesc>& ds) {
    if (ds.empty()) return;
    for (const auto& d : ds)
        KMH_CHECK(d.r <= CAP_MAX && d.ld >= 2 * d.r && d.ldr >= d.r, KMH_E_BAD_ARG, "capacitance: block beyond its limits");
    CapDesc* dd = c.upload_vec(ds);
    KMH_LAUNCH(c, k_capacitance, dim3((unsigned)ds.size()), dim3(256), dd);
}

void trsm_left(Ctx& c, bool trans, const std::vector<TrsmProb>& probs) {
    // FP32 right-hand sides (left lower solves only) in launches of their own
    bool any32 = false, any64 = false;
    for (const auto& p : probs) (p.eb ? any32 : any64) = true;
    if (any32 && any64) {
        std::vector<TrsmProb> a, b;
        for (const auto& p : probs) (p.eb ? a : b).push_back(p);
        trsm_left(c, trans, b);
        trsm_left(c, trans, a);
        return;
    }
    KMH_CHECK(!any32 || !trans, KMH_E_BAD_ARG, "FP32 right-hand sides only for left lower solves");
    int maxblk = 0;
    for (const auto& p : probs) maxblk = std::max<int>(maxblk, (int)ceil_div(p.n, PB));
    for (int s = 0; s < maxblk; ++s) {
        std::vector<LdiagDesc> diag;
        std::vector<GemmProb> upd;
        int chunks = 0;
        for (const auto& p : probs) {
            const int nblk = (int)ceil_div(p.n, PB);
            if (s >= nblk || p.ncols <= 0) continue;
            const int kb = trans ? nblk - 1 - s : s;
            const int k0 = kb * PB;
            const int nb = std::min(PB, p.n - k0);
            diag.push_back({p.L + k0 + (size_t)k0 * p.ldl, nb, p.ldl, adv(p.B, k0, p.eb), p.ncols, p.ldb, chunks, 0});
            chunks += (int)ceil_div(p.ncols, PB);
            if (!trans) {
                const int rest = p.n - k0 - nb;
                if (rest > 0) {
                    GemmProb g{rest, p.ncols, nb, p.L + (k0 + nb) + (size_t)k0 * p.ldl, p.ldl,
                               adv(p.B, k0, p.eb), p.ldb, adv(p.B, k0 + nb, p.eb), p.ldb, -1.0, 1.0, 0};
                    g.eb = g.ec = p.eb;
                    upd.push_back(g);
                }
            } else if (k0 > 0) {
                upd.push_back({k0, p.ncols, nb, p.L + k0, p.ldl, p.B + k0, p.ldb, p.B, p.ldb, -1.0, 1.0, 0});
            }
        }
        if (diag.empty()) continue;
        launch_trsm_diag(c, trans ? 1 : 0, diag, chunks, any32);
        if (!upd.empty()) gemm(c, trans, false, upd);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Sum over the 256 threads of a block; every thread gets the same value.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    return ((sh[0] + sh[1]) + (sh[2] + sh[3]));
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
template <bool TR, class ES, class ED>
__global__ __launch_bounds__(256) void k_copy2d(const Copy2D* __restrict__ ds) {
    const Copy2D d = ds[blockIdx.y];
    const ES* src = reinterpret_cast<const ES*>(d.src);
    ED* dst = reinterpret_cast<ED*>(d.dst);
    // 256-row segments of one column per step (uniform index arithmetic)
    const int rb = (d.m + 255) >> 8;
    const long long segs = (long long)rb * d.n;
    for (long long t = blockIdx.x; t < segs; t += gridDim.x) {
        const int j = (int)(t / rb), i = (int)(t - (long long)j * rb) * 256 + (int)threadIdx.x;
        if (i >= d.m) continue;
        const double v = d.alpha * (double)src[i + (size_t)j * d.lds];
        if (TR) dst[j + (size_t)i * d.ldd] = (ED)v;
        else dst[i + (size_t)j * d.ldd] = (ED)v;
    }
}

void copy2d(Ctx& c, const std::vector<Copy2D>& probs, bool transpose) {
    if (probs.empty()) return;
    // one launch group per storage-type pair
    int seen = 0;
    for (const auto& p : probs) {
        KMH_CHECK(!transpose || (!p.es && !p.ed), KMH_E_BAD_ARG, "copy2d: no FP32 transposes");
        seen |= 1 << ((p.es ? 1 : 0) | (p.ed ? 2 : 0));
    }
    for (int tc = 0; tc < 4; ++tc) {
        if (!((seen >> tc) & 1)) continue;
        std::vector<Copy2D> sel;
        for (const auto& p : probs)
            if (((p.es ? 1 : 0) | (p.ed ? 2 : 0)) == tc) sel.push_back(p);
        for (size_t off = 0; off < sel.size(); off += 65535) {
            std::vector<Copy2D> part(sel.begin() + off, sel.begin() + std::min(sel.size(), off + 65535));
            Copy2D* d = c.upload_vec(part);
            long long mx = 0;
            for (auto& p : part) mx = std::max<long long>(mx, (long long)p.m * p.n);
            const unsigned gx = (unsigned)std::min<long long>(std::max<long long>(1, ceil_div(mx, 256)), 1024);
            const dim3 g(gx, (unsigned)part.size());
            if (transpose) KMH_LAUNCH(c, (k_copy2d<true, double, double>), g, dim3(256), d);
            else if (tc == 0) KMH_LAUNCH(c, (k_copy2d<false, double, double>), g, dim3(256), d);
            else if (tc == 1) KMH_LAUNCH(c, (k_copy2d<false, float, double>), g, dim3(256), d);
            else if (tc == 2) KMH_LAUNCH(c, (k_copy2d<false, double, float>), g, dim3(256), d);
            else KMH_LAUNCH(c, (k_copy2d<false, float, float>), g, dim3(256), d);
        }
    }
}

__global__ __launch_bounds__(256) void k_axpby2d(const Axpby2D* __restrict__ ds) {
    const Axpby2D d = ds[blockIdx.y];
    const int rb = (d.m + 255) >> 8;
    const long long segs = (long long)rb * d.n;
    for (long long t = blockIdx.x; t < segs; t += gridDim.x) {
        const int j = (int)(t / rb), i = (int)(t - (long long)j * rb) * 256 + (int)threadIdx.x;
        if (i >= d.m) continue;
        double* p = d.dst + i + (size_t)j * d.ldd;
        const double s = d.src[i + (size_t)j * d.lds];
        *p = (d.a == 0.0 ? 0.0 : d.a * (*p)) + d.b * s;
    }
}

void axpby2d(Ctx& c, const std::vector<Axpby2D>& probs) {
    if (probs.empty()) return;
    for (size_t off = 0; off < probs.size(); off += 65535) {
        std::vector<Axpby2D> part(probs.begin() + off, probs.begin() + std::min(probs.size(), off + 65535));
        Axpby2D* d = c.upload_vec(part);
        long long mx = 0;
        for (auto& p : part) mx = std::max<long long>(mx, (long long)p.m * p.n);
        const unsigned gx = (unsigned)std::min<long long>(std::max<long long>(1, ceil_div(mx, 256)), 1024);
        KMH_LAUNCH(c, k_axpby2d, dim3(gx, (unsigned)part.size()), dim3(256), d);
    }
}

__global__ __launch_bounds__(256) void k_add2(const Add2* __restrict__ ds) {
    const Add2 d = ds[blockIdx.y];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * 256)
        d.out[i] = d.x[i] + d.y[i];
}

void add2(Ctx& c, const std::vector<Add2>& probs) {
    if (probs.empty()) return;
    for (size_t off = 0; off < probs.size(); off += 65535) {
        std::vector<Add2> part(probs.begin() + off, probs.begin() + std::min(probs.size(), off + 65535));
        Add2* d = c.upload_vec(part);
        int64_t mx = 0;
        for (auto& p : part) mx = std::max(mx, p.n);
        const unsigned gx = (unsigned)std::min<int64_t>(std::max<int64_t>(1, ceil_div(mx, 256)), 1024);
        KMH_LAUNCH(c, k_add2, dim3(gx, (unsigned)part.size()), dim3(256), d);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_gather(const T* const* __restrict__ src, T* __restrict__ dst, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = *src[i];
}

__global__ __launch_bounds__(256) void k_zero_ints(int* const* __restrict__ p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) *p[i] = 0;
}

void gather_doubles(Ctx& c, const std::vector<const double*>& src, double* d_out) {
    if (src.empty()) return;
    const double* const* d = c.upload_vec(src);
    KMH_LAUNCH(c, k_gather<double>, dim3((unsigned)ceil_div((int64_t)src.size(), 256)), dim3(256), d, d_out,
               (int)src.size());
}

void gather_ints(Ctx& c, const std::vector<const int*>& src, int* d_out) {
    if (src.empty()) return;
    const int* const* d = c.upload_vec(src);
    KMH_LAUNCH(c, k_gather<int>, dim3((unsigned)ceil_div((int64_t)src.size(), 256)), dim3(256), d, d_out,
               (int)src.size());
}

void zero_ints(Ctx& c, const std::vector<int*>& ptrs) {
    if (ptrs.empty()) return;
    int* const* d = c.upload_vec(ptrs);
    KMH_LAUNCH(c, k_zero_ints, dim3((unsigned)ceil_div((int64_t)ptrs.size(), 256)), dim3(256), d, (int)ptrs.size());
}

template <bool SET>
__global__ __launch_bounds__(256) void k_diag_op(const DiagOp* __restrict__ ds) {
    const DiagOp d = ds[blockIdx.y];
    if (SET) {
        const long long tot = (long long)d.n * d.n;
        for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
            const int i = (int)(e % d.n), j = (int)(e / d.n);
            d.A[i + (size_t)j * d.ld] = (i == j) ? d.v : 0.0;
        }
    } else {
        for (int i = blockIdx.x * 256 + threadIdx.x; i < d.n; i += gridDim.x * 256) d.A[i + (size_t)i * d.ld] += d.v;
    }
}

static void diag_launch(Ctx& c, const std::vector<DiagOp>& mats, bool set) {
    if (mats.empty()) return;
    for (size_t off = 0; off < mats.size(); off += 65535) {
        std::vector<DiagOp> part(mats.begin() + off, mats.begin() + std::min(mats.size(), off + 65535));
        DiagOp* d = c.upload_vec(part);
        long long mx = 0;
        for (auto& p : part) mx = std::max<long long>(mx, set ? (long long)p.n * p.n : p.n);
        const unsigned gx = (unsigned)std::min<long long>(std::max<long long>(1, ceil_div(mx, 256)), 1024);
        if (set) KMH_LAUNCH(c, (k_diag_op<true>), dim3(gx, (unsigned)part.size()), dim3(256), d);
        else KMH_LAUNCH(c, (k_diag_op<false>), dim3(gx, (unsigned)part.size()), dim3(256), d);
    }
}
void set_identity(Ctx& c, const std::vector<DiagOp>& mats) { diag_launch(c, mats, true); }
void add_diag(Ctx& c, const std::vector<DiagOp>& mats) { diag_launch(c, mats, false); }

struct LogDiagDesc {
    const double* A;
    int n, ld;
    double* out;
};

__global__ __launch_bounds__(256) void k_log_diag(const LogDiagDesc* __restrict__ ms) {
    __shared__ double sh[4];
    const LogDiagDesc m = ms[blockIdx.x];
    double s = 0.0;
    for (int i = threadIdx.x; i < m.n; i += 256) s += log(m.A[i + (size_t)i * m.ld]);
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) *m.out = t;
}

void log_diag_sums(Ctx& c, const std::vector<SqMat>& mats, const std::vector<double*>& outs) {
    if (mats.empty()) return;
    std::vector<LogDiagDesc> ds;
    for (size_t j = 0; j < mats.size(); ++j) ds.push_back({mats[j].A, mats[j].n, mats[j].ld, outs[j]});
    LogDiagDesc* d = c.upload_vec(ds);
    KMH_LAUNCH(c, k_log_diag, dim3((unsigned)ds.size()), dim3(256), d);
}

void log_diag_sums(Ctx& c, const std::vector<SqMat>& mats, double* d_out) {
    std::vector<double*> outs;
    for (size_t j = 0; j < mats.size(); ++j) outs.push_back(d_out + j);
    log_diag_sums(c, mats, outs);
}

__global__ __launch_bounds__(256) void k_sum_vec(const SumJob* __restrict__ jobs) {
    __shared__ double sh[4];
    const SumJob j = jobs[blockIdx.x];
    double s = 0.0;
    for (int i = threadIdx.x; i < j.count; i += 256) s += j.in[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) *j.out = t;
}

void sum_vectors(Ctx& c, const std::vector<SumJob>& jobs) {
    if (jobs.empty()) return;
    SumJob* d = c.upload_vec(jobs);
    KMH_LAUNCH(c, k_sum_vec, dim3((unsigned)jobs.size()), dim3(256), d);
}

void sum_vector(Ctx& c, const double* d_in, int count, double* d_out) { sum_vectors(c, {{d_in, count, d_out}}); }

// Dot products in two deterministic stages: fixed 8192-element slices, one
// workgroup each, then one workgroup per job adds its slices in order.
constexpr int DOT_SLICE = 8192;

struct DotSlice {
    const double* x;
    const double* y;
    int len;
    int pad;
    double* part;
};

__global__ __launch_bounds__(256) void k_dot_slices(const DotSlice* __restrict__ ds) {
    __shared__ double sh[4];
    const DotSlice d = ds[blockIdx.x];
    double s0 = 0.0, s1 = 0.0;
    int i = threadIdx.x;
    for (; i + 256 < d.len; i += 512) {
        s0 += d.x[i] * d.y[i];
        s1 += d.x[i + 256] * d.y[i + 256];
    }
    if (i < d.len) s0 += d.x[i] * d.y[i];
    const double t = block_sum(s0 + s1, sh);
    if (threadIdx.x == 0) *d.part = t;
}

struct DotSum {
    const double* part;
    int count;
    int pad;
    double* out;
};

__global__ __launch_bounds__(64) void k_dot_sums(const DotSum* __restrict__ ds) {
    const DotSum d = ds[blockIdx.x];
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int q = 0; q < d.count; ++q) s += d.part[q];
    *d.out = s;
}

void dots_batch(Ctx& c, const std::vector<DotJob>& jobs) {
    if (jobs.empty()) return;
    std::vector<DotSlice> sl;
    std::vector<DotSum> sm;
    size_t np = 0;
    for (const auto& j : jobs) np += std::max<int64_t>(1, ceil_div(j.m, DOT_SLICE));
    double* part = static_cast<double*>(c.scratch(np * sizeof(double)));
    size_t q = 0;
    for (const auto& j : jobs) {
        const int ns = (int)std::max<int64_t>(1, ceil_div(j.m, DOT_SLICE));
        sm.push_back({part + q, ns, 0, j.out});
        for (int s = 0; s < ns; ++s) {
            const int64_t i0 = (int64_t)s * DOT_SLICE;
            sl.push_back({j.x + i0, j.y + i0, (int)std::max<int64_t>(0, std::min<int64_t>(DOT_SLICE, j.m - i0)), 0,
                          part + q + s});
        }
        q += ns;
    }
    DotSlice* dsl = c.upload_vec(sl);
    KMH_LAUNCH(c, k_dot_slices, dim3((unsigned)sl.size()), dim3(256), dsl);
    DotSum* dsm = c.upload_vec(sm);
    KMH_LAUNCH(c, k_dot_sums, dim3((unsigned)sm.size()), dim3(64), dsm);
    c.scratch_free(part);
}

void column_dots(Ctx& c, const double* X, int ldx, const double* Y, int ldy, int m, int ncols,
                 double* d_out) {
    std::vector<DotJob> jobs;
    for (int j = 0; j < ncols; ++j) jobs.push_back({X + (size_t)j * ldx, Y + (size_t)j * ldy, m, d_out + j});
    dots_batch(c, jobs);
}

}  // namespace kmh
