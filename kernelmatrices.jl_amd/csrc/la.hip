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
//    blocks factored or solved by one workgroup, trailing updates by GEMM.
//  * QR: Householder QR of <=256 x 64 tiles held in LDS (130 KB), combined
//    by a binary TSQR tree; columns beyond 64 by block Gram-Schmidt with one
//    reorthogonalisation pass.
#include <algorithm>
#include <cmath>

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
};

constexpr int GT = 64;
constexpr int GK = 16;

// Lane layout of V_MFMA_F64_16X16X4F64 on gfx950 (measured with one-hot
// operands): A(i,k) is held by lane i + 16k, B(k,j) by lane j + 16k, and
// accumulator register r of lane j + 16g holds D(4r + g, j).  (The FP32
// 16x16x4 form instead holds D(4g + r, j).)
__device__ __forceinline__ int mfma_d_row(int g, int r) { return 4 * r + g; }

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc* __restrict__ ds, int nd) {
    __shared__ double As[GK][GT + 1];
    __shared__ double Bs[GK][GT + 1];
    const int t = blockIdx.x;
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ds[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const GemmDesc d = ds[lo];
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

    for (int k0 = 0; k0 < d.k; k0 += GK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int i, kk;
            if (!TA) { i = tid & 63; kk = (tid >> 6) + 4 * q; }
            else     { kk = tid & 15; i = (tid >> 4) + 16 * q; }
            const int gi = row0 + i, gk = k0 + kk;
            double v = 0.0;
            if (gi < d.m && gk < d.k)
                v = TA ? d.A[gk + (size_t)gi * d.lda] : d.A[gi + (size_t)gk * d.lda];
            As[kk][i] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int j, kk;
            if (!TB) { kk = tid & 15; j = (tid >> 4) + 16 * q; }
            else     { j = tid & 63; kk = (tid >> 6) + 4 * q; }
            const int gj = col0 + j, gk = k0 + kk;
            double v = 0.0;
            if (gj < d.n && gk < d.k)
                v = TB ? d.B[gj + (size_t)gk * d.ldb] : d.B[gk + (size_t)gj * d.ldb];
            Bs[kk][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            const double a0 = As[kk][wr + (lane & 15)];
            const double a1 = As[kk][wr + 16 + (lane & 15)];
            const double b0 = Bs[kk][wc + (lane & 15)];
            const double b1 = Bs[kk][wc + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = row0 + wr + mb * 16 + mfma_d_row(lane >> 4, j);
                const int col = col0 + wc + nb * 16 + (lane & 15);
                if (row < d.m && col < d.n) {
                    double* cp = d.C + row + (size_t)col * d.ldc;
                    double v = d.alpha * acc[mb][nb][j];
                    if (d.beta != 0.0) v += d.beta * (*cp);
                    *cp = v;
                }
            }
}

struct RedDesc {
    double* C;
    const double* W;
    int ldc, m, n, s;
    double alpha, beta;
    long long elem0;
};

__global__ __launch_bounds__(256) void k_splitk_reduce(const RedDesc* __restrict__ ds, int nd,
                                                       long long total) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ds[mid].elem0 <= e) lo = mid;
        else hi = mid - 1;
    }
    const RedDesc d = ds[lo];
    const long long loc = e - d.elem0;
    const int i = (int)(loc % d.m), j = (int)(loc / d.m);
    const size_t mn = (size_t)d.m * d.n;
    double s = 0.0;
    for (int q = 0; q < d.s; ++q) s += d.W[q * mn + loc];
    double* cp = d.C + i + (size_t)j * d.ldc;
    double v = d.alpha * s;
    if (d.beta != 0.0) v += d.beta * (*cp);
    *cp = v;
}

void gemm(Ctx& c, bool ta, bool tb, const std::vector<GemmProb>& probs) {
    long long total_tiles = 0;
    for (const auto& p : probs)
        if (p.m > 0 && p.n > 0) total_tiles += ceil_div(p.m, GT) * ceil_div(p.n, GT);
    if (total_tiles == 0) return;
    // Split-k decision from the problem's own shape only, so a problem gives
    // bitwise the same result whatever else shares the launch.
    std::vector<int> split(probs.size(), 1), kchunk(probs.size(), 0);
    size_t wsz = 0;
    for (size_t i = 0; i < probs.size(); ++i) {
        const auto& p = probs[i];
        if (p.m <= 0 || p.n <= 0 || p.lower) continue;
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
    double* W = wsz ? static_cast<double*>(c.scratch(wsz * sizeof(double))) : nullptr;
    std::vector<GemmDesc> ds;
    std::vector<RedDesc> rs;
    int tiles = 0;
    size_t woff = 0;
    long long relem = 0;
    for (size_t i = 0; i < probs.size(); ++i) {
        const auto& p = probs[i];
        if (p.m <= 0 || p.n <= 0) continue;
        const int tn = (int)ceil_div(p.n, GT);
        const int nt = (int)ceil_div(p.m, GT) * tn;
        if (split[i] == 1) {
            ds.push_back({p.A, p.B, p.C, p.m, p.n, p.k, p.lda, p.ldb, p.ldc, p.alpha, p.beta,
                          tiles, tn, p.lower, 0});
            tiles += nt;
        } else {
            const int s = split[i];
            const int kc = kchunk[i];
            for (int q = 0; q < s; ++q) {
                const int k0 = q * kc;
                const int kk = std::min(kc, p.k - k0);
                const double* A = ta ? p.A + k0 : p.A + (size_t)k0 * p.lda;
                const double* B = tb ? p.B + (size_t)k0 * p.ldb : p.B + k0;
                ds.push_back({A, B, W + woff + (size_t)q * p.m * p.n, p.m, p.n, kk, p.lda, p.ldb,
                              p.m, 1.0, 0.0, tiles, tn, 0, 0});
                tiles += nt;
            }
            rs.push_back({p.C, W + woff, p.ldc, p.m, p.n, s, p.alpha, p.beta, relem});
            woff += (size_t)s * p.m * p.n;
            relem += (long long)p.m * p.n;
        }
    }
    GemmDesc* dd = c.upload_vec(ds);
    const int nd = (int)ds.size();
    if (c.prof) {
        double fl = 0.0;
        for (const auto& d : ds) fl += 2.0 * d.m * d.n * (double)d.k;
        c.ev_flops.push_back(fl);
        KMH_HIP(hipEventRecord(c.prof_event(), c.stream));
    }
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), dim3(tiles), dim3(256), 0, c.stream, dd, nd);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), dim3(tiles), dim3(256), 0, c.stream, dd, nd);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), dim3(tiles), dim3(256), 0, c.stream, dd, nd);
    else hipLaunchKernelGGL((k_gemm<true, true>), dim3(tiles), dim3(256), 0, c.stream, dd, nd);
    KMH_HIP(hipGetLastError());
    if (c.prof) KMH_HIP(hipEventRecord(c.prof_event(), c.stream));
    if (!rs.empty()) {
        RedDesc* dr = c.upload_vec(rs);
        hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)ceil_div(relem, 256)), dim3(256), 0, c.stream,
                           dr, (int)rs.size(), relem);
        KMH_HIP(hipGetLastError());
    }
    if (W) c.scratch_free(W);
}

// ---------------------------------------------------------------------------
// POTRF (lower) and TRSM
// ---------------------------------------------------------------------------
constexpr int PB = 64;

__global__ __launch_bounds__(256) void k_potrf_diag(const SqMat* __restrict__ ms, int* info0) {
    __shared__ double S[PB][PB + 1];   // S[col][row]
    const SqMat m = ms[blockIdx.x];
    int* info = m.info ? m.info : info0;
    const int n = m.n, tid = threadIdx.x;
    for (int e = tid; e < PB * PB; e += 256) {
        const int i = e % PB, j = e / PB;
        S[j][i] = (i < n && j < n && i >= j) ? m.A[i + (size_t)j * m.ld] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (tid == 0) {
            const double djj = S[j][j];
            if (!(djj > 0.0)) atomicExch(info, 1);
            S[j][j] = sqrt(djj);
        }
        __syncthreads();
        const double ljj = S[j][j];
        for (int i = j + 1 + tid; i < n; i += 256) S[j][i] = S[j][i] / ljj;
        __syncthreads();
        const int w = n - j - 1;
        for (int e = tid; e < w * w; e += 256) {
            const int l = j + 1 + e / w, i = j + 1 + e % w;
            if (i >= l) S[l][i] -= S[j][i] * S[j][l];
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += 256) {
        const int i = e % n, j = e / n;
        m.A[i + (size_t)j * m.ld] = (i >= j) ? S[j][i] : 0.0;
    }
}

// X L^T = B for the rows of B (one thread per row), L is nb x nb lower.
struct RltDesc {
    const double* L;
    int nb, ldl;
    double* B;
    int m, ldb;
    int chunk0, pad;
};

__global__ __launch_bounds__(256) void k_trsm_rlt(const RltDesc* __restrict__ ds, int nd) {
    __shared__ double Ls[PB][PB + 1];   // Ls[j][l] = L(j, l), identity padding
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ds[mid].chunk0 <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const RltDesc d = ds[lo];
    const int tid = threadIdx.x;
    for (int e = tid; e < PB * PB; e += 256) {
        const int j = e % PB, l = e / PB;
        double v;
        if (j < d.nb && l < d.nb) v = (l <= j) ? d.L[j + (size_t)l * d.ldl] : 0.0;
        else v = (j == l) ? 1.0 : 0.0;
        Ls[j][l] = v;
    }
    __syncthreads();
    const int row = ((int)blockIdx.x - d.chunk0) * 256 + tid;
    if (row >= d.m) return;
    double x[PB];
#pragma unroll
    for (int j = 0; j < PB; ++j) x[j] = (j < d.nb) ? d.B[row + (size_t)j * d.ldb] : 0.0;
#pragma unroll
    for (int j = 0; j < PB; ++j) {
        double s = x[j];
#pragma unroll
        for (int l = 0; l < j; ++l) s -= x[l] * Ls[j][l];
        x[j] = s / Ls[j][j];
    }
#pragma unroll
    for (int j = 0; j < PB; ++j)
        if (j < d.nb) d.B[row + (size_t)j * d.ldb] = x[j];
}

// Diagonal-block solve for the left TRSM: one thread per right-hand column.
struct LdiagDesc {
    const double* L;
    int nb, ldl;
    double* B;
    int ncols, ldb;
    int chunk0, pad;
};

template <bool TRANS>
__global__ __launch_bounds__(256) void k_trsm_ldiag(const LdiagDesc* __restrict__ ds, int nd) {
    __shared__ double Ls[PB][PB + 1];   // Ls[j][l] = L(j, l), identity padding
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ds[mid].chunk0 <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const LdiagDesc d = ds[lo];
    const int tid = threadIdx.x;
    for (int e = tid; e < PB * PB; e += 256) {
        const int j = e % PB, l = e / PB;
        double v;
        if (j < d.nb && l < d.nb) v = (l <= j) ? d.L[j + (size_t)l * d.ldl] : 0.0;
        else v = (j == l) ? 1.0 : 0.0;
        Ls[j][l] = v;
    }
    __syncthreads();
    const int col = ((int)blockIdx.x - d.chunk0) * 256 + tid;
    if (col >= d.ncols) return;
    double* b = d.B + (size_t)col * d.ldb;
    double x[PB];
#pragma unroll
    for (int j = 0; j < PB; ++j) x[j] = (j < d.nb) ? b[j] : 0.0;
    if (!TRANS) {
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            double s = x[j];
#pragma unroll
            for (int l = 0; l < j; ++l) s -= Ls[j][l] * x[l];
            x[j] = s / Ls[j][j];
        }
    } else {
#pragma unroll
        for (int j = PB - 1; j >= 0; --j) {
            double s = x[j];
#pragma unroll
            for (int l = j + 1; l < PB; ++l) s -= Ls[l][j] * x[l];
            x[j] = s / Ls[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < PB; ++j)
        if (j < d.nb) b[j] = x[j];
}

__global__ __launch_bounds__(256) void k_zero_upper(const SqMat* __restrict__ ms) {
    const SqMat m = ms[blockIdx.y];
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < (long long)m.n * m.n;
         e += (long long)gridDim.x * 256) {
        const int i = (int)(e % m.n), j = (int)(e / m.n);
        if (i < j) m.A[i + (size_t)j * m.ld] = 0.0;
    }
}

void potrf(Ctx& c, const std::vector<SqMat>& mats) {
    if (mats.empty()) return;
    int nmax = 0;
    for (const auto& m : mats) nmax = std::max(nmax, m.n);
    for (int k0 = 0; k0 < nmax; k0 += PB) {
        std::vector<SqMat> diag;
        std::vector<RltDesc> pan;
        std::vector<GemmProb> upd;
        int chunks = 0;
        for (const auto& m : mats) {
            if (m.n <= k0) continue;
            const int nb = std::min(PB, m.n - k0);
            double* Akk = m.A + k0 + (size_t)k0 * m.ld;
            diag.push_back({Akk, nb, m.ld, m.info});
            const int rest = m.n - k0 - nb;
            if (rest > 0) {
                pan.push_back({Akk, nb, m.ld, Akk + nb, rest, m.ld, chunks, 0});
                chunks += (int)ceil_div(rest, 256);
                upd.push_back({rest, rest, nb, Akk + nb, m.ld, Akk + nb, m.ld,
                               Akk + nb + (size_t)nb * m.ld, m.ld, -1.0, 1.0, 1});
            }
        }
        SqMat* dd = c.upload_vec(diag);
        hipLaunchKernelGGL(k_potrf_diag, dim3((unsigned)diag.size()), dim3(256), 0, c.stream, dd, c.d_info);
        KMH_HIP(hipGetLastError());
        if (!pan.empty()) {
            RltDesc* dp = c.upload_vec(pan);
            hipLaunchKernelGGL(k_trsm_rlt, dim3(chunks), dim3(256), 0, c.stream, dp, (int)pan.size());
            KMH_HIP(hipGetLastError());
            gemm(c, false, true, upd);
        }
    }
    SqMat* dm = c.upload_vec(mats);
    hipLaunchKernelGGL(k_zero_upper, dim3(16, (unsigned)mats.size()), dim3(256), 0, c.stream, dm);
    KMH_HIP(hipGetLastError());
}

void trsm_left(Ctx& c, bool trans, const std::vector<TrsmProb>& probs) {
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
            diag.push_back({p.L + k0 + (size_t)k0 * p.ldl, nb, p.ldl, p.B + k0, p.ncols, p.ldb, chunks, 0});
            chunks += (int)ceil_div(p.ncols, 256);
            if (!trans) {
                const int rest = p.n - k0 - nb;
                if (rest > 0)
                    upd.push_back({rest, p.ncols, nb, p.L + (k0 + nb) + (size_t)k0 * p.ldl, p.ldl,
                                   p.B + k0, p.ldb, p.B + k0 + nb, p.ldb, -1.0, 1.0, 0});
            } else if (k0 > 0) {
                upd.push_back({k0, p.ncols, nb, p.L + k0, p.ldl, p.B + k0, p.ldb, p.B, p.ldb, -1.0, 1.0, 0});
            }
        }
        if (diag.empty()) continue;
        LdiagDesc* dd = c.upload_vec(diag);
        if (!trans)
            hipLaunchKernelGGL((k_trsm_ldiag<false>), dim3(chunks), dim3(256), 0, c.stream, dd, (int)diag.size());
        else
            hipLaunchKernelGGL((k_trsm_ldiag<true>), dim3(chunks), dim3(256), 0, c.stream, dd, (int)diag.size());
        KMH_HIP(hipGetLastError());
        if (!upd.empty()) gemm(c, trans, false, upd);
    }
}

// ---------------------------------------------------------------------------
// Householder QR of small tiles (TSQR building block)
// ---------------------------------------------------------------------------
constexpr int QT = 256;   // max rows of a tile
constexpr int QC = 64;    // max columns

struct SqrDesc {
    const double* src0; int rows0, ld0;
    const double* src1; int rows1, ld1;
    double* qout; int ldq;
    double* rout; int ldr;
    int nc, pad;
};

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

__global__ __launch_bounds__(256) void k_small_qr(const SqrDesc* __restrict__ ds) {
    __shared__ double A[QC][QT + 1];   // A[col][row]
    __shared__ double tau[QC];
    __shared__ double w[QC];
    __shared__ double part[4][QC];
    __shared__ double sh[4];
    const SqrDesc d = ds[blockIdx.x];
    const int mt = d.rows0 + d.rows1, nc = d.nc, tid = threadIdx.x;
    for (int e = tid; e < mt * nc; e += 256) {
        const int i = e % mt, j = e / mt;
        A[j][i] = (i < d.rows0) ? d.src0[i + (size_t)j * d.ld0] : d.src1[(i - d.rows0) + (size_t)j * d.ld1];
    }
    __syncthreads();
    const int cc = tid & 63, pt = tid >> 6;
    for (int j = 0; j < nc; ++j) {
        // xnorm^2 over rows j+1..mt-1 of column j
        double s = 0.0;
        for (int i = j + 1 + tid; i < mt; i += 256) s += A[j][i] * A[j][i];
        const double x2 = block_sum(s, sh);
        const double alpha = A[j][j];
        double tj, beta, scal;
        if (x2 == 0.0) {
            tj = 0.0; beta = alpha; scal = 0.0;
        } else {
            beta = -copysign(sqrt(alpha * alpha + x2), alpha);
            tj = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        __syncthreads();
        if (x2 != 0.0)
            for (int i = j + 1 + tid; i < mt; i += 256) A[j][i] *= scal;
        if (tid == 0) { A[j][j] = beta; tau[j] = tj; }
        __syncthreads();
        if (tj != 0.0 && j + 1 < nc) {
            // w_c = A(j,c) + sum_{i>j} v_i A(i,c)
            const int c = j + 1 + cc;
            double p = 0.0;
            if (c < nc)
                for (int i = j + 1 + pt; i < mt; i += 4) p += A[j][i] * A[c][i];
            part[pt][cc] = p;
            __syncthreads();
            if (tid < 64 && c < nc) w[cc] = A[c][j] + ((part[0][cc] + part[1][cc]) + (part[2][cc] + part[3][cc]));
            __syncthreads();
            for (int i = j + tid; i < mt; i += 256) {
                const double vi = (i == j) ? 1.0 : A[j][i];
                const double tv = tj * vi;
                for (int c2 = j + 1; c2 < nc; ++c2) A[c2][i] -= tv * w[c2 - j - 1];
            }
            __syncthreads();
        }
    }
    // R (nc x nc upper; lower zeroed)
    for (int e = tid; e < nc * nc; e += 256) {
        const int i = e % nc, j = e / nc;
        d.rout[i + (size_t)j * d.ldr] = (i <= j) ? A[j][i] : 0.0;
    }
    __syncthreads();
    // explicit Q in place (LAPACK dorg2r order)
    for (int j = nc - 1; j >= 0; --j) {
        const double tj = tau[j];
        if (j + 1 < nc && tj != 0.0) {
            const int c = j + 1 + cc;
            double p = 0.0;
            if (c < nc)
                for (int i = j + 1 + pt; i < mt; i += 4) p += A[j][i] * A[c][i];
            part[pt][cc] = p;
            __syncthreads();
            if (tid < 64 && c < nc) w[cc] = A[c][j] + ((part[0][cc] + part[1][cc]) + (part[2][cc] + part[3][cc]));
            __syncthreads();
            for (int i = j + tid; i < mt; i += 256) {
                const double vi = (i == j) ? 1.0 : A[j][i];
                const double tv = tj * vi;
                for (int c2 = j + 1; c2 < nc; ++c2) A[c2][i] -= tv * w[c2 - j - 1];
            }
            __syncthreads();
        } else if (j + 1 < nc) {
            // H_j = I: the rows above and at j of later columns are untouched
            __syncthreads();
        }
        for (int i = tid; i < mt; i += 256) {
            double v;
            if (i < j) v = 0.0;
            else if (i == j) v = 1.0 - tj;
            else v = -tj * A[j][i];
            A[j][i] = v;
        }
        __syncthreads();
    }
    for (int e = tid; e < mt * nc; e += 256) {
        const int i = e % mt, j = e / mt;
        d.qout[i + (size_t)j * d.ldq] = A[j][i];
    }
}

struct TsqrPanel {
    double* A; int m; int nb; int lda;
    double* R; int ldr;
};

// Thin QR of m x nb panels (nb <= 64, m >= nb) by TSQR.
static void tsqr(Ctx& c, const std::vector<TsqrPanel>& ps) {
    if (ps.empty()) return;
    struct Node { int panel; const double* rptr; int rld; };
    struct Tile { int panel; int row0; int rows; int node; };
    std::vector<Tile> tiles;
    std::vector<std::vector<int>> cur(ps.size());
    int nnodes = 0;
    for (size_t p = 0; p < ps.size(); ++p) {
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
    // count internal nodes
    int ninternal = 0;
    for (size_t p = 0; p < ps.size(); ++p) {
        int k = (int)cur[p].size();
        while (k > 1) { ninternal += k / 2; k = k - k / 2; }
    }
    const int total_nodes = nnodes + ninternal;
    const size_t RS = (size_t)QC * QC;          // R slot, ld 64
    const size_t QS = (size_t)2 * QC * QC;      // Q' slot, 128 x 64, ld 128
    double* RT = static_cast<double*>(c.scratch(std::max<size_t>(1, total_nodes) * RS * sizeof(double)));
    double* QTb = static_cast<double*>(c.scratch(std::max(1, ninternal) * QS * sizeof(double)));
    double* CS = static_cast<double*>(c.scratch(std::max<size_t>(1, total_nodes) * RS * sizeof(double)));
    // tile QRs
    {
        std::vector<SqrDesc> ds;
        for (const auto& t : tiles) {
            const auto& P = ps[t.panel];
            const bool single = cur[t.panel].size() == 1;
            double* a = P.A + t.row0;
            ds.push_back({a, t.rows, P.lda, nullptr, 0, 0, a, P.lda,
                          single ? P.R : RT + (size_t)t.node * RS, single ? P.ldr : QC, P.nb, 0});
        }
        SqrDesc* dd = c.upload_vec(ds);
        hipLaunchKernelGGL(k_small_qr, dim3((unsigned)ds.size()), dim3(256), 0, c.stream, dd);
        KMH_HIP(hipGetLastError());
    }
    // reduction tree
    struct Comb { int panel; int node; int a; int b; int level; };
    std::vector<Comb> combs;
    int next_internal = 0;
    int level = 0;
    std::vector<int> root(ps.size(), -1);
    for (size_t p = 0; p < ps.size(); ++p)
        if (cur[p].size() == 1) root[p] = cur[p][0];
    for (;;) {
        std::vector<SqrDesc> ds;
        for (size_t p = 0; p < ps.size(); ++p) {
            auto& L = cur[p];
            if (L.size() <= 1) continue;
            std::vector<int> nxt;
            const bool final_level = (L.size() == 2);
            for (size_t i = 0; i + 1 < L.size(); i += 2) {
                const int node = nnodes + next_internal;
                const int qi = next_internal++;
                combs.push_back({(int)p, node, L[i], L[i + 1], level});
                const bool is_root = final_level;
                ds.push_back({RT + (size_t)L[i] * RS, ps[p].nb, QC, RT + (size_t)L[i + 1] * RS, ps[p].nb, QC,
                              QTb + (size_t)qi * QS, 2 * QC,
                              is_root ? ps[p].R : RT + (size_t)node * RS, is_root ? ps[p].ldr : QC,
                              ps[p].nb, 0});
                nxt.push_back(node);
                if (is_root) root[p] = node;
            }
            if (L.size() % 2) nxt.push_back(L.back());
            L = nxt;
        }
        if (ds.empty()) break;
        SqrDesc* dd = c.upload_vec(ds);
        hipLaunchKernelGGL(k_small_qr, dim3((unsigned)ds.size()), dim3(256), 0, c.stream, dd);
        KMH_HIP(hipGetLastError());
        ++level;
    }
    // down pass: coefficient of each node (pointer + ld); root = identity
    std::vector<const double*> cptr(total_nodes, nullptr);
    std::vector<int> cld(total_nodes, 0);
    std::vector<char> is_ident(total_nodes, 0);
    for (size_t p = 0; p < ps.size(); ++p) is_ident[root[p]] = 1;
    for (int lv = level - 1; lv >= 0; --lv) {
        std::vector<GemmProb> gp;
        for (const auto& cb : combs) {
            if (cb.level != lv) continue;
            const int nb = ps[cb.panel].nb;
            const int qi = cb.node - nnodes;
            const double* Q = QTb + (size_t)qi * QS;
            for (int h = 0; h < 2; ++h) {
                const int child = h ? cb.b : cb.a;
                if (is_ident[cb.node]) {
                    cptr[child] = Q + h * nb;
                    cld[child] = 2 * QC;
                } else {
                    double* dst = CS + (size_t)child * RS;
                    gp.push_back({nb, nb, nb, Q + h * nb, 2 * QC, cptr[cb.node], cld[cb.node], dst, QC,
                                  1.0, 0.0, 0});
                    cptr[child] = dst;
                    cld[child] = QC;
                }
            }
        }
        gemm(c, false, false, gp);
    }
    // tiles: A_tile <- Q_tile * C_tile (in place; n = nb <= 64 keeps each
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

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
template <bool TR>
__global__ __launch_bounds__(256) void k_copy2d(const Copy2D* __restrict__ ds) {
    const Copy2D d = ds[blockIdx.y];
    const long long tot = (long long)d.m * d.n;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int i = (int)(e % d.m), j = (int)(e / d.m);
        const double v = d.alpha * d.src[i + (size_t)j * d.lds];
        if (TR) d.dst[j + (size_t)i * d.ldd] = v;
        else d.dst[i + (size_t)j * d.ldd] = v;
    }
}

void copy2d(Ctx& c, const std::vector<Copy2D>& probs, bool transpose) {
    if (probs.empty()) return;
    for (size_t off = 0; off < probs.size(); off += 65535) {
        std::vector<Copy2D> part(probs.begin() + off, probs.begin() + std::min(probs.size(), off + 65535));
        Copy2D* d = c.upload_vec(part);
        long long mx = 0;
        for (auto& p : part) mx = std::max<long long>(mx, (long long)p.m * p.n);
        const unsigned gx = (unsigned)std::min<long long>(std::max<long long>(1, ceil_div(mx, 256)), 1024);
        if (transpose) hipLaunchKernelGGL((k_copy2d<true>), dim3(gx, (unsigned)part.size()), dim3(256), 0, c.stream, d);
        else hipLaunchKernelGGL((k_copy2d<false>), dim3(gx, (unsigned)part.size()), dim3(256), 0, c.stream, d);
        KMH_HIP(hipGetLastError());
    }
}

__global__ __launch_bounds__(256) void k_axpby2d(const Axpby2D* __restrict__ ds) {
    const Axpby2D d = ds[blockIdx.y];
    const long long tot = (long long)d.m * d.n;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int i = (int)(e % d.m), j = (int)(e / d.m);
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
        hipLaunchKernelGGL(k_axpby2d, dim3(gx, (unsigned)part.size()), dim3(256), 0, c.stream, d);
        KMH_HIP(hipGetLastError());
    }
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
        if (set) hipLaunchKernelGGL((k_diag_op<true>), dim3(gx, (unsigned)part.size()), dim3(256), 0, c.stream, d);
        else hipLaunchKernelGGL((k_diag_op<false>), dim3(gx, (unsigned)part.size()), dim3(256), 0, c.stream, d);
        KMH_HIP(hipGetLastError());
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
    hipLaunchKernelGGL(k_log_diag, dim3((unsigned)ds.size()), dim3(256), 0, c.stream, d);
    KMH_HIP(hipGetLastError());
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
    hipLaunchKernelGGL(k_sum_vec, dim3((unsigned)jobs.size()), dim3(256), 0, c.stream, d);
    KMH_HIP(hipGetLastError());
}

void sum_vector(Ctx& c, const double* d_in, int count, double* d_out) { sum_vectors(c, {{d_in, count, d_out}}); }

__global__ __launch_bounds__(256) void k_coldots(const double* X, int ldx, const double* Y, int ldy, int m,
                                                 double* out) {
    __shared__ double sh[4];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += 256) s += X[i + (size_t)j * ldx] * Y[i + (size_t)j * ldy];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) out[j] = t;
}

void column_dots(Ctx& c, const double* X, int ldx, const double* Y, int ldy, int m, int ncols,
                 double* d_out) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_coldots, dim3(ncols), dim3(256), 0, c.stream, X, ldx, Y, ldy, m, d_out);
    KMH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// panel QR: block Gram-Schmidt (two passes) over 64-column blocks + TSQR
// ---------------------------------------------------------------------------
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
