// Kernel-matrix assembly and partial-pivot ACA on gfx950.
//
//  * k_eval: dense blocks K(rows, cols) of uneven sizes in one launch; the
//    64 column points of a tile are staged in LDS, each thread owns rows.
//  * k_aca:  one workgroup (1024 threads) runs the full ACA of one
//    off-diagonal block, following src/factorizations.jl:9-84 step by step:
//    pivot search with the first-index tie rule of restrictedmaxabs
//    (src/utils.jl:2-15), sequential residual subtraction, and the norm
//    update accumulated in the reference's order.  Rows and columns are
//    written straight into the level's panel buffer.
#include <algorithm>

#include "assemble.h"

namespace kmh {

__global__ __launch_bounds__(256) void k_eval(const EvalDesc* __restrict__ ds, int nd) {
    __shared__ double cx[64], cy[64];
    int lo = 0, hi = nd - 1;
    const int t = blockIdx.x;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ds[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const EvalDesc d = ds[lo];
    const KDev k = d.k;
    const int lt = t - d.tile0;
    const int tn = (d.n + 63) / 64;
    const int row0 = (lt / tn) * 64, col0 = (lt % tn) * 64;
    const int tid = threadIdx.x;
    if (tid < 64) {
        const int j = col0 + tid;
        cx[tid] = (j < d.n) ? d.cx[j] : 0.0;
        cy[tid] = (j < d.n) ? d.cy[j] : 0.0;
    }
    __syncthreads();
    const int i = row0 + (tid & 63);
    if (i >= d.m) return;
    const double xi = d.rx[i], yi = d.ry[i];
    const int jmax = min(64, d.n - col0);
    for (int jj = tid >> 6; jj < jmax; jj += 4) {
        const double v = kernel_eval(k, xi, yi, cx[jj], cy[jj]);
        d.dst[i + (size_t)(col0 + jj) * d.ld] = v;
    }
}

void eval_blocks(Ctx& c, const kmh_kernel& kern, std::vector<EvalDesc> ds) {
    const KDev k = to_kdev(kern);
    for (auto& d : ds) d.k = k;
    eval_blocks(c, std::move(ds));
}

void eval_blocks(Ctx& c, std::vector<EvalDesc> ds) {
    int tiles = 0;
    std::vector<EvalDesc> act;
    for (auto d : ds) {
        if (d.m <= 0 || d.n <= 0) continue;
        d.tile0 = tiles;
        tiles += (int)(ceil_div(d.m, 64) * ceil_div(d.n, 64));
        act.push_back(d);
    }
    if (act.empty()) return;
    EvalDesc* dd = c.upload_vec(act);
    KMH_LAUNCH(c, k_eval, dim3(tiles), dim3(256), dd, (int)act.size());
}

// ---------------------------------------------------------------------------
// ACA
// ---------------------------------------------------------------------------
constexpr int AT = 1024;          // threads per ACA workgroup
constexpr int AW = AT / 64;       // waves

struct ArgMax {
    double a;   // |value|
    double v;   // signed value
    int i;      // index
};

__device__ __forceinline__ bool better(const ArgMax& x, const ArgMax& y) {
    // larger |value| wins; ties go to the lower index
    return (x.a > y.a) || (x.a == y.a && x.i < y.i);
}

__device__ __forceinline__ ArgMax wave_argmax(ArgMax m) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ArgMax o;
        o.a = __shfl_xor(m.a, off);
        o.v = __shfl_xor(m.v, off);
        o.i = __shfl_xor(m.i, off);
        if (better(o, m)) m = o;
    }
    return m;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}

struct AcaShared {
    double sd[AW];
    double sa[AW], sv[AW];
    int si[AW], sm[AW];
};

__device__ double block_sum_aca(double v, AcaShared& sh) {
    v = wave_sum_d(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh.sd[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < AW; ++q) s += sh.sd[q];
    return s;
}

__device__ __forceinline__ bool bit_get(const unsigned* m, int i) { return (m[i >> 5] >> (i & 31)) & 1u; }

// restrictedmaxabs over x[0..len) (global, ld-contiguous) excluding mask.
// Returns (signed value, index); (0, first free index) if all are zero.
__device__ void restricted_maxabs(const double* x, int len, const unsigned* mask, AcaShared& sh,
                                  double* val, int* idx) {
    ArgMax m{-1.0, 0.0, 0x7fffffff};
    int ffree = 0x7fffffff;
    for (int j = threadIdx.x; j < len; j += AT) {
        if (bit_get(mask, j)) continue;
        if (j < ffree) ffree = j;
        const double v = x[j];
        const double a = fabs(v);
        if (a > m.a) { m.a = a; m.v = v; m.i = j; }
    }
    m = wave_argmax(m);
    ffree = wave_min_i(ffree);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { sh.sa[w] = m.a; sh.sv[w] = m.v; sh.si[w] = m.i; sh.sm[w] = ffree; }
    __syncthreads();
    ArgMax b{sh.sa[0], sh.sv[0], sh.si[0]};
    int ff = sh.sm[0];
    for (int q = 1; q < AW; ++q) {
        ArgMax o{sh.sa[q], sh.sv[q], sh.si[q]};
        if (better(o, b)) b = o;
        ff = min(ff, sh.sm[q]);
    }
    if (ff == 0x7fffffff) ff = len - 1;   // every index excluded
    if (b.a > 0.0) { *val = b.v; *idx = b.i; }
    else { *val = 0.0; *idx = ff; }
}

__global__ __launch_bounds__(AT) void k_aca(const AcaDesc* __restrict__ ds) {
    extern __shared__ char smem[];
    __shared__ AcaShared sh;
    __shared__ double s_z;
    const AcaDesc d = ds[blockIdx.x];
    const KDev k = d.k;
    const double rtol = d.rtol;
    const double* px = d.px;
    const double* py = d.py;
    const int tid = threadIdx.x;
    const int wr = (d.nro + 31) / 32, wc = (d.ncl + 31) / 32;
    unsigned* urow = reinterpret_cast<unsigned*>(smem);
    unsigned* ucol = urow + wr;
    double* coef = reinterpret_cast<double*>(smem + (((size_t)(wr + wc) * 4 + 15) & ~size_t(15)));
    double* du = coef + d.cap;
    double* dv = du + d.cap;
    for (int i = tid; i < wr + wc; i += AT) urow[i] = 0u;
    const double* rx = px + d.a;
    const double* ry = py + d.a;
    const double* cxp = px + d.c;
    const double* cyp = py + d.c;
    double* U = d.U;
    double* V = d.V;
    const size_t ld = (size_t)d.ld;
    __syncthreads();

    // first row: skip all-zero rows (src/factorizations.jl:22-27)
    int strt = 0;
    for (;;) {
        const double xr = rx[strt], yr = ry[strt];
        double mx = 0.0;
        for (int j = tid; j < d.ncl; j += AT) {
            const double v = kernel_eval(k, xr, yr, cxp[j], cyp[j]);
            V[j] = v;
            mx = fmax(mx, fabs(v));
        }
        // block max
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        __syncthreads();
        if ((tid & 63) == 0) sh.sd[tid >> 6] = mx;
        __syncthreads();
        double m = 0.0;
        for (int q = 0; q < AW; ++q) m = fmax(m, sh.sd[q]);
        if (m == 0.0 && strt + 1 < d.maxsz) { ++strt; continue; }
        break;
    }
    if (strt + 1 == d.maxsz) {
        if (tid == 0) { *d.status = 1; *d.rank = 0; }
        return;
    }
    if (tid == 0) urow[strt >> 5] |= 1u << (strt & 31);
    __syncthreads();
    double mv;
    int mi;
    restricted_maxabs(V, d.ncl, ucol, sh, &mv, &mi);
    double rs = 0.0;
    for (int j = tid; j < d.ncl; j += AT) {
        const double v = V[j] / mv;
        V[j] = v;
        rs += v * v;
    }
    double cs = 0.0;
    {
        const double xc = cxp[mi], yc = cyp[mi];
        for (int i = tid; i < d.nro; i += AT) {
            const double v = kernel_eval(k, rx[i], ry[i], xc, yc);
            U[i] = v;
            cs += v * v;
        }
    }
    double rn2 = block_sum_aca(rs, sh);
    double cn2 = block_sum_aca(cs, sh);
    double znorm2 = rn2 * cn2;
    if (tid == 0) ucol[mi >> 5] |= 1u << (mi & 31);
    int frnk = 1;
    int status = 0;
    __syncthreads();

    while (sqrt(rn2) * sqrt(cn2) > rtol * sqrt(znorm2) && frnk < d.maxsz) {
        if (frnk >= d.cap) { status = 2; break; }
        const double* ucur = U + (size_t)(frnk - 1) * ld;
        double* vnew = V + (size_t)frnk * ld;
        double* unew = U + (size_t)frnk * ld;
        // next row index (src/factorizations.jl:45-46)
        int ri;
        restricted_maxabs(ucur, d.nro, urow, sh, &mv, &ri);
        if (tid == 0) urow[ri >> 5] |= 1u << (ri & 31);
        for (int t = tid; t < frnk; t += AT) coef[t] = U[ri + (size_t)t * ld];
        __syncthreads();
        // residual row (src/factorizations.jl:48-50)
        {
            const double xr = rx[ri], yr = ry[ri];
            for (int j = tid; j < d.ncl; j += AT) {
                double v = kernel_eval(k, xr, yr, cxp[j], cyp[j]);
                for (int t = 0; t < frnk; ++t) v = __dsub_rn(v, __dmul_rn(coef[t], V[j + (size_t)t * ld]));
                vnew[j] = v;
            }
        }
        __syncthreads();
        int ci;
        restricted_maxabs(vnew, d.ncl, ucol, sh, &mv, &ci);
        if (mv == 0.0) break;                         // :55
        if (tid == 0) ucol[ci >> 5] |= 1u << (ci & 31);
        rs = 0.0;
        for (int j = tid; j < d.ncl; j += AT) {
            const double v = vnew[j] / mv;
            vnew[j] = v;
            rs += v * v;
        }
        for (int t = tid; t < frnk; t += AT) coef[t] = V[ci + (size_t)t * ld];
        __syncthreads();
        // residual column (src/factorizations.jl:58-61)
        cs = 0.0;
        {
            const double xc = cxp[ci], yc = cyp[ci];
            for (int i = tid; i < d.nro; i += AT) {
                double v = kernel_eval(k, rx[i], ry[i], xc, yc);
                for (int t = 0; t < frnk; ++t) v = __dsub_rn(v, __dmul_rn(coef[t], U[i + (size_t)t * ld]));
                unew[i] = v;
                cs += v * v;
            }
        }
        rn2 = block_sum_aca(rs, sh);
        cn2 = block_sum_aca(cs, sh);   // includes a barrier: unew/vnew complete
        // norm update terms (src/factorizations.jl:63-67): one wave per term
        {
            const int lane = tid & 63, w = tid >> 6;
            for (int t = w; t < frnk; t += AW) {
                const double* ut = U + (size_t)t * ld;
                const double* vt = V + (size_t)t * ld;
                double a = 0.0, b = 0.0;
                for (int i = lane; i < d.nro; i += 64) a += ut[i] * unew[i];
                for (int j = lane; j < d.ncl; j += 64) b += vt[j] * vnew[j];
                a = wave_sum_d(a);
                b = wave_sum_d(b);
                if (lane == 0) { du[t] = a; dv[t] = b; }
            }
        }
        __syncthreads();
        if (tid == 0) {
            double z = znorm2;
            for (int t = 0; t < frnk; ++t) z += 2.0 * fabs(du[t]) * fabs(dv[t]);
            z += rn2 * cn2;
            s_z = z;
        }
        __syncthreads();
        znorm2 = s_z;
        ++frnk;
    }
    if (tid == 0) { *d.status = status; *d.rank = frnk; }
}

size_t aca_shared_bytes(int nro, int ncl, int cap) {
    const size_t words = (size_t)(nro + 31) / 32 + (size_t)(ncl + 31) / 32;
    return ((words * 4 + 15) & ~size_t(15)) + (size_t)3 * cap * sizeof(double);
}

void aca_launch(Ctx& c, const std::vector<AcaDesc>& ds) {
    if (ds.empty()) return;
    size_t shm = 0;
    for (const auto& d : ds) shm = std::max(shm, aca_shared_bytes(d.nro, d.ncl, d.cap));
    KMH_CHECK(shm <= 120 * 1024, KMH_E_BAD_ARG, "ACA block too large for one workgroup's LDS");
    AcaDesc* dd = c.upload_vec(ds);
    KMH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_aca),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    c.launch_kernel(k_aca, dim3((unsigned)ds.size()), dim3(AT), (unsigned)shm, dd);
}

}  // namespace kmh
