// Kernel-matrix assembly and partial-pivot ACA on gfx950.
//
//  * k_eval: dense blocks K(rows, cols) of uneven sizes in one launch; the
//    64 column points of a tile are staged in LDS, each thread owns rows.
//  * AcaRun: partial-pivot ACA of every off-diagonal block of a build, run
//    step by step with all blocks' steps grouped into shared launches
//    (src/factorizations.jl:9-84).  Rows and columns are written straight
//    into the level's panel buffer.
#include <algorithm>
#include <cstring>

#include "assemble.h"

namespace kmh {

__global__ __launch_bounds__(256) void k_eval(const EvalDesc* __restrict__ ds, const int* __restrict__ map) {
    __shared__ double cx[64], cy[64];
    const int t = blockIdx.x;
    const EvalDesc d = ds[map[t]];
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
    std::vector<int> map;
    for (auto d : ds) {
        if (d.m <= 0 || d.n <= 0) continue;
        d.tile0 = tiles;
        const int nt = (int)(ceil_div(d.m, 64) * ceil_div(d.n, 64));
        map_append(map, (int)act.size(), nt);
        tiles += nt;
        act.push_back(d);
    }
    if (act.empty()) return;
    EvalDesc* dd = c.upload_vec(act);
    const int* dm = c.upload_vec(map);
    KMH_LAUNCH(c, k_eval, dim3(tiles), dim3(256), dd, dm);
}

__device__ __forceinline__ bool bit_get(const unsigned* m, int i) { return (m[i >> 5] >> (i & 31)) & 1u; }
__device__ __forceinline__ void bit_set(unsigned* m, int i) { m[i >> 5] |= 1u << (i & 31); }

// ---------------------------------------------------------------------------
// ACA (src/factorizations.jl:9-84), step-parallel across workgroups.
//
// Every block of every level runs the reference's loop.  One loop iteration
// of all running blocks is five grouped launches, so a large block's
// residual rows and columns and its norm-update products spread over many
// workgroups:
//   k_aca_row   residual row at the row pivot (:48-50), per 1024-column
//               chunk, with the chunk's argmax of |v| over unused columns
//   k_aca_piv   column pivot (:52-55, restrictedmaxabs over the chunks)
//   k_aca_col   v /= pivot (:56) and its squared norm per column chunk; the
//               residual column at the column pivot (:58-61), its squared
//               norm and the argmax of |u| over unused rows per row chunk
//   k_aca_dots  u.U_t and v.V_t for the earlier terms t (:63-65), one
//               workgroup per term (four waves over quarters of the vectors)
//   k_aca_fin   ||Z||_F^2 update (:63-67), rank + 1, stop test, next row
//               pivot (:43-46)
// The element arithmetic, pivot tie rule and stop test are the reference's;
// sums over chunks are added in chunk order, so results are deterministic.
// ---------------------------------------------------------------------------
constexpr int ACH = 256;    // rows or columns per chunk (one per thread): a
                            // 32768-row block spreads over 128 workgroups
constexpr int ACA_COEF = 4096;   // LDS coefficients: the rank limit

struct AcaArg {
    double a, v;   // |value|, value
    int i, ff;     // index of the maximum, first admissible index
};

__device__ __forceinline__ bool arg_better(double a1, int i1, double a2, int i2) {
    return (a1 > a2) || (a1 == a2 && i1 < i2);
}

__device__ __forceinline__ AcaArg arg_none() { return AcaArg{-1.0, 0.0, 0x7fffffff, 0x7fffffff}; }

struct AcaState {
    const double *rx, *ry, *cx, *cy;
    double *U, *V;
    int nro, ncl, ld, maxsz, cap, rch, cch, pad0;
    int* rank;
    int* status;
    double rtol;
    KDev k;
    unsigned *urow, *ucol;
    AcaArg *rpart, *cpart;   // per row / column chunk argmax
    double *rsum, *csum;     // per column / row chunk squared norms
    double *du, *dv;         // norm-update products of the current step
    // evolving state
    int frnk, done, st, ri, ci, pad1;
    double mv, rn2, cn2, znorm2;
};

// A running block and the offsets of its chunks in the grouped grids.
struct AcaPlan {
    int b, cch0, rch0, pad;
};

// workgroup-wide argmax over (a, i) and min over ff; every thread gets it
__device__ AcaArg block_argmax(AcaArg m, AcaArg* sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        AcaArg o;
        o.a = __shfl_xor(m.a, off);
        o.v = __shfl_xor(m.v, off);
        o.i = __shfl_xor(m.i, off);
        o.ff = __shfl_xor(m.ff, off);
        if (arg_better(o.a, o.i, m.a, m.i)) { m.a = o.a; m.v = o.v; m.i = o.i; }
        m.ff = min(m.ff, o.ff);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = m;
    __syncthreads();
    AcaArg b = sh[0];
    for (int q = 1; q < nw; ++q) {
        if (arg_better(sh[q].a, sh[q].i, b.a, b.i)) { b.a = sh[q].a; b.v = sh[q].v; b.i = sh[q].i; }
        b.ff = min(b.ff, sh[q].ff);
    }
    return b;
}

// workgroup-wide sum in a fixed order; every thread gets it
__device__ double block_sum_fixed(double v, double* sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += sh[q];
    return s;
}

// restrictedmaxabs (src/utils.jl:2-15) from chunk partials: the signed value
// and index of the largest |x| over admissible indices, ties to the lower
// index; (0, first admissible index) if all are zero, where the first
// admissible index is the last index when every index is excluded.
__device__ void pick(const AcaArg* part, int n, int len, double* val, int* idx) {
    double ba = -1.0, bv = 0.0;
    int bi = 0x7fffffff, ff = 0x7fffffff;
    for (int q = 0; q < n; ++q) {
        if (arg_better(part[q].a, part[q].i, ba, bi)) { ba = part[q].a; bv = part[q].v; bi = part[q].i; }
        ff = min(ff, part[q].ff);
    }
    if (ff == 0x7fffffff) ff = len - 1;
    if (ba > 0.0) { *val = bv; *idx = bi; }
    else { *val = 0.0; *idx = ff; }
}

__device__ void aca_stop(AcaState& s, int st) {
    s.done = 1;
    s.st = st;
    *s.status = st;
    *s.rank = s.frnk;
}

// loop head (:43-46): stop test, capacity, next row pivot
__device__ void aca_loop_head(AcaState& s) {
    if (!(sqrt(s.rn2) * sqrt(s.cn2) > s.rtol * sqrt(s.znorm2) && s.frnk < s.maxsz)) {
        aca_stop(s, 0);
        return;
    }
    if (s.frnk >= s.cap) {   // more columns than the panel holds: the host widens it
        aca_stop(s, 2);
        return;
    }
    double v;
    int ri;
    pick(s.rpart, s.rch, s.nro, &v, &ri);
    s.ri = ri;
    bit_set(s.urow, ri);
}

__device__ __forceinline__ int find_plan(const AcaPlan* pl, int np, int x, bool rows) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((rows ? pl[mid].rch0 : pl[mid].cch0) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// First pair (:18-41) and the first loop head, one workgroup per block.
__global__ __launch_bounds__(256) void k_aca_init(AcaState* __restrict__ sts) {
    __shared__ AcaArg sha[4];
    __shared__ double shd[4];
    AcaState& s = sts[blockIdx.x];
    const int tid = threadIdx.x;
    const int wr = (s.nro + 31) / 32, wc = (s.ncl + 31) / 32;
    for (int i = tid; i < wr; i += 256) s.urow[i] = 0u;
    for (int i = tid; i < wc; i += 256) s.ucol[i] = 0u;
    int strt = 0;
    for (;;) {   // :22-25 first non-zero row
        const double xr = s.rx[strt], yr = s.ry[strt];
        double mx = 0.0;
        for (int j = tid; j < s.ncl; j += 256) {
            const double v = kernel_eval(s.k, xr, yr, s.cx[j], s.cy[j]);
            s.V[j] = v;
            mx = fmax(mx, fabs(v));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        __syncthreads();
        if ((tid & 63) == 0) shd[tid >> 6] = mx;
        __syncthreads();
        const double m = fmax(fmax(shd[0], shd[1]), fmax(shd[2], shd[3]));
        __syncthreads();
        if (m == 0.0 && strt + 1 < s.maxsz) {
            ++strt;
            continue;
        }
        break;
    }
    if (strt + 1 == s.maxsz) {   // :26
        if (tid == 0) {
            s.frnk = 0;
            aca_stop(s, 1);
        }
        return;
    }
    __syncthreads();   // the bit masks are cleared
    if (tid == 0) bit_set(s.urow, strt);
    AcaArg m = arg_none();
    for (int j = tid; j < s.ncl; j += 256) {
        if (j < m.ff) m.ff = j;
        const double v = s.V[j], a = fabs(v);
        if (a > m.a) { m.a = a; m.v = v; m.i = j; }
    }
    const AcaArg bc = block_argmax(m, sha);
    double mv;
    int mi;
    pick(&bc, 1, s.ncl, &mv, &mi);   // :33
    double rs = 0.0;
    for (int j = tid; j < s.ncl; j += 256) {
        const double v = s.V[j] / mv;
        s.V[j] = v;
        rs += v * v;
    }
    double cs = 0.0;
    AcaArg mr = arg_none();
    const double xc = s.cx[mi], yc = s.cy[mi];
    for (int i = tid; i < s.nro; i += 256) {
        const double v = kernel_eval(s.k, s.rx[i], s.ry[i], xc, yc);
        s.U[i] = v;
        cs += v * v;
        if (i == strt) continue;
        if (i < mr.ff) mr.ff = i;
        const double a = fabs(v);
        if (a > mr.a) { mr.a = a; mr.v = v; mr.i = i; }
    }
    const double rn2 = block_sum_fixed(rs, shd);
    const double cn2 = block_sum_fixed(cs, shd);
    const AcaArg br = block_argmax(mr, sha);
    if (tid == 0) {
        bit_set(s.ucol, mi);
        s.rn2 = rn2;
        s.cn2 = cn2;
        s.znorm2 = rn2 * cn2;
        s.frnk = 1;
        s.st = 0;
        s.done = 0;
        s.rpart[0] = br;   // the next row pivot comes from U[:, 0]
        for (int q = 1; q < s.rch; ++q) s.rpart[q] = arg_none();
        aca_loop_head(s);
    }
}

// A block stopped at capacity continues with wider panels.
__global__ void k_aca_resume(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl) {
    if (threadIdx.x != 0) return;
    AcaState& s = sts[pl[blockIdx.x].b];
    s.done = 0;
    s.st = 0;
    aca_loop_head(s);
}

__global__ __launch_bounds__(256) void k_aca_row(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl,
                                                 int np) {
    __shared__ double coef[ACA_COEF];
    __shared__ AcaArg sha[4];
    const int p = find_plan(pl, np, blockIdx.x, false);
    AcaState& s = sts[pl[p].b];
    if (s.done) return;
    const int chunk = blockIdx.x - pl[p].cch0;
    const int f = s.frnk, ri = s.ri;
    const size_t ld = (size_t)s.ld;
    for (int t = threadIdx.x; t < f; t += 256) coef[t] = s.U[ri + (size_t)t * ld];
    __syncthreads();
    const double xr = s.rx[ri], yr = s.ry[ri];
    double* vnew = s.V + (size_t)f * ld;
    AcaArg m = arg_none();
    const int j0 = chunk * ACH, j1 = min(s.ncl, j0 + ACH);
    for (int j = j0 + threadIdx.x; j < j1; j += 256) {
        double v = kernel_eval(s.k, xr, yr, s.cx[j], s.cy[j]);
        for (int t = 0; t < f; ++t) v = __dsub_rn(v, __dmul_rn(coef[t], s.V[j + (size_t)t * ld]));
        vnew[j] = v;
        if (bit_get(s.ucol, j)) continue;
        if (j < m.ff) m.ff = j;
        const double a = fabs(v);
        if (a > m.a) { m.a = a; m.v = v; m.i = j; }
    }
    const AcaArg r = block_argmax(m, sha);
    if (threadIdx.x == 0) s.cpart[chunk] = r;
}

__global__ void k_aca_piv(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl) {
    if (threadIdx.x != 0) return;
    AcaState& s = sts[pl[blockIdx.x].b];
    if (s.done) return;
    double mv;
    int ci;
    pick(s.cpart, s.cch, s.ncl, &mv, &ci);
    if (mv == 0.0) {   // :55
        aca_stop(s, 0);
        return;
    }
    s.mv = mv;
    s.ci = ci;
    bit_set(s.ucol, ci);
}

__global__ __launch_bounds__(256) void k_aca_col(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl,
                                                 int np, int ncch) {
    __shared__ double coef[ACA_COEF];
    __shared__ AcaArg sha[4];
    __shared__ double shd[4];
    const bool colpart = (int)blockIdx.x < ncch;
    const int x = colpart ? (int)blockIdx.x : (int)blockIdx.x - ncch;
    const int p = find_plan(pl, np, x, !colpart);
    AcaState& s = sts[pl[p].b];
    if (s.done) return;
    const int f = s.frnk;
    const size_t ld = (size_t)s.ld;
    if (colpart) {
        const int chunk = x - pl[p].cch0;
        double* vnew = s.V + (size_t)f * ld;
        const double mv = s.mv;
        double rs = 0.0;
        const int j0 = chunk * ACH, j1 = min(s.ncl, j0 + ACH);
        for (int j = j0 + threadIdx.x; j < j1; j += 256) {
            const double v = vnew[j] / mv;
            vnew[j] = v;
            rs += v * v;
        }
        const double t = block_sum_fixed(rs, shd);
        if (threadIdx.x == 0) s.rsum[chunk] = t;
        return;
    }
    const int chunk = x - pl[p].rch0;
    const int ci = s.ci;
    for (int t = threadIdx.x; t < f; t += 256) coef[t] = s.V[ci + (size_t)t * ld];
    __syncthreads();
    const double xc = s.cx[ci], yc = s.cy[ci];
    double* unew = s.U + (size_t)f * ld;
    double cs = 0.0;
    AcaArg m = arg_none();
    const int i0 = chunk * ACH, i1 = min(s.nro, i0 + ACH);
    for (int i = i0 + threadIdx.x; i < i1; i += 256) {
        double v = kernel_eval(s.k, s.rx[i], s.ry[i], xc, yc);
        for (int t = 0; t < f; ++t) v = __dsub_rn(v, __dmul_rn(coef[t], s.U[i + (size_t)t * ld]));
        unew[i] = v;
        cs += v * v;
        if (bit_get(s.urow, i)) continue;
        if (i < m.ff) m.ff = i;
        const double a = fabs(v);
        if (a > m.a) { m.a = a; m.v = v; m.i = i; }
    }
    const double t = block_sum_fixed(cs, shd);
    const AcaArg r = block_argmax(m, sha);
    if (threadIdx.x == 0) {
        s.csum[chunk] = t;
        s.rpart[chunk] = r;
    }
}

// workgroup g of the grid: plan entry g / fb, earlier term t = g % fb; wave
// w sums quarter w of the vectors (four interleaved partial sums per lane),
// and the quarters are added in order
__device__ __forceinline__ double quarter_dot(const double* __restrict__ x, const double* __restrict__ y, int n,
                                              int w, int lane) {
    const int q = (n + 3) >> 2, i0 = w * q, i1 = min(n, i0 + q);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = i0 + lane;
    for (; i + 192 < i1; i += 256) {
        a0 += x[i] * y[i];
        a1 += x[i + 64] * y[i + 64];
        a2 += x[i + 128] * y[i + 128];
        a3 += x[i + 192] * y[i + 192];
    }
    for (; i < i1; i += 64) a0 += x[i] * y[i];
    double a = (a0 + a1) + (a2 + a3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    return a;
}

__global__ __launch_bounds__(256) void k_aca_dots(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl,
                                                  int np, int fb) {
    __shared__ double sh[2][4];
    const int g = (int)blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p = g / fb, t = g - p * fb;
    if (p >= np) return;   // uniform over the workgroup
    AcaState& s = sts[pl[p].b];
    if (s.done || t >= s.frnk) return;   // uniform over the workgroup
    const size_t ld = (size_t)s.ld;
    const double a = quarter_dot(s.U + (size_t)t * ld, s.U + (size_t)s.frnk * ld, s.nro, w, lane);
    const double c = quarter_dot(s.V + (size_t)t * ld, s.V + (size_t)s.frnk * ld, s.ncl, w, lane);
    if (lane == 0) {
        sh[0][w] = a;
        sh[1][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s.du[t] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        s.dv[t] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    }
}

__global__ void k_aca_fin(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl) {
    if (threadIdx.x != 0) return;
    AcaState& s = sts[pl[blockIdx.x].b];
    if (s.done) return;
    double rn2 = 0.0, cn2 = 0.0;
    for (int q = 0; q < s.cch; ++q) rn2 += s.rsum[q];
    for (int q = 0; q < s.rch; ++q) cn2 += s.csum[q];
    double z = s.znorm2;
    for (int t = 0; t < s.frnk; ++t) z += 2.0 * fabs(s.du[t]) * fabs(s.dv[t]);
    z += rn2 * cn2;
    s.rn2 = rn2;
    s.cn2 = cn2;
    s.znorm2 = z;
    s.frnk += 1;
    aca_loop_head(s);
}

struct AcaRun::Impl {
    Ctx& c;
    int nb = 0;
    char* mem = nullptr;
    AcaState* d_st = nullptr;
    std::vector<AcaState> h;
    std::vector<int> widen;   // blocks to continue
    explicit Impl(Ctx& cc) : c(cc) {}
};

AcaRun::AcaRun(Ctx& c, const std::vector<AcaDesc>& ds) : m_(new Impl(c)) {
    Impl& m = *m_;
    m.nb = (int)ds.size();
    if (m.nb == 0) return;
    size_t bits = 0, rparts = 0, cparts = 0, ts = 0;
    m.h.resize(m.nb);
    for (int b = 0; b < m.nb; ++b) {
        const auto& d = ds[b];
        KMH_CHECK(d.cap <= ACA_COEF, KMH_E_BAD_ARG,
                  "ACA rank exceeds the supported maximum of 4096");
        AcaState& s = m.h[b];
        std::memset(&s, 0, sizeof(s));
        s.rx = d.px + d.a; s.ry = d.py + d.a;
        s.cx = d.qx + d.c; s.cy = d.qy + d.c;
        s.U = d.U; s.V = d.V;
        s.nro = d.nro; s.ncl = d.ncl; s.ld = d.ld; s.maxsz = d.maxsz; s.cap = d.cap;
        s.rch = (int)ceil_div(d.nro, ACH);
        s.cch = (int)ceil_div(d.ncl, ACH);
        s.rank = d.rank; s.status = d.status;
        s.rtol = d.rtol; s.k = d.k;
        bits += (size_t)(d.nro + 31) / 32 + (size_t)(d.ncl + 31) / 32;
        rparts += s.rch;
        cparts += s.cch;
        ts += std::min(d.maxsz, ACA_COEF);
    }
    auto al = [](size_t n) { return (n + 255) & ~size_t(255); };
    const size_t bytes = al(sizeof(AcaState) * m.nb) + al(bits * 4) + al((rparts + cparts) * sizeof(AcaArg)) +
                         al((rparts + cparts) * sizeof(double)) + al(2 * ts * sizeof(double));
    m.mem = static_cast<char*>(c.scratch(bytes));
    char* w = m.mem;
    auto take = [&](size_t n) { char* p = w; w += al(n); return p; };
    m.d_st = reinterpret_cast<AcaState*>(take(sizeof(AcaState) * m.nb));
    unsigned* bitp = reinterpret_cast<unsigned*>(take(bits * 4));
    AcaArg* rp = reinterpret_cast<AcaArg*>(take((rparts + cparts) * sizeof(AcaArg)));
    AcaArg* cp = rp + rparts;
    double* rsum = reinterpret_cast<double*>(take((rparts + cparts) * sizeof(double)));
    double* csum = rsum + cparts;
    double* du = reinterpret_cast<double*>(take(2 * ts * sizeof(double)));
    double* dv = du + ts;
    for (auto& s : m.h) {
        s.urow = bitp; bitp += (s.nro + 31) / 32;
        s.ucol = bitp; bitp += (s.ncl + 31) / 32;
        s.rpart = rp; rp += s.rch;
        s.cpart = cp; cp += s.cch;
        s.rsum = rsum; rsum += s.cch;
        s.csum = csum; csum += s.rch;
        s.du = du; du += std::min(s.maxsz, ACA_COEF);
        s.dv = dv; dv += std::min(s.maxsz, ACA_COEF);
    }
    KMH_HIP(hipMemcpyAsync(m.d_st, m.h.data(), sizeof(AcaState) * m.nb, hipMemcpyHostToDevice, c.q()));
    KMH_LAUNCH(c, k_aca_init, dim3(m.nb), dim3(256), m.d_st);
}

AcaRun::~AcaRun() {
    if (m_->mem) m_->c.scratch_free(m_->mem);
}

void AcaRun::widen(int b, double* U, double* V, int cap) {
    Impl& m = *m_;
    KMH_CHECK(b >= 0 && b < m.nb && cap <= ACA_COEF, KMH_E_BAD_ARG,
              "ACA rank exceeds the supported maximum of 4096");
    AcaState& s = m.h[b];
    s.U = U;
    s.V = V;
    s.cap = cap;
    m.widen.push_back(b);
}

void AcaRun::run() {
    Impl& m = *m_;
    if (m.nb == 0) return;
    Ctx& c = m.c;
    constexpr int SYNC_STEPS = 16;
    const size_t st_bytes = sizeof(AcaState) * m.nb;
    if (!m.widen.empty()) {
        // the device holds the evolving state; patch the widened panels in
        std::vector<AcaState> cur(m.nb);
        KMH_HIP(hipMemcpyAsync(cur.data(), m.d_st, st_bytes, hipMemcpyDeviceToHost, c.q()));
        c.sync();
        std::vector<AcaPlan> pl;
        for (int b : m.widen) {
            KMH_CHECK(cur[b].st == 2 && cur[b].done, KMH_E_BAD_ARG, "ACA widen of a block that did not stop at capacity");
            cur[b].U = m.h[b].U;
            cur[b].V = m.h[b].V;
            cur[b].cap = m.h[b].cap;
            pl.push_back({b, 0, 0, 0});
        }
        m.widen.clear();
        KMH_HIP(hipMemcpyAsync(m.d_st, cur.data(), st_bytes, hipMemcpyHostToDevice, c.q()));
        c.sync();   // cur is a host temporary
        AcaPlan* d_pl = c.upload_vec(pl);
        KMH_LAUNCH(c, k_aca_resume, dim3((unsigned)pl.size()), dim3(64), m.d_st, d_pl);
    }
    std::vector<AcaState> cur(m.nb);
    for (;;) {
        KMH_HIP(hipMemcpyAsync(cur.data(), m.d_st, st_bytes, hipMemcpyDeviceToHost, c.q()));
        c.sync();
        std::vector<AcaPlan> pl;
        int cch = 0, rch = 0, fb = 1;
        for (int b = 0; b < m.nb; ++b) {
            if (cur[b].done) continue;
            pl.push_back({b, cch, rch, 0});
            cch += cur[b].cch;
            rch += cur[b].rch;
            fb = std::max(fb, cur[b].frnk);
        }
        if (pl.empty()) break;
        const int np = (int)pl.size();
        AcaPlan* d_pl = c.upload_vec(pl);
        for (int q = 0; q < SYNC_STEPS; ++q, ++fb) {
            KMH_LAUNCH(c, k_aca_row, dim3(cch), dim3(256), m.d_st, d_pl, np);
            KMH_LAUNCH(c, k_aca_piv, dim3(np), dim3(64), m.d_st, d_pl);
            KMH_LAUNCH(c, k_aca_col, dim3(cch + rch), dim3(256), m.d_st, d_pl, np, cch);
            KMH_LAUNCH(c, k_aca_dots, dim3((unsigned)((int64_t)np * fb)), dim3(256), m.d_st, d_pl, np, fb);
            KMH_LAUNCH(c, k_aca_fin, dim3(np), dim3(64), m.d_st, d_pl);
        }
    }
}

}  // namespace kmh
