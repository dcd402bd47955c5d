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
#include "wavesum.h"

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
// residual rows and columns spread over many workgroups:
//   k_aca_row   residual row at the row pivot (:48-50), per 256-column
//               chunk, with the chunk's argmax of |v| over unused columns
//   k_aca_piv   column pivot (:52-55, restrictedmaxabs over the chunks)
//   k_aca_col   v /= pivot (:56) and its squared norm per column chunk; the
//               residual column at the column pivot (:58-61), its squared
//               norm and the argmax of |u| over unused rows per row chunk
//   k_aca_dred  the norm-update products u.U_t and v.V_t (:63-65) of the
//               previous term, summed from the per-wave partials that the
//               row and column passes left
//   k_aca_fin   ||Z||_F^2 update (:63-67), stop test, rank + 1, next row
//               pivot (:43-46)
// The products of a term against the earlier ones are formed one step
// late: the next step's residual passes stream the earlier terms' columns
// anyway (the reference's loop reads U and V once for the residual and
// once more for the products), so they multiply each element by the
// previous term's element on the way.  That makes the next step
// speculative: it runs before the stop test that would have allowed it,
// and when the test (evaluated with the now complete ||Z||_F^2) stops the
// loop, its row and column stay unused beyond the block's rank.  The
// element arithmetic, pivot tie rule and stop test are the reference's; sums
// over chunks and waves are added in a fixed order, so results are
// deterministic.
// ---------------------------------------------------------------------------
constexpr int ACH = 256;    // rows or columns per chunk (one per thread): a
                            // 32768-row block spreads over 128 workgroups
constexpr int ACA_COEF = 4096;   // LDS coefficients: the rank limit
constexpr int ACA_B = 16;        // terms per batch of a residual pass

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
    double *du, *dv;         // norm-update products of the pending term
    double *pdu, *pdv;       // their per-wave partials: [t][wave slot], rows of plu / plv
    int plu, plv;            // 4 x row chunks, 4 x column chunks
    // evolving state: frnk accepted terms; the last one's norms rn2, cn2 and
    // products are pending, znorm2 holds ||Z||^2 before its contribution
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

// pick() by the 64 lanes of one wave, every lane getting the result: the
// largest entry under arg_better's total order (|value| descending, then
// index ascending) and the least ff do not depend on the scan order, so this
// equals the sequential scan.
__device__ void pick_wave(const AcaArg* part, int n, int len, double* val, int* idx) {
    const int lane = threadIdx.x & 63;
    double ba = -1.0, bv = 0.0;
    int bi = 0x7fffffff, ff = 0x7fffffff;
    for (int q = lane; q < n; q += 64) {
        const AcaArg p = part[q];
        if (arg_better(p.a, p.i, ba, bi)) { ba = p.a; bv = p.v; bi = p.i; }
        ff = min(ff, p.ff);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double oa = __shfl_xor(ba, off), ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off), of = __shfl_xor(ff, off);
        if (arg_better(oa, oi, ba, bi)) { ba = oa; bv = ov; bi = oi; }
        ff = min(ff, of);
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


// before a (speculative) step: the loop's rank limit (:43) and the panel's
// capacity for the step's column; false when the block stopped
__device__ bool aca_can_step(AcaState& s) {
    if (s.frnk >= s.maxsz) {
        aca_stop(s, 0);
        return false;
    }
    if (s.frnk >= s.cap) {   // more columns than the panel holds: the host widens it
        aca_stop(s, 2);
        return false;
    }
    return true;
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
        s.znorm2 = 0.0;   // the first term's rn2 * cn2 is added by the first k_aca_fin
        s.frnk = 1;
        s.st = 0;
        s.done = 0;
        if (aca_can_step(s)) {   // the next row pivot comes from U[:, 0]
            double v;
            int ri;
            pick(&br, 1, s.nro, &v, &ri);
            s.ri = ri;
            bit_set(s.urow, ri);
        }
    }
}

// One lane's residual chain and its share of the previous term's products:
//   returns v - c_0 x_0 - c_1 x_1 - ... - c_{f-1} x_{f-1}, x_t = p[t ld],
//   each product rounded and subtracted in term order (src/factorizations.jl
//   :49-50, 59-60: tmprow .-= U[j][mi] * V[j] for j in 1:frnk);
//   and for t < f - 1 stores the wave's sum of x_t x_{f-1} (the products of
//   the last term against the earlier ones, :63-65) at part[t * pl].
// All lanes of the wave take part (an inactive lane, act false, contributes
// zero products; its p must still point into the panel).  coef[f .. f +
// ACA_B) must be zero.  Loads go ACA_B
// terms at a time, the next batch issued before the current one is used: a
// lane's chain runs over up to thousands of column-strided loads, and a
// block of 32768 columns has only 512 waves, so the loads in flight per
// wave set the bandwidth.  The batch's products are summed over the wave by
// column sums in LDS (pw: 64 x 17 doubles of this wave), fixed order.
__device__ __forceinline__ double aca_pass(double v, const double* __restrict__ coef, const double* __restrict__ p,
                                           size_t ld, int f, bool act, double* pw, int lane, double* part,
                                           int pl) {
    typedef const __attribute__((address_space(1))) double* GP;
    const GP g = (GP)p;
    const int k = f - 1;   // the previous term; products against t < k
    const double last = (act && k > 0) ? g[(size_t)k * ld] : 0.0;
    double xa[ACA_B], xb[ACA_B];
    // terms t >= f load a valid element masked to +0 (a bit mask, not a
    // select, which would become a branch around the load); with coef[t] = 0
    // there, v - (+0 * 0) = v + (-0) = v exactly, so the chain needs no
    // condition
    // (the mask is applied where the value is used, so the next batch's
    // loads stay in flight while the current one is consumed)
    auto load = [&](double (&x)[ACA_B], int t0) {
#pragma unroll
        for (int u = 0; u < ACA_B; ++u) x[u] = g[(size_t)min(t0 + u, f - 1) * ld];
    };
    auto use = [&](double (&x)[ACA_B], int t0) {
#pragma unroll
        for (int u = 0; u < ACA_B; ++u) {
            const unsigned long long mk = (t0 + u < f) ? ~0ull : 0ull;
            x[u] = __longlong_as_double(__double_as_longlong(x[u]) & mk);
            v = __dsub_rn(v, __dmul_rn(coef[t0 + u], x[u]));
        }
        if (t0 < k) {   // uniform
            double pr[ACA_B];
#pragma unroll
            for (int u = 0; u < ACA_B; ++u) pr[u] = (t0 + u < k) ? x[u] * last : 0.0;
            const double sm = lds_colsum<ACA_B>(pr, pw, lane);
            if (lane < ACA_B && t0 + lane < k) part[(size_t)(t0 + lane) * pl] = sm;
        }
    };
    if (f <= 0) return v;
    // the next batch is loaded unconditionally (addresses clamped, values
    // masked where used): a conditional load would make the wait before the
    // current batch's first use a wait for every load in flight
    load(xa, 0);
    for (int t0 = 0; t0 < f; t0 += ACA_B) {
        load(xb, t0 + ACA_B);
        use(xa, t0);
#pragma unroll
        for (int u = 0; u < ACA_B; ++u) xa[u] = xb[u];
    }
    return v;
}

// A block stopped at capacity continues with wider panels: pick its next
// row pivot (one wave per block).
__global__ __launch_bounds__(64) void k_aca_resume(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl) {
    AcaState& s = sts[pl[blockIdx.x].b];
    const int ok = __builtin_amdgcn_readfirstlane(s.frnk < s.cap && s.frnk < s.maxsz ? 1 : 0);
    if (!ok) return;
    double v;
    int ri;
    pick_wave(s.rpart, s.rch, s.nro, &v, &ri);
    if (threadIdx.x == 0) {
        s.done = 0;
        s.st = 0;
        s.ri = ri;
        bit_set(s.urow, ri);
    }
}

__global__ __launch_bounds__(256) void k_aca_row(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl,
                                                 int np) {
    __shared__ double coef[ACA_COEF + ACA_B];
    __shared__ double pws[4][64 * (ACA_B + 1)];
    __shared__ AcaArg sha[4];
    const int p = find_plan(pl, np, blockIdx.x, false);
    AcaState& s = sts[pl[p].b];
    if (s.done) return;
    const int chunk = blockIdx.x - pl[p].cch0;
    const int f = s.frnk, ri = s.ri;
    const size_t ld = (size_t)s.ld;
    for (int t = threadIdx.x; t < f + ACA_B; t += 256) coef[t] = t < f ? s.U[ri + (size_t)t * ld] : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double xr = s.rx[ri], yr = s.ry[ri];
    double* vnew = s.V + (size_t)f * ld;
    const int j0 = chunk * ACH, j1 = min(s.ncl, j0 + ACH);
    const int j = j0 + (int)threadIdx.x;
    const bool act = j < j1;
    const int jc = act ? j : j0;
    const double v = aca_pass(act ? kernel_eval(s.k, xr, yr, s.cx[jc], s.cy[jc]) : 0.0, coef, s.V + jc, ld, f,
                              act, pws[w], lane, s.pdv + chunk * 4 + w, s.plv);
    AcaArg m = arg_none();
    if (act) {
        vnew[j] = v;
        if (!bit_get(s.ucol, j)) {
            m.ff = j;
            m.a = fabs(v);
            m.v = v;
            m.i = j;
        }
    }
    const AcaArg r = block_argmax(m, sha);
    if (threadIdx.x == 0) s.cpart[chunk] = r;
}

// one wave per block
__global__ __launch_bounds__(64) void k_aca_piv(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl) {
    AcaState& s = sts[pl[blockIdx.x].b];
    if (s.done) return;   // uniform
    double mv;
    int ci;
    pick_wave(s.cpart, s.cch, s.ncl, &mv, &ci);
    if (threadIdx.x != 0) return;
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
    __shared__ double coef[ACA_COEF + ACA_B];
    __shared__ double pws[4][64 * (ACA_B + 1)];
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
    for (int t = threadIdx.x; t < f + ACA_B; t += 256) coef[t] = t < f ? s.V[ci + (size_t)t * ld] : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double xc = s.cx[ci], yc = s.cy[ci];
    double* unew = s.U + (size_t)f * ld;
    const int i0 = chunk * ACH, i1 = min(s.nro, i0 + ACH);
    const int i = i0 + (int)threadIdx.x;
    const bool act = i < i1;
    const int ic = act ? i : i0;
    const double v = aca_pass(act ? kernel_eval(s.k, s.rx[ic], s.ry[ic], xc, yc) : 0.0, coef, s.U + ic, ld, f,
                              act, pws[w], lane, s.pdu + chunk * 4 + w, s.plu);
    double cs = 0.0;
    AcaArg m = arg_none();
    if (act) {
        unew[i] = v;
        cs = v * v;
        if (!bit_get(s.urow, i)) {
            m.ff = i;
            m.a = fabs(v);
            m.v = v;
            m.i = i;
        }
    }
    const double t = block_sum_fixed(cs, shd);
    const AcaArg r = block_argmax(m, sha);
    if (threadIdx.x == 0) {
        s.csum[chunk] = t;
        s.rpart[chunk] = r;
    }
}

// The previous term's products u_k.U_t and v_k.V_t for t < k (k = frnk - 1)
// from the per-wave partials of the residual passes: wave w of workgroup g
// takes term t = 4 (g % fb4) + w of plan entry g / fb4; its lanes add the
// slots lane, lane + 64, ... in order and the lanes meet by xor-shuffles.
__global__ __launch_bounds__(256) void k_aca_dred(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl,
                                                  int np, int fb4) {
    const int g = (int)blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p = g / fb4;
    if (p >= np) return;
    AcaState& s = sts[pl[p].b];
    const int t = 4 * (g - p * fb4) + w;
    if (s.done || t >= s.frnk - 1) return;   // uniform over the wave
    const double* ru = s.pdu + (size_t)t * s.plu;
    const double* rv = s.pdv + (size_t)t * s.plv;
    double a = 0.0, c = 0.0;
    for (int q = lane; q < s.plu; q += 64) a += ru[q];
    for (int q = lane; q < s.plv; q += 64) c += rv[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        c += __shfl_xor(c, off);
    }
    if (lane == 0) {
        s.du[t] = a;
        s.dv[t] = c;
    }
}

// One 256-thread workgroup per block: ||Z||^2 with the previous term k
// (:63-67), the stop test of the step just taken (:43), and if it stands,
// rank + 1 and the next row pivot.  The chunk norms and the product terms
// are staged in LDS by all threads and summed by one in the sequential
// order (z starts from the stored ||Z||^2 and takes the terms t = 0, 1, ...
// as :63-66 do); the next row pivot is one wave's pick.
constexpr int FIN_STAGE = 1024;
__global__ __launch_bounds__(256) void k_aca_fin(AcaState* __restrict__ sts, const AcaPlan* __restrict__ pl) {
    __shared__ double sa[FIN_STAGE], sb[FIN_STAGE];
    __shared__ int go;
    AcaState& s = sts[pl[blockIdx.x].b];
    if (s.done) return;   // uniform
    const int tid = threadIdx.x;
    const int k = s.frnk - 1;
    double z = s.znorm2;
    for (int b0 = 0; b0 < k; b0 += FIN_STAGE) {
        const int m = min(FIN_STAGE, k - b0);
        for (int q = tid; q < m; q += 256) {
            sa[q] = s.du[b0 + q];
            sb[q] = s.dv[b0 + q];
        }
        __syncthreads();
        if (tid == 0)
            for (int q = 0; q < m; ++q) z += 2.0 * fabs(sa[q]) * fabs(sb[q]);
        __syncthreads();
    }
    if (tid == 0) {
        z += s.rn2 * s.cn2;
        // :43, the test the speculative step should have passed
        go = (sqrt(s.rn2) * sqrt(s.cn2) > s.rtol * sqrt(z)) ? 1 : 0;
        if (!go) aca_stop(s, 0);
        else s.znorm2 = z;
    }
    __syncthreads();
    if (!go) return;
    double rn2 = 0.0, cn2 = 0.0;
    for (int b0 = 0; b0 < s.cch; b0 += FIN_STAGE) {
        const int m = min(FIN_STAGE, s.cch - b0);
        for (int q = tid; q < m; q += 256) sa[q] = s.rsum[b0 + q];
        __syncthreads();
        if (tid == 0)
            for (int q = 0; q < m; ++q) rn2 += sa[q];
        __syncthreads();
    }
    for (int b0 = 0; b0 < s.rch; b0 += FIN_STAGE) {
        const int m = min(FIN_STAGE, s.rch - b0);
        for (int q = tid; q < m; q += 256) sa[q] = s.csum[b0 + q];
        __syncthreads();
        if (tid == 0)
            for (int q = 0; q < m; ++q) cn2 += sa[q];
        __syncthreads();
    }
    if (tid == 0) {
        s.rn2 = rn2;
        s.cn2 = cn2;
        s.frnk = k + 2;
        go = aca_can_step(s) ? 1 : 0;
    }
    __syncthreads();
    if (!go || tid >= 64) return;
    double v;
    int ri;
    pick_wave(s.rpart, s.rch, s.nro, &v, &ri);
    if (tid == 0) {
        s.ri = ri;
        bit_set(s.urow, ri);
    }
}

struct AcaRun::Impl {
    Ctx& c;
    int nb = 0;
    char* mem = nullptr;
    AcaState* d_st = nullptr;
    std::vector<AcaState> h;
    std::vector<int> widen;   // blocks to continue
    std::vector<char*> extra;  // product partials of widened blocks
    explicit Impl(Ctx& cc) : c(cc) {}
};

AcaRun::AcaRun(Ctx& c, const std::vector<AcaDesc>& ds) : m_(new Impl(c)) {
    Impl& m = *m_;
    m.nb = (int)ds.size();
    if (m.nb == 0) return;
    size_t bits = 0, rparts = 0, cparts = 0, ts = 0, pp = 0;
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
        s.plu = 4 * s.rch;
        s.plv = 4 * s.cch;
        s.rank = d.rank; s.status = d.status;
        s.rtol = d.rtol; s.k = d.k;
        bits += (size_t)(d.nro + 31) / 32 + (size_t)(d.ncl + 31) / 32;
        rparts += s.rch;
        cparts += s.cch;
        ts += std::min(d.maxsz, ACA_COEF);
        pp += (size_t)d.cap * (s.plu + s.plv);   // products of terms t < frnk - 1 < cap
    }
    auto al = [](size_t n) { return (n + 255) & ~size_t(255); };
    const size_t bytes = al(sizeof(AcaState) * m.nb) + al(bits * 4) + al((rparts + cparts) * sizeof(AcaArg)) +
                         al((rparts + cparts) * sizeof(double)) + al(2 * ts * sizeof(double)) +
                         al(pp * sizeof(double));
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
    double* pd = reinterpret_cast<double*>(take(pp * sizeof(double)));
    for (auto& s : m.h) {
        s.pdu = pd; pd += (size_t)s.cap * s.plu;
        s.pdv = pd; pd += (size_t)s.cap * s.plv;
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
    for (char* p : m_->extra) m_->c.scratch_free(p);
    if (m_->mem) m_->c.scratch_free(m_->mem);
}

void AcaRun::widen(int b, double* U, double* V, int cap) {
    Impl& m = *m_;
    KMH_CHECK(b >= 0 && b < m.nb && cap <= ACA_COEF, KMH_E_BAD_ARG,
              "ACA rank exceeds the supported maximum of 4096");
    AcaState& s = m.h[b];
    if (cap > s.cap) {   // the product partials grow with the capacity
        char* pd = static_cast<char*>(m.c.scratch((size_t)cap * (s.plu + s.plv) * sizeof(double)));
        m.extra.push_back(pd);
        s.pdu = reinterpret_cast<double*>(pd);
        s.pdv = s.pdu + (size_t)cap * s.plu;
    }
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
            cur[b].pdu = m.h[b].pdu;
            cur[b].pdv = m.h[b].pdv;
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
            const int fb4 = (fb + 3) / 4;
            KMH_LAUNCH(c, k_aca_dred, dim3((unsigned)((int64_t)np * fb4)), dim3(256), m.d_st, d_pl, np, fb4);
            KMH_LAUNCH(c, k_aca_fin, dim3(np), dim3(256), m.d_st, d_pl);
        }
    }
}

}  // namespace kmh
