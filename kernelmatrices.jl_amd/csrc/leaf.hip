// Fused leaf stage for leaves of at most 256 rows (symfact.jl:17-21).
//
//  * k_leaf_chol: one 512-thread workgroup per leaf.  The leaf's lower
//    16 x 16 tiles (136 for 256 rows) are evaluated from the points straight
//    into registers, dealt round-robin to the 8 waves (up to 17 per wave),
//    and factored by a right-looking blocked Cholesky: per block column p
//    the owner of the diagonal tile factors it in registers (dpotf2 order,
//    values of other rows by lane broadcasts) and inverts it, the owners of
//    the tiles below it form L_ip = S_ip L_pp^-T on MFMA and publish them in
//    LDS, and every wave updates its trailing tiles S_ij -= L_ip L_jp^T on
//    MFMA.  The leaf of K is never written: the kernel writes L once, with
//    the inverses of its diagonal tiles and the log-diagonal sum.
//  * k_leaf_solve: B <- L^-1 B for the leaf's rows of the panels, one
//    workgroup per leaf and 80-column chunk: wave w keeps row tile w of B
//    in registers; per block column p, wave p forms X_p = L_pp^-1 B_p with
//    the stored inverse and publishes it, and every wave below it subtracts
//    L_wp X_p, its tile of L loaded one step ahead.
//
// Tile layout: element (r, c) of a 16 x 16 tile is held by lane r + 16 (c & 3)
// in register c >> 2, stored at (c >> 2) * 64 + r + 16 (c & 3).  With
// v_mfma_f64_16x16x4f64 (A(i,k) in lane i + 16k, B(k,j) in lane j + 16k,
// accumulator r of lane j + 16g holding D(4r + g, j)), register q of a tile
// in this layout is the A operand of the product's k-chunk q as it stands,
// is the B operand of chunk q of a product with the tile's transpose, and the
// accumulator of D = Z^T is Z in this layout.  Right-hand tiles use the
// natural layout (element (r, c) in lane c + 16 (r & 3), register r >> 2),
// the accumulator layout of D = Z.
#include "leaf.h"

namespace kmh {

namespace {
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int LF_NT = LEAF_FUSED_MAX / 16;               // tiles per dimension
constexpr int LF_TILES = LF_NT * (LF_NT - 1) / 2;         // tiles below the diagonal
constexpr int LF_WAVES = 8;
constexpr int LF_SLOTS = (LF_TILES + LF_WAVES - 1) / LF_WAVES;
constexpr int LF_RS = LF_SLOTS - 2;   // slots held in registers; the last two in LDS
constexpr int LS_CW = 80;                                 // columns per solve workgroup
constexpr int LS_CB = LS_CW / 16;

__device__ __forceinline__ double bcast(double v, int lane) {
    // value of v held by `lane` (a compile-time constant after unrolling)
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// tile id (row-major over the tiles below the diagonal) -> block row i > column j
__device__ __forceinline__ void tile_ij(int id, int& i, int& j) {
    int r = 1;
    while (r * (r + 1) / 2 <= id) ++r;
    i = r;
    j = id - r * (r - 1) / 2;
}

// The owned tiles of K (identity beyond m) for one covariance family: with
// the family and the derivative code fixed, kernel_eval folds to one
// formula, so the fully unrolled loop stays small.
template <int FAM>
__device__ __forceinline__ void eval_tiles(const KDev& k0, d4 (&R)[LF_SLOTS], const int (&ti)[LF_SLOTS],
                                           const int (&tj)[LF_SLOTS], const double (*pts)[LEAF_FUSED_MAX], int m,
                                           int li, int lg, int wave, int nt, double (*dtile)[256], double* lt) {
    KDev k = k0;
    k.family = FAM;
    k.deriv = KMH_D0;
    const int lane = li + 16 * lg;
    for (int p = wave; p < nt; p += LF_WAVES) {   // the wave's diagonal tiles, to LDS
        const int row = 16 * p + li;
        const double xr = pts[0][row], yr = pts[1][row];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = 16 * p + 4 * q + lg;
            dtile[p][q * 64 + lane] = (row < m && col < m) ? kernel_eval(k, xr, yr, pts[0][col], pts[1][col])
                                                           : (row == col ? 1.0 : 0.0);
        }
    }
#pragma unroll
    for (int t = 0; t < LF_SLOTS; ++t) {
        if (ti[t] < 0) continue;
        const int row = 16 * ti[t] + li;
        const double xr = pts[0][row], yr = pts[1][row];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = 16 * tj[t] + 4 * q + lg;
            const double v = (row < m && col < m) ? kernel_eval(k, xr, yr, pts[0][col], pts[1][col])
                                                  : (row == col ? 1.0 : 0.0);
            if (t < LF_RS) R[t][q] = v;
            else lt[(t - LF_RS) * 256 + q * 64 + lane] = v;
        }
        __builtin_amdgcn_sched_barrier(0);   // one tile's evaluations at a time
    }
}
}  // namespace

__global__ __launch_bounds__(512) void k_leaf_chol(const LeafChol* __restrict__ ds) {
    __shared__ double pts[2][LEAF_FUSED_MAX];
    __shared__ double pan[LF_NT][256];   // L_ip of the current block column, tile layout
    __shared__ double dinv[256];         // its inverse, tile layout
    __shared__ double trs[16 * 17];      // a transposition of the diagonal owner
    __shared__ double dtile[LF_NT][256];      // the diagonal tiles, tile layout
    __shared__ double ltile[LF_WAVES][(LF_SLOTS - LF_RS) * 256];   // each wave's last tile slots, tile layout
    __shared__ double pivs[LEAF_FUSED_MAX];   // diag(L), for the log sum at the end
    __shared__ double lsum[64 * LF_WAVES];
    const LeafChol d = ds[blockIdx.x];
    const int m = d.m, nt = (m + 15) >> 4;
    const int tid = threadIdx.x, lane = tid & 63, li = lane & 15, lg = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const KDev k = d.k;
    for (int e = tid; e < 16 * nt; e += 64 * LF_WAVES) {
        pts[0][e] = e < m ? d.px[e] : 0.0;
        pts[1][e] = e < m ? d.py[e] : 0.0;
    }
    __syncthreads();
    const int ntile = nt * (nt - 1) / 2;
    double* __restrict__ L = d.L;
    int ti[LF_SLOTS], tj[LF_SLOTS];
    d4 R[LF_SLOTS];
    // the owned tiles (block row -1: none); zero the mirrored upper tiles of L
#pragma unroll
    for (int t = 0; t < LF_SLOTS; ++t) {
        const int id = wave + LF_WAVES * t;
        ti[t] = tj[t] = -1;
        R[t] = d4{0.0, 0.0, 0.0, 0.0};
        if (id < ntile) {
            tile_ij(id, ti[t], tj[t]);
            const int ur = 16 * tj[t] + li;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int uc = 16 * ti[t] + 4 * q + lg;
                if (ur < m && uc < m) L[ur + (size_t)uc * m] = 0.0;
            }
        }
    }
    double* const lt = ltile[wave];
    switch (k.family) {
        case KMH_RATIONAL2: eval_tiles<KMH_RATIONAL2>(k, R, ti, tj, pts, m, li, lg, wave, nt, dtile, lt); break;
        case KMH_CAUCHY: eval_tiles<KMH_CAUCHY>(k, R, ti, tj, pts, m, li, lg, wave, nt, dtile, lt); break;
        case KMH_MATERN12: eval_tiles<KMH_MATERN12>(k, R, ti, tj, pts, m, li, lg, wave, nt, dtile, lt); break;
        default: eval_tiles<KMH_MATERN32>(k, R, ti, tj, pts, m, li, lg, wave, nt, dtile, lt); break;
    }
    bool bad = false;
    for (int p = 0; p < nt; ++p) {
        const int o = 16 * p;
        // the tile coordinates pass through an empty asm each step, so the
        // LDS addresses built from them are formed where they are used rather
        // than held in registers across the loop
#pragma unroll
        for (int t = 0; t < LF_SLOTS; ++t) __asm__ volatile("" : "+s"(ti[t]), "+s"(tj[t]));
        // 1. the diagonal tile: factor, write out, invert
        if (wave == p % LF_WAVES) {
            d4 st;   // the diagonal tile, tile layout
#pragma unroll
            for (int q = 0; q < 4; ++q) st[q] = dtile[p][q * 64 + lane];
            d4 tmb, tmbt;   // blockdiag(M_0..M_3) and its transpose, tile layout
            // four 4-column steps; step jb: the 4 x 4 diagonal block by
            // broadcasts, its Cholesky factor and inverse M_jb in uniform
            // arithmetic, the 16 x 4 column block below it as (M_jb S_col^T)^T
            // and the rank-4 update of the columns right of it on MFMA
#pragma unroll
            for (int jb = 0; jb < 4; ++jb) {
                const int o4 = 4 * jb;
                double sv[4][4];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr)
#pragma unroll
                    for (int gg = 0; gg <= rr; ++gg) sv[rr][gg] = bcast(st[jb], o4 + rr + 16 * gg);
                double lv[4][4] = {}, iv[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    double piv = sv[c][c];
#pragma unroll
                    for (int k2 = 0; k2 < c; ++k2) piv = __fma_rn(-lv[c][k2], lv[c][k2], piv);
                    bad |= (o + o4 + c < m) && !(piv > 0.0);
                    iv[c] = rsqrt(piv);
                    lv[c][c] = piv * iv[c];
                    if (lane == 0) pivs[o + o4 + c] = lv[c][c];
#pragma unroll
                    for (int rr = c + 1; rr < 4; ++rr) {
                        double t2 = sv[rr][c];
#pragma unroll
                        for (int k2 = 0; k2 < c; ++k2) t2 = __fma_rn(-lv[rr][k2], lv[c][k2], t2);
                        lv[rr][c] = t2 * iv[c];
                    }
                }
                // M = L_D^-1 by rows: M(r,c) = -(sum_{c<=k<r} L(r,k) M(k,c)) / L(r,r)
                double mv[4][4] = {};
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    mv[rr][rr] = iv[rr];
#pragma unroll
                    for (int c = 0; c < rr; ++c) {
                        double t2 = 0.0;
#pragma unroll
                        for (int k2 = c; k2 < rr; ++k2) t2 = __fma_rn(lv[rr][k2], mv[k2][c], t2);
                        mv[rr][c] = -t2 * iv[rr];
                    }
                }
                // lane-dependent picks of L_D's and M's entries (lower triangles)
                double apan = 0.0, bt = 0.0, bm = 0.0, ld = 0.0;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr)
#pragma unroll
                    for (int c = 0; c <= rr; ++c) {
                        apan = (li == rr && lg == c) ? mv[rr][c] : apan;        // M(i, k), i = li < 4
                        bt = (lg == rr && (li & 3) == c) ? mv[rr][c] : bt;      // Mb(4jb + k, j)
                        bm = ((li & 3) == rr && lg == c) ? mv[rr][c] : bm;      // Mb(i, 4jb + k)
                        ld = (li == o4 + rr && lg == c) ? lv[rr][c] : ld;       // L_D(i - o4, k)
                    }
                tmbt[jb] = (li >> 2) == jb ? bt : 0.0;
                tmb[jb] = (li >> 2) == jb ? bm : 0.0;
                // P = S_col L_D^-T (pc[0] = P(li, lg)), refined once:
                // P += (S_col - P L_D^T) L_D^-T, with P L_D^T = (L_D P^T)^T
                double al = 0.0;   // L_D(i, k), i = li < 4
#pragma unroll
                for (int rr = 0; rr < 4; ++rr)
#pragma unroll
                    for (int c = 0; c <= rr; ++c) al = (li == rr && lg == c) ? lv[rr][c] : al;
                const d4 z4 = d4{0.0, 0.0, 0.0, 0.0};
                const d4 pc = mfma(apan, st[jb], z4);
                const d4 pl = mfma(al, pc[0], z4);
                const d4 dp = mfma(apan, st[jb] - pl[0], z4);
                const double v = li > o4 + 3 ? pc[0] + dp[0] : ld;   // below the block, the block, zero above
                st[jb] = v;
                const double lmv = li > o4 + 3 ? v : 0.0;
                st = mfma(-lmv, lmv, st);
            }
            // L_pp (zero above its diagonal) to L
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = o + li, col = o + 4 * q + lg;
                if (row < m && col < m) L[row + (size_t)col * m] = st[q];
            }
            // L_pp^-1 = Mb (I + N)(I + N^2) with N = I - L_pp Mb (strictly
            // block lower, N^4 = 0), as four products on MFMA; N^T is
            // transposed through LDS
            d4 acc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = mfma(st[q], tmbt[q], acc);   // (L Mb)^T
            d4 tnt, tn, eye;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                eye[q] = li == 4 * q + lg ? 1.0 : 0.0;
                tnt[q] = eye[q] - acc[q];                                   // N^T
                trs[li * 17 + 4 * q + lg] = tnt[q];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) tn[q] = trs[(4 * q + lg) * 17 + li];   // N
            d4 n2 = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) n2 = mfma(tn[q], tnt[q], n2);           // (N^2)^T
            d4 qt = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) qt = mfma(eye[q] + tn[q], eye[q] + n2[q], qt);   // ((I+N)(I+N^2))^T
            d4 linv = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) linv = mfma(qt[q], tmb[q], linv);       // X = L_pp^-1
            // one Newton step X += X (I - L X), as X^T through LDS, (L X)^T and
            // X (I - L X) = ((I - L X)^T X^T)^T on MFMA
#pragma unroll
            for (int q = 0; q < 4; ++q) trs[li * 17 + 4 * q + lg] = linv[q];
            d4 xt;
#pragma unroll
            for (int q = 0; q < 4; ++q) xt[q] = trs[(4 * q + lg) * 17 + li];
            d4 lx = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) lx = mfma(st[q], xt[q], lx);           // (L X)^T
            d4 dx = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) dx = mfma(eye[q] - lx[q], linv[q], dx);   // X (I - L X)
            linv += dx;
            double* gl = d.linv + (size_t)p * 256;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                dinv[q * 64 + lane] = linv[q];
                gl[q * 64 + lane] = linv[q];
            }
        }
        __syncthreads();
        // 2. the tiles below it: L_ip = S_ip L_pp^-T, as (L_ip)^T = L_pp^-1 S_ip^T
#pragma unroll
        for (int t = 0; t < LF_SLOTS; ++t) {
            if (tj[t] == p && ti[t] > p) {
                d4 cur;
#pragma unroll
                for (int q = 0; q < 4; ++q) cur[q] = t < LF_RS ? R[t][q] : lt[(t - LF_RS) * 256 + q * 64 + lane];
                d4 acc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = mfma(dinv[q * 64 + lane], cur[q], acc);
                if (t < LF_RS) R[t] = acc;
                else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) lt[(t - LF_RS) * 256 + q * 64 + lane] = acc[q];
                }
                const int row = 16 * ti[t] + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    pan[ti[t]][q * 64 + lane] = acc[q];
                    const int col = o + 4 * q + lg;
                    if (row < m && col < m) L[row + (size_t)col * m] = acc[q];
                }
            }
        }
        __syncthreads();
        // 3. trailing tiles (i >= j > p): S_ij -= L_ip L_jp^T, as S_ij^T -= L_jp L_ip^T;
        //    the diagonal ones in LDS, each by the wave that factors it
        for (int j = p + 1 + (wave + LF_WAVES - (p + 1) % LF_WAVES) % LF_WAVES; j < nt; j += LF_WAVES) {
            d4 acc;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = dtile[j][q * 64 + lane];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = mfma(-pan[j][q * 64 + lane], pan[j][q * 64 + lane], acc);
#pragma unroll
            for (int q = 0; q < 4; ++q) dtile[j][q * 64 + lane] = acc[q];
        }
#pragma unroll
        for (int t = 0; t < LF_SLOTS; ++t) {
            if (tj[t] > p) {
                d4 acc;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = t < LF_RS ? R[t][q] : lt[(t - LF_RS) * 256 + q * 64 + lane];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc = mfma(-pan[tj[t]][q * 64 + lane], pan[ti[t]][q * 64 + lane], acc);
                if (t < LF_RS) R[t] = acc;
                else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) lt[(t - LF_RS) * 256 + q * 64 + lane] = acc[q];
                }
            }
        }
    }
    if (bad && lane == 0) atomicExch(d.info, 1);
    // sum of log diag(L): one log per thread, the partial sums in thread order
    __syncthreads();
    lsum[tid] = tid < m ? log(pivs[tid]) : 0.0;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int e = 0; e < 64 * LF_WAVES; ++e) s += lsum[e];
        *d.part = s;
    }
}

void leaf_chol(Ctx& c, const std::vector<LeafChol>& ls) {
    if (ls.empty()) return;
    for (const auto& l : ls)
        KMH_CHECK(l.m > 0 && l.m <= LEAF_FUSED_MAX, KMH_E_BAD_ARG, "leaf_chol: leaf beyond the fused size");
    LeafChol* dd = c.upload_vec(ls);
    KMH_LAUNCH(c, k_leaf_chol, dim3((unsigned)ls.size()), dim3(64 * LF_WAVES), dd);
}

__global__ __launch_bounds__(1024) void k_leaf_solve(const LeafSolve* __restrict__ ds) {
    __shared__ double xs[2][LS_CB][256];   // X_p, natural layout, by step parity
    const LeafSolve d = ds[blockIdx.x];
    const int m = d.m, nt = (m + 15) >> 4, nc = d.ncols;
    const int tid = threadIdx.x, lane = tid & 63, li = lane & 15, lg = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool own = w < nt;
    double* __restrict__ B = d.B;
    const double* __restrict__ L = d.L;
    // row tile w of B, natural layout: register q of column tile cb holds B(16w + 4q + lg, 16cb + li)
    d4 Bt[LS_CB];
#pragma unroll
    for (int cb = 0; cb < LS_CB; ++cb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * w + 4 * q + lg, col = 16 * cb + li;
            Bt[cb][q] = (own && row < m && col < nc) ? B[row + (size_t)col * d.ldb] : 0.0;
        }
    d4 inv = d4{0.0, 0.0, 0.0, 0.0};   // L_ww^-1, tile layout
    if (own) {
#pragma unroll
        for (int q = 0; q < 4; ++q) inv[q] = d.linv[(size_t)w * 256 + q * 64 + lane];
    }
    // L_wp, tile layout, loaded one step ahead
    const int lrow = 16 * w + li;
    auto load_l = [&](int p) {
        d4 v;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = 16 * p + 4 * q + lg;
            v[q] = (lrow < m && col < m) ? L[lrow + (size_t)col * m] : 0.0;
        }
        return v;
    };
    d4 lnext = d4{0.0, 0.0, 0.0, 0.0};
    if (own && w > 0) lnext = load_l(0);
    for (int p = 0; p < nt; ++p) {
        double* xp = &xs[p & 1][0][0];
        if (w == p) {
#pragma unroll
            for (int cb = 0; cb < LS_CB; ++cb) {
                d4 acc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = mfma(inv[q], Bt[cb][q], acc);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    xp[cb * 256 + q * 64 + lane] = acc[q];
                    const int row = 16 * p + 4 * q + lg, col = 16 * cb + li;
                    if (row < m && col < nc) B[row + (size_t)col * d.ldb] = acc[q];
                }
            }
        }
        __syncthreads();
        if (own && w > p) {
            const d4 lw = lnext;
            if (p + 1 < w) lnext = load_l(p + 1);
#pragma unroll
            for (int cb = 0; cb < LS_CB; ++cb) {
                d4 acc = Bt[cb];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = mfma(-lw[q], xp[cb * 256 + q * 64 + lane], acc);
                Bt[cb] = acc;
            }
        }
    }
}

void leaf_solve(Ctx& c, const std::vector<LeafSolve>& ls) {
    std::vector<LeafSolve> ds;
    for (const auto& l : ls) {
        KMH_CHECK(l.m > 0 && l.m <= LEAF_FUSED_MAX, KMH_E_BAD_ARG, "leaf_solve: leaf beyond the fused size");
        for (int c0 = 0; c0 < l.ncols; c0 += LS_CW) {
            LeafSolve s = l;
            s.B = l.B + (size_t)c0 * l.ldb;
            s.ncols = std::min(LS_CW, l.ncols - c0);
            ds.push_back(s);
        }
    }
    if (ds.empty()) return;
    LeafSolve* dd = c.upload_vec(ds);
    KMH_LAUNCH(c, k_leaf_solve, dim3((unsigned)ds.size()), dim3(1024), dd);
}

}  // namespace kmh
