/*
 * kmhodlr -- C ABI of the MI355X (gfx950) HODLR Gaussian-process engine.
 *
 * The reference (KernelMatrices.jl 0.9.0) has no FFI: its boundary is the
 * Julia method surface of the HODLR path.  Each entry point below names the
 * reference method it replaces (file:line under the reference tree).  A Julia
 * shim binds these with ccall; the Python binding uses ctypes.
 *
 * Conventions
 *   - Every call returns an int status (KMH_OK = 0).  On failure
 *     kmh_last_error() gives a thread-local message that bindings rethrow
 *     with the reference's exception text.
 *   - Points are n x d doubles, row-major (point i at pts[i*d .. i*d+d-1]);
 *     d must be 2.
 *   - Vectors / right-hand sides are n x nrhs doubles, column-major.
 *   - Functions without a suffix take HOST pointers and are synchronous.
 *     Functions with the suffix _dev take DEVICE pointers (on the context's
 *     device) and are ordered on the context's stream; call
 *     kmh_ctx_synchronize() before reading results on the host.
 *   - The library owns all device memory of a handle.  A handle is not
 *     thread-safe; distinct handles on one context share its stream.
 */
#ifndef KMHODLR_H
#define KMHODLR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define KMH_OK               0
#define KMH_E_BAD_ARG        1  /* invalid argument (sizes, ranks, kernel)      */
#define KMH_E_NOT_PD         2  /* Cholesky met a non-positive pivot (PosDefException) */
#define KMH_E_NOT_FACTORIZED 3  /* solve/logdet/gradient before factorization   */
#define KMH_E_OOM            4  /* device allocation failed                      */
#define KMH_E_BUDGET         5  /* request exceeds the context's HBM budget      */
#define KMH_E_DEVICE         6  /* HIP runtime error                             */
#define KMH_E_ACA            7  /* "Incompatible maxrank or matrix dimensions."  */
#define KMH_E_STATE          8  /* operation invalid in the handle's state       */

/* ---- kernel families (the reference's kernel closures) ---- */
#define KMH_RATIONAL2  1  /* p1^2/(1+(r/p2)^2)^2          README.md:42-48              */
#define KMH_CAUCHY     2  /* p1^2/(1+(p2 r)^2)            example/basicexample.jl:62-68 */
#define KMH_MATERN12   3  /* t0*exp(-r/rho)               src/covariancefunctions.jl:258-266, nu=1/2 */
#define KMH_MATERN32   4  /* t0*(1+sqrt3 r/rho)exp(-sqrt3 r/rho), nu=3/2, same lines */

/* ---- derivative codes (src/covariancefunctions.jl:268-330, 424-472) ---- */
#define KMH_D0   0   /* the kernel itself (nugget applied)   */
#define KMH_D1   1   /* d/dp1                                */
#define KMH_D2   2   /* d/dp2                                */
#define KMH_D11 11   /* d2/dp1^2                             */
#define KMH_D12 12   /* d2/dp1 dp2                           */
#define KMH_D22 22   /* d2/dp2^2                             */

typedef struct kmh_kernel {
    int32_t family;     /* KMH_RATIONAL2 ... KMH_MATERN32                     */
    int32_t deriv;      /* KMH_D0 ... ; derivative kernels carry no nugget    */
    double  params[4];  /* (p1, p2) = (t0 = sigma^2, rho) for Matern          */
    double  nugget;     /* added where the two points are equal (x == y)      */
} kmh_kernel;

/* ---- build options (constructor.jl:6-8 keyword/positional arguments) ---- */
#define KMH_ACA        0
#define KMH_NYSTROM    1
#define KMH_FIXED_LEVEL 0   /* structstypes.jl:30-32 */
#define KMH_LOG_LEVEL   1   /* structstypes.jl:35-37 */
#define KMH_FP64       0
#define KMH_MIXED      1   /* FP32 off-diagonal panels, cascade panels and Q
                              factors (FP64 MFMA accumulation); FP64 leaves,
                              capacitance and logdet; solves refined to FP64
                              against the FP64 HODLR matvec                  */
#define KMH_MIXED_Q    2   /* as KMH_MIXED with the Nystrom cascade panels in
                              FP64: only the panel QR and its Q factors are
                              FP32 (ACA panels are FP32 in both modes)       */

typedef struct kmh_build_opts {
    int32_t method;      /* KMH_ACA or KMH_NYSTROM (kwarg nystrom)            */
    int32_t level_kind;  /* KMH_FIXED_LEVEL or KMH_LOG_LEVEL                  */
    int32_t level;       /* lv of FixedLevel(lv) / LogLevel(lv)               */
    int32_t maxrank;     /* maxrank (0 = no limit for ACA; Nystrom rank)      */
    double  tol;         /* ep: ACA relative tolerance                        */
    int32_t precision;   /* KMH_FP64, KMH_MIXED or KMH_MIXED_Q                */
    int32_t reserved[3];
} kmh_build_opts;

typedef struct kmh_info {
    int64_t n;           /* matrix size                                       */
    int32_t levels;      /* number of off-diagonal levels                     */
    int32_t nleaves;     /* 2^levels                                          */
    int32_t factorized;  /* 1 after kmh_hodlr_factorize                       */
    int32_t method;      /* KMH_ACA / KMH_NYSTROM                             */
    int64_t device_bytes;/* device memory held by the handle                  */
} kmh_info;

/* factor-apply operations on W (K = W W^T), HODLR/src/baseoverloads.jl:95-150 */
#define KMH_W      0   /* y = W x      */
#define KMH_WT     1   /* y = W^T x    */
#define KMH_WINV   2   /* y = W \ x    */
#define KMH_WTINV  3   /* y = W^T \ x  */

typedef struct kmh_ctx kmh_ctx;
typedef struct kmh_hodlr kmh_hodlr;

/* ---- errors and devices ---- */
const char* kmh_last_error(void);
int kmh_device_count(int* count);

/* ---- context: device, stream, HBM budget for handles (0 = unlimited) ----
 * HODLR handles and servers keep their context alive: destroying a context
 * that still has them only marks it; it is torn down with the last one.   */
int kmh_ctx_create(int device, uint64_t hbm_budget_bytes, kmh_ctx** out);
int kmh_ctx_destroy(kmh_ctx* ctx);
int kmh_ctx_synchronize(kmh_ctx* ctx);
int kmh_ctx_bytes_in_use(kmh_ctx* ctx, uint64_t* out);
/* Grow the context's device pool so that it holds a free block of at least
 * `bytes` (one device allocation, nothing when such a block exists); the
 * memory stays with the context and serves later handles and transients.
 * No reference counterpart: for callers that size the pool before a
 * latency-sensitive phase, so that it makes no device allocation.          */
int kmh_ctx_reserve(kmh_ctx* ctx, uint64_t bytes);
/* Measurement hooks (no counterpart in the reference): between begin and end
 * every FP64 GEMM kernel launch and every 24-column tile QR launch is
 * bracketed by HIP event markers on the context's stream, so a span is one
 * launch (its duration plus the launch latency; each marker pair idles the
 * queue for several microseconds, see kmh_ctx_profile_config).  end returns
 * the KMH_PROF_GEMM totals (algorithmic
 * flops 2mnk, summed kernel duration in ms, launch count) of the timed GEMM
 * launches; query returns any tag's totals of the last begin..end window:
 *   KMH_PROF_GEMM     every GEMM kernel launch (split-k reductions included,
 *                     with 0 flops), amount = flops
 *   KMH_PROF_NN_WIDE  k_gemm_nn_wide launches, amount = algorithmic HBM bytes
 *                     (A, B and C written, plus C read when beta != 0)
 *   KMH_PROF_TN_TALL  k_gemm_tn_tall launches, amount = bytes of A, B, C
 *   KMH_PROF_NN_ROWS  k_gemm_nn_rows launches, amount = bytes as for
 *                     KMH_PROF_NN_WIDE
 *   KMH_PROF_TILE_QR  k_tile_qr<24> launches (Householder QR of <= 128 x 24
 *                     tiles, Q left in compact-WY form), amount = bytes of
 *                     the tiles read and of the Householder vectors and R
 *                     written                                                */
enum {
    KMH_PROF_GEMM = 0,
    KMH_PROF_NN_WIDE = 1,
    KMH_PROF_TN_TALL = 2,
    KMH_PROF_NN_ROWS = 3,
    KMH_PROF_TILE_QR = 4,
    KMH_PROF_TAGS = 5
};
int kmh_ctx_profile_begin(kmh_ctx* ctx);
int kmh_ctx_profile_end(kmh_ctx* ctx, double* gemm_flops, double* gemm_ms, int64_t* gemm_launches);
int kmh_ctx_profile_query(kmh_ctx* ctx, int32_t tag, double* amount, double* ms, int64_t* spans);
/* Which launches the hooks time (default: every tag, every launch): the tags
 * whose bit (1 << tag) is set in tag_mask, each at every stride-th launch of
 * that tag, counted from begin.  A timed launch leaves several microseconds
 * of idle around it, so a throughput run times a sample.  The KMH_PROF_GEMM
 * totals of end cover the timed launches only.  Not inside begin..end. */
int kmh_ctx_profile_config(kmh_ctx* ctx, uint32_t tag_mask, int32_t stride);

/* GEMM autotuner (no reference counterpart).  For each of count shapes
 * (6 ints each: ta, tb, m, n, k, problems per launch) time every kernel
 * class that can take the shape (generic MFMA tiles, tall TN, row-per-thread
 * NN, wide NN), median of reps launches, and keep the fastest in the
 * context's table keyed by (ta, tb, ceil(log2) of m, n, k); the table is
 * written to path (text) when path is not null.  tuning_load replaces the
 * context's table by a file's (null path: clear it; the built-in shape
 * rules apply to every class the table does not name).                     */
int kmh_gemm_autotune(kmh_ctx* ctx, int32_t count, const int32_t* shapes, int32_t reps, const char* path);
int kmh_gemm_tuning_load(kmh_ctx* ctx, const char* path);
int kmh_gemm_tuning_size(kmh_ctx* ctx, int32_t* entries);

/* ---- HODLR.KernelHODLR(K, ep, maxrank, lvl; nystrom)  constructor.jl:6-58 ----
 * K = KernelMatrix(pts, pts, params, kernel)  src/structstypes.jl:17-23      */
int kmh_hodlr_build(kmh_ctx* ctx, const double* pts, int64_t n, int32_t d,
                    const kmh_kernel* kernel, const kmh_build_opts* opts,
                    kmh_hodlr** out);
int kmh_hodlr_build_dev(kmh_ctx* ctx, const double* pts_dev, int64_t n, int32_t d,
                        const kmh_kernel* kernel, const kmh_build_opts* opts,
                        kmh_hodlr** out);
int kmh_hodlr_destroy(kmh_hodlr* h);
int kmh_hodlr_info(kmh_hodlr* h, kmh_info* out);
/* wait for the handle's queued device work (its context's stream)          */
int kmh_hodlr_synchronize(kmh_hodlr* h);

/* leaf boundaries: nleaves+1 offsets (indices.jl:29-37)                      */
int kmh_hodlr_leaf_offsets(kmh_hodlr* h, int64_t* offsets);
/* per-block ranks, level-major: level 1 (finest) first, 2^(L-l) blocks each;
 * total 2^L - 1 entries (length of each U[l][k], V[l][k] column set)         */
int kmh_hodlr_ranks(kmh_hodlr* h, int32_t* ranks);
/* dense full(HK) before factorization, n x n column-major
 * (HODLR/src/baseoverloads.jl:313-326; for tests)                            */
int kmh_hodlr_full(kmh_hodlr* h, double* out);
/* U and V of block k at level l (1-based level): U is rows x rank, V is
 * cols x rank, column-major; only before factorization                      */
int kmh_hodlr_block_uv(kmh_hodlr* h, int32_t level, int32_t block, double* U, double* V);

/* ---- HK * x  (HODLR/src/baseoverloads.jl:152-179) ---- */
int kmh_hodlr_matvec(kmh_hodlr* h, const double* x, double* y, int64_t nrhs);
int kmh_hodlr_matvec_dev(kmh_hodlr* h, const double* x, double* y, int64_t nrhs);

/* ---- HODLR.symmetricfactorize!(HK)  HODLR/src/symfact.jl:6-50 ---- */
int kmh_hodlr_factorize(kmh_hodlr* h);

/* ---- mul!/ldiv!(t, HK.W or adjoint(HK.W), v)  HODLR/src/baseoverloads.jl:95-150 */
int kmh_hodlr_factor_apply(kmh_hodlr* h, int32_t op, const double* x, double* y, int64_t nrhs);
int kmh_hodlr_factor_apply_dev(kmh_hodlr* h, int32_t op, const double* x, double* y, int64_t nrhs);

/* ---- HK \ b  (HODLR/src/baseoverloads.jl:181-192, 252-263) ---- */
int kmh_hodlr_solve(kmh_hodlr* h, const double* b, double* x, int64_t nrhs);
int kmh_hodlr_solve_dev(kmh_hodlr* h, const double* b, double* x, int64_t nrhs);
/* kmh_hodlr_solve_dev on an FP64 handle replays a hipGraph of the solve for
 * repeated (b, x, nrhs): the second solve with the same device buffers
 * captures the launches (descriptors and temporaries in blocks the graph
 * owns), later ones replay them; results are bitwise those of the uncaptured
 * solve.  At most 8 graphs per handle, freed with it.  enable: 1 on
 * (default), 0 off (frees the graphs), other values leave it; returns the
 * captures and replays so far.  No reference counterpart.                   */
int kmh_hodlr_solve_graphs(kmh_hodlr* h, int32_t enable, int64_t* captures, int64_t* replays);

/* The same solve without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * stream capture, serving loops, harnesses that forbid hipMalloc).  Every
 * temporary lives in one fixed 32 MiB device array per GPU, held until the
 * process ends; a call whose temporaries could exceed it (large nrhs or
 * ranks) fails with KMH_E_BAD_ARG before any work is queued.  Calls that
 * overlap on one GPU (different contexts or threads) race on that array.  */
int kmh_hodlr_solve_dev_noalloc(kmh_hodlr* h, const double* b, double* x, int64_t nrhs);

/* Iterative refinement of the last solve of a KMH_MIXED handle (no
 * reference counterpart): corrections made and the final largest relative
 * residual |b - K x| / |b| over the right-hand sides, K the FP64 HODLR
 * operator (leaves as L L^T, off-diagonal U V^T); 0 and 0 for FP64 handles */
int kmh_hodlr_refinement(kmh_hodlr* h, int32_t* iters, double* relres);

/* ---- logdet(HK)  HODLR/src/baseoverloads.jl:194-209 ---- */
int kmh_hodlr_logdet(kmh_hodlr* h, double* out);

/* ---- batched form: independent problems (different n, kernels, ranks)
 * advance together, one grouped launch per step across all of them.  No
 * reference counterpart: the reference loops problems one by one.  Each
 * request gets its own status; a failing request does not stop the others. */
typedef struct kmh_build_req {
    const double* pts;      /* n x d row-major; host or device (on_device)    */
    int64_t n;
    int32_t d;
    int32_t on_device;
    kmh_kernel kernel;
    kmh_build_opts opts;
} kmh_build_req;
int kmh_hodlr_build_batch(kmh_ctx* ctx, int32_t count, const kmh_build_req* reqs,
                          kmh_hodlr** out, int32_t* status);
int kmh_hodlr_factorize_batch(kmh_ctx* ctx, int32_t count, kmh_hodlr* const* hs, int32_t* status);
/* x_i = K_i^{-1} b_i for device blocks (n_i x nrhs_i); b_i may equal x_i    */
int kmh_hodlr_solve_batch_dev(kmh_ctx* ctx, int32_t count, kmh_hodlr* const* hs,
                              const double* const* b, double* const* x, const int64_t* nrhs);

/* ---- serving: continuous batching of GP likelihood requests ----
 * One request = nll_objective (maxlikfunctions.jl:29-50): build, factorize,
 * negloglik and, when derivative kernels are given, stoch_gradient.  No
 * reference counterpart (the reference evaluates one objective at a time).
 * kmh_server_step admits waiting requests whose HBM footprint fits the budget
 * (0 = unlimited) while fewer than max_inflight run, then advances every
 * in-flight request by one stage (build, leaf stage, one tree level, finish)
 * with the stages of all requests in shared grouped launches.  Inputs are
 * copied at submit; results are kept until read with kmh_server_result.
 * The in-flight requests are spread over lanes (a new request joins the lane
 * with the fewest): lane 0 runs on the server's context, every other lane on
 * a private context (own HIP stream and scratch pool, no HBM budget of its
 * own: bound a multi-lane server with hbm_budget_bytes), and a step advances
 * the lanes on parallel host threads so that their work overlaps on the
 * device.  Default: 2 lanes when the context has no HBM budget and
 * max_inflight >= 2, else 1.  Results do not depend on the lane count.    */
#define KMH_SERVE_MAX_PARAMS 4
#define KMH_SERVE_MAX_LANES 4
typedef struct kmh_server kmh_server;
typedef struct kmh_serve_req {
    const double* pts;        /* n x 2 row-major (host, or device if on_device) */
    int64_t n;
    int32_t d;
    int32_t ndk;              /* number of derivative kernels (0: no gradient) */
    int32_t on_device;        /* pts, y and saa are device pointers the caller
                                 keeps alive until the request completes      */
    int32_t keep_factor;      /* 1: keep the factorized HODLR resident after
                                 completion for follow-up solves
                                 (kmh_server_factor) until kmh_server_release */
    kmh_kernel kernel;
    kmh_build_opts opts;
    const double* y;          /* data vector, n                             */
    const double* saa;        /* n x nv SAA vectors (column-major) if ndk > 0 */
    int64_t nv;
    kmh_kernel dks[KMH_SERVE_MAX_PARAMS];
} kmh_serve_req;
typedef struct kmh_serve_result {
    int32_t status;           /* KMH_OK or the request's error              */
    int32_t ngrad;
    double  nll;              /* 0.5 logdet + 0.5 y' K^-1 y                 */
    double  logdet;
    double  grad[KMH_SERVE_MAX_PARAMS];
    double  queue_us;         /* submit -> admission                        */
    double  latency_us;       /* submit -> completion                       */
    int32_t resident;         /* 1: the factor is resident (keep_factor)    */
    int32_t reserved;
    int64_t factor_bytes;     /* device bytes the resident factor holds     */
} kmh_serve_result;
int kmh_server_create(kmh_ctx* ctx, int32_t max_inflight, uint64_t hbm_budget_bytes, kmh_server** out);
/* 1..KMH_SERVE_MAX_LANES lanes; KMH_E_STATE while requests are queued or
 * running (no reference counterpart)                                        */
int kmh_server_set_lanes(kmh_server* srv, int32_t lanes);
int kmh_server_destroy(kmh_server* srv);
/* Size every lane's pool for `requests` requests of n points built with
 * `opts` and carrying nv SAA vectors, dealt evenly over the lanes, by the
 * server's own per-request footprint estimate (kmh_ctx_reserve per lane), so
 * that serving them makes no device allocation.  Call after
 * kmh_server_set_lanes.  No reference counterpart.                          */
int kmh_server_reserve(kmh_server* srv, int64_t n, const kmh_build_opts* opts, int32_t nv, int32_t requests);
int kmh_server_submit(kmh_server* srv, const kmh_serve_req* req, int64_t* id);
/* A failure inside a step (device allocation, budget, HIP error) fails every
 * request of the lane it happened on: their results carry the status and
 * message, their device memory and reservations are released, and the step
 * returns KMH_OK; other lanes and later steps proceed.  ACA requests are
 * admitted on an estimate of their ranks; after the build the reservation
 * follows the actual ranks, and a request that then exceeds the budget fails
 * alone with KMH_E_BUDGET.  completed counts finished requests, failed ones
 * included.                                                                 */
int kmh_server_step(kmh_server* srv, int32_t* completed);
int kmh_server_pending(kmh_server* srv, int32_t* waiting, int32_t* inflight, uint64_t* reserved_bytes);
/* KMH_E_STATE while the request is still queued or running.  A request
 * without a resident factor is forgotten once its result is read; one with
 * a resident factor stays until kmh_server_release.                        */
int kmh_server_result(kmh_server* srv, int64_t id, kmh_serve_result* out);
/* The resident factor of a completed keep_factor request, as a HODLR handle
 * lent by the server: the factorized HK after nll_objective, for follow-up
 * HK \ b, logdet and W applies (HODLR/src/baseoverloads.jl:95-209) with the
 * kmh_hodlr_* entry points.  The handle lives on the context of the lane
 * that served the request (the server's context for lane 0): its host-pointer
 * calls are synchronous, its _dev calls are ordered on that lane's stream
 * (kmh_hodlr_synchronize waits for them).  Valid until kmh_server_release
 * or kmh_server_destroy; kmh_hodlr_destroy refuses it.  Its bytes stay
 * reserved against the server's HBM budget.  (No reference counterpart.) */
int kmh_server_factor(kmh_server* srv, int64_t id, kmh_hodlr** out);
/* Forget a completed request: frees its resident factor and reservation.  */
int kmh_server_release(kmh_server* srv, int64_t id);

/* ---- HODLR.stoch_gradient(K, HK, y, dfuns, vecs)  stochasticgradient.jl:30-45
 * dks: p derivative kernels; y: n; saa: n x nv column-major; grad: p        */
int kmh_stoch_gradient(kmh_hodlr* h, const kmh_kernel* dks, int32_t p,
                       const double* y, const double* saa, int64_t nv, double* grad);
/* same with y and saa in device memory (SAA vectors are fixed across an
 * optimisation, so callers keep them resident); grad stays on the host      */
int kmh_stoch_gradient_dev(kmh_hodlr* h, const kmh_kernel* dks, int32_t p,
                           const double* y, const double* saa, int64_t nv, double* grad);

/* ---- HODLR.stoch_profile_gradient(K, HK, y, dfuns, vecs)  stochasticgradient.jl:48-86
 * gradient of the profile negative log-likelihood
 * 0.5 (logdet K + n log(y' K^-1 y)):
 * grad_j = 0.5 (trace_j - n y' K^-1 dK_j K^-1 y / y' K^-1 y)              */
int kmh_stoch_profile_gradient(kmh_hodlr* h, const kmh_kernel* dks, int32_t p,
                               const double* y, const double* saa, int64_t nv, double* grad);

/* ---- HODLR.stoch_fisher(K, HK, DKs, vecs)  stochastichessian.jl:91-115
 * F: p x p column-major (symmetric)                                         */
int kmh_stoch_fisher(kmh_hodlr* h, const kmh_kernel* dks, int32_t p, const double* saa, int64_t nv, double* F);

/* ---- HODLR.stoch_hessian(K, HK, y, d1funs, d2funs, vecs)  stochastichessian.jl:117-147
 * d2: p(p+1)/2 kernels in the order (1,1),(1,2)..(1,p),(2,2)..: family 0
 * stands for the reference's ZeroFunction.  exact_second = 0 reproduces the
 * reference's second-derivative block terms (HODLR/src/utils.jl:110-155);
 * exact_second = 1 uses the exact second derivative of the Nystrom blocks
 * (the reference's terms differ from it off the diagonal blocks).
 * Hout: p x p column-major (symmetric)                                      */
int kmh_stoch_hessian(kmh_hodlr* h, const kmh_kernel* d1, int32_t p, const kmh_kernel* d2,
                      const double* y, const double* saa, int64_t nv, int32_t exact_second, double* Hout);

/* ---- point ordering: maxlikdata(pts=..., sortmethod=:KDTree)
 * HODLR/src/constructor.jl:171-181 (NearestNeighbors.KDTree(pts).data).
 * Host-only (no device work).  perm: n indices; pts[perm[i]] is the i-th
 * sorted point.  Cells of more than leafsize points are split at their
 * midpoint after a stable sort along the coordinate of largest spread.      */
int kmh_kd_order(const double* pts, int64_t n, int32_t d, int32_t leafsize, int64_t* perm);

/* ---- HODLR.RKernelHODLR(K, tol, maxrank, lvl)  constructor.jl:130-161 ------
 * Recursive HODLR of K(x1, x2) (x1 and x2 may differ, also in size) for
 * matvecs: rows and columns split at floor(len/2), ACA (opts->tol,
 * opts->maxrank) of both off-diagonal blocks of every node, dense diagonal
 * blocks at the last level; level 0 is the dense matrix (:141).  opts:
 * level_kind/level as for kmh_hodlr_build (LogLevel(k) -> FixedLevel(
 * floor(log2(min(n1, n2))) - k), :133-136); the level is clamped to
 * [0, floor(log2(min(n1, n2)))]; opts->method is ignored.  pts1 n1 x d,
 * pts2 n2 x d row-major (host), d = 2.                                    */
typedef struct kmh_rhodlr kmh_rhodlr;
int kmh_rhodlr_build(kmh_ctx* ctx, const double* pts1, int64_t n1, const double* pts2, int64_t n2, int32_t d,
                     const kmh_kernel* kernel, const kmh_build_opts* opts, kmh_rhodlr** out);
/* levels and the matrix size (any output may be null)                        */
int kmh_rhodlr_levels(kmh_rhodlr* h, int32_t* levels, int64_t* n1, int64_t* n2);
/* ACA ranks, depth by depth from the root, nodes left to right, (K12, K21)
 * per node: 2 (2^levels - 1) entries                                          */
int kmh_rhodlr_ranks(kmh_rhodlr* h, int32_t* ranks);
/* R * x  baseoverloads.jl:299-303: x n2 x nrhs, y n1 x nrhs column-major;
 * host buffers, or device buffers stream-ordered on the context (_dev)      */
int kmh_rhodlr_matvec(kmh_rhodlr* h, const double* x, double* y, int64_t nrhs);
int kmh_rhodlr_matvec_dev(kmh_rhodlr* h, const double* x, double* y, int64_t nrhs);
int kmh_rhodlr_destroy(kmh_rhodlr* h);

#ifdef __cplusplus
}
#endif
#endif /* KMHODLR_H */
