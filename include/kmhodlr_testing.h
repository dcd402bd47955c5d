/*
 * Test hooks of libkmhodlr: direct access to the batched dense primitives
 * (GEMM, Cholesky, triangular solve, panel QR) on one matrix, host buffers
 * in and out.  They have no counterpart in the reference; they exist so the
 * GPU test-suite can check each primitive against numpy before the HODLR
 * algorithms that are built from them.
 */
#ifndef KMHODLR_TESTING_H
#define KMHODLR_TESTING_H

#include "kmhodlr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C = alpha op(A) op(B) + beta C, column-major; lower != 0 updates only the
 * 64x64 tiles on or below the diagonal */
int kmh_test_gemm(kmh_ctx* ctx, int32_t ta, int32_t tb, int32_t m, int32_t n, int32_t k,
                  const double* A, int32_t lda, const double* B, int32_t ldb,
                  double* Cm, int32_t ldc, double alpha, double beta, int32_t lower);
/* in-place lower Cholesky; returns KMH_E_NOT_PD on a non-positive pivot */
int kmh_test_potrf(kmh_ctx* ctx, int32_t n, double* A, int32_t lda);
/* B <- L^{-1} B (trans = 0) or L^{-T} B (trans = 1) */
int kmh_test_trsm(kmh_ctx* ctx, int32_t trans, int32_t n, const double* Lm, int32_t ldl,
                  int32_t ncols, double* B, int32_t ldb);
/* thin QR of an m x r panel: A <- Q, R (r x r) */
int kmh_test_qr(kmh_ctx* ctx, int32_t m, int32_t r, double* A, int32_t lda, double* R, int32_t ldr);
/* capacitance factors of one lrsymfact block (r <= 80): from R1, R2 (r x r
 * upper, ld ldr) the last r rows of X = chol(C) - I and G = I - chol(C)^-1
 * (2r x 2r, ld ld >= 2r; the first r rows are left as given), *logsum =
 * sum log diag(chol(I - Bm^T Bm)), Bm = R1 R2^T; returns KMH_E_NOT_PD when
 * I - Bm^T Bm is not positive definite                                   */
int kmh_test_capacitance(kmh_ctx* ctx, int32_t r, const double* R1, const double* R2, int32_t ldr, double* X,
                         double* G, int32_t ld, double* logsum);
/* descriptor-arena ring: `count` uploads of `words` 64-bit words, each read
 * back by a kernel that checks its contents; *bad = uploads whose kernel saw
 * other data (0 when the ring's fences hold) */
int kmh_test_arena(kmh_ctx* ctx, int32_t count, int32_t words, int64_t* bad);
/* the timed launches of the open profiling window (kmh_ctx_profile_begin),
 * in launch order: tag, algorithmic amount, flops, kernel ms, workgroups;
 * up to cap entries, *count = all of them                                */
int kmh_test_profile_spans(kmh_ctx* ctx, int32_t cap, int32_t* tags, double* amounts, double* flops, double* ms,
                           int32_t* grids, int32_t* count);

/* failure injection: the context's next `after` scratch allocations succeed
 * and the one after fails with KMH_E_OOM on the host (no device call);
 * after < 0 turns it off                                                    */
int kmh_test_fail_scratch(kmh_ctx* ctx, int64_t after);
/* bytes of the context's scratch pool handed out and not returned          */
int kmh_test_scratch_live(kmh_ctx* ctx, uint64_t* bytes);
/* number of blocks the context's scratch pool has handed out so far        */
int kmh_test_pool_allocs(kmh_ctx* ctx, uint64_t* count);
/* replace the descriptor arena by one of `bytes` (after draining the
 * stream), so that the next large uploads make it grow                     */
int kmh_test_arena_shrink(kmh_ctx* ctx, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* KMHODLR_TESTING_H */
