// Batched dense linear algebra on gfx950 (FP64): the primitives the HODLR
// build, factorization, solve and gradient are composed of.
//
// All routines take lists of problems of uneven sizes and run each step as
// one grouped launch on the context's stream.  Matrices are column-major.
#pragma once
#include "common.h"

namespace kmh {

// C = alpha * op(A) * op(B) + beta * C, op(A) is m x k, op(B) is k x n.
// beta == 0 never reads C.  lower != 0 skips 64x64 output tiles strictly
// above the diagonal (symmetric rank-k updates of a lower triangle).
// Cin (optional, same leading dimension as C): the beta term reads Cin
// instead of C (out-of-place update); only for untransposed products with
// n <= 32 and k <= 64 (k <= 128 for n <= 16, k <= 256 for n <= 8).
// Storage types: an operand flagged f32 (ea, eb, ec; Cin follows C) is
// FP32 in memory behind the double pointer (reinterpret it with f32p / f64p
// below); it is widened on load and the product accumulates in FP64 on the
// FP64 MFMA (these products are HBM-bound, so FP32 storage halves their
// bytes while FP64 accumulation costs nothing); an FP32 C is rounded on
// store.  Supported (ea, eb, ec): all FP64, (0,1,1), (1,1,0), (1,0,1),
// (1,0,0).
struct GemmProb {
    int m, n, k;
    const double* A; int lda;
    const double* B; int ldb;
    double* C; int ldc;
    double alpha, beta;
    int lower;
    const double* Cin = nullptr;
    uint8_t ea = 0, eb = 0, ec = 0;
};
// An FP32 array passed through a double-typed descriptor field.
inline const double* f64p(const float* p) { return reinterpret_cast<const double*>(p); }
inline double* f64p(float* p) { return reinterpret_cast<double*>(p); }
inline float* f32p(double* p) { return reinterpret_cast<float*>(p); }
inline const float* f32p(const double* p) { return reinterpret_cast<const float*>(p); }
void gemm(Ctx& c, bool ta, bool tb, const std::vector<GemmProb>& probs);

// Global-memory view of a pointer.  Pointers a kernel reads from a
// descriptor are generic, and generic (flat) loads and stores count on the
// LDS counter as well, so every LDS wait after one also waits for it (a
// prefetch stops overlapping the compute that reads LDS); global accesses
// leave LDS waits alone.
template <class T> using Glob = __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ Glob<T> glob(T* p) { return (Glob<T>)p; }

// v where keep holds, else +0.0, by a bit mask.  Kernels load a clamped,
// always valid element and zero it with this instead of `keep ? load : 0`:
// the compiler turns that select into a branch around the load with a wait
// for it inside, which serialises a row of loads on the memory latency.
// The empty asm keeps the mask opaque, so the AND is not folded back into
// the select.
__device__ __forceinline__ double keep_or_zero(double v, bool keep) {
    const unsigned long long m = keep ? ~0ull : 0ull;
    return __longlong_as_double(__double_as_longlong(v) & m);
}

// A square matrix (or its leading n x n part) with leading dimension ld.
struct SqMat {
    double* A;
    int n;
    int ld;
    int* info = nullptr;   // failure flag for potrf (null: the context's flag)
};
// In-place lower Cholesky of each matrix; the strict upper triangle is
// zeroed.  A non-positive or NaN pivot sets *info (or ctx.d_info[0]) to 1.
void potrf(Ctx& c, const std::vector<SqMat>& mats);

// B <- L^{-1} B (trans = false) or L^{-T} B (trans = true), L lower n x n.
struct TrsmProb {
    const double* L; int n; int ldl;
    double* B; int ncols; int ldb;
    uint8_t eb = 0;   // B stored in FP32 (solved in FP64, rounded on store); left solves only
};
void trsm_left(Ctx& c, bool trans, const std::vector<TrsmProb>& probs);

// Thin QR of each m x r panel (m >= r): A <- Q (orthonormal columns),
// R (r x r upper, leading dimension ldr) written in full (lower part zero).
// Block classical Gram-Schmidt with reorthogonalisation across 24-column
// blocks, TSQR (single-wave Householder tiles, fan-in-4 tree) inside.  With
// src set, the panel is read from src (leading dimension lds) and only Q is
// written to A.
// e32: the panel A (the Q output) and src are FP32 in memory (FP64
// arithmetic, Q rounded on store); R stays FP64.
struct QrPanel {
    double* A; int m; int r; int lda;
    double* R; int ldr;
    const double* src = nullptr;
    int lds = 0;
    uint8_t e32 = 0;
};
void qr_panels(Ctx& c, const std::vector<QrPanel>& panels);

// One block Gram-Schmidt pass of the panel QR fused with the projection of
// the next: X = S - Q T1 (written to X; S may be X itself) and T2 = Q^T X,
// for Q m x kq (kq <= CGS_KQ_MAX), S and X m x nb (nb <= CGS_NB_MAX),
// T1 kq x nb.  FP64 only.
constexpr int CGS_KQ_MAX = 48;
constexpr int CGS_NB_MAX = 32;
struct CgsProb {
    const double* Q; int ldq; int m; int kq;
    const double* S; int lds;
    const double* T1; int ldt1;
    double* X; int ldx;
    int nb;
    double* T2; int ldt2;
};
void cgs_pass(Ctx& c, const std::vector<CgsProb>& ps);

// The capacitance factors of one block of lrsymfact (HODLR/src/utils.jl:29-35)
// for r <= CAP_MAX, in one workgroup per block: with Bm = R1 R2^T and
// C = [I, Bm; Bm^T, I] = chol(C) chol(C)^T, chol(C) = [I, 0; Bm^T, L22] and
// L22 = chol(I - Bm^T Bm), it writes rows r..2r-1 of the 2r x 2r blocks
//   X = chol(C) - I    = [Bm^T, L22 - I]
//   G = I - chol(C)^-1 = [L22^-1 Bm^T, I - L22^-1]
// (their first r rows are zero and are neither written nor read), the sum of
// log diag(L22) to *part, and sets *info to 1 when I - Bm^T Bm is not
// positive definite.
constexpr int CAP_MAX = 80;
struct CapDesc {
    const double* R1; const double* R2; int ldr;   // R_V and R_U, r x r upper (zero below)
    int r;
    double* X; double* G; int ld;                  // ld >= 2r
    double* part;
    int* info;
};
void capacitance(Ctx& c, const std::vector<CapDesc>& ds);

// Batched 2-D copy / transpose / elementwise helpers.
struct Copy2D {
    const double* src; int lds;
    double* dst; int ldd;
    int m, n;
    double alpha;   // dst = alpha * src (transposed: dst(j,i) = alpha * src(i,j))
    uint8_t es = 0, ed = 0;   // src / dst stored in FP32 (not with transpose)
};
void copy2d(Ctx& c, const std::vector<Copy2D>& probs, bool transpose = false);
// dst(i,j) = a * dst(i,j) + b * src(i,j) for i < m, j < n
struct Axpby2D {
    const double* src; int lds;
    double* dst; int ldd;
    int m, n;
    double a, b;
};
void axpby2d(Ctx& c, const std::vector<Axpby2D>& probs);
// out[i] = x[i] + y[i] for i < n (contiguous vectors; out may alias neither)
struct Add2 {
    const double* x;
    const double* y;
    double* out;
    int64_t n;
};
void add2(Ctx& c, const std::vector<Add2>& probs);
// d_out[i] = *src[i] for every pointer (one launch), e.g. per-handle flags or
// sums read back with one copy.
void gather_doubles(Ctx& c, const std::vector<const double*>& src, double* d_out);
void gather_ints(Ctx& c, const std::vector<const int*>& src, int* d_out);
// *p = 0 for every pointer (one launch).
void zero_ints(Ctx& c, const std::vector<int*>& ptrs);
// dst = identity (n x n) ; or dst(i,i) += v
struct DiagOp {
    double* A; int n; int ld;
    double v;
};
void set_identity(Ctx& c, const std::vector<DiagOp>& mats);   // A = v * I
void add_diag(Ctx& c, const std::vector<DiagOp>& mats);       // A += v * I

// out[j] = sum_i log(A_j(i,i)) for each matrix j (deterministic order).
void log_diag_sums(Ctx& c, const std::vector<SqMat>& mats, double* d_out);
// *outs[j] = sum_i log(A_j(i,i)), one output address per matrix.
void log_diag_sums(Ctx& c, const std::vector<SqMat>& mats, const std::vector<double*>& outs);
// out = sum of d_in[0..count) in a fixed order (one workgroup).
void sum_vector(Ctx& c, const double* d_in, int count, double* d_out);
struct SumJob {
    const double* in;
    int count;
    double* out;
};
void sum_vectors(Ctx& c, const std::vector<SumJob>& jobs);
// *out = dot(x[0..m), y[0..m)) for each job, in a fixed order (slices of
// 8192 elements, then the slices in sequence).
struct DotJob {
    const double* x;
    const double* y;
    int64_t m;
    double* out;
};
void dots_batch(Ctx& c, const std::vector<DotJob>& jobs);
// out[j] = dot(X[:,j], Y[:,j]) for j < ncols (as dots_batch).
void column_dots(Ctx& c, const double* X, int ldx, const double* Y, int ldy,
                 int m, int ncols, double* d_out);

}  // namespace kmh
