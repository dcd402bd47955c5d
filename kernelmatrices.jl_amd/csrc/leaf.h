// Fused leaf stage of the symmetric factorization for leaves of at most
// LEAF_FUSED_MAX rows (symfact.jl:17-21): one workgroup evaluates a leaf of
// K from its points, factors it and keeps the inverses of its 16 x 16
// diagonal blocks; a second kernel applies L^{-1} to the leaf's rows of the
// right-hand panels.
#pragma once
#include "common.h"
#include "kernelfn.h"

namespace kmh {

constexpr int LEAF_FUSED_MAX = 256;

// Doubles of the diagonal-block inverses of an m-row leaf (m <= LEAF_FUSED_MAX).
inline size_t leaf_linv_size(int m) { return (size_t)256 * ((m + 15) / 16); }

// K(x_leaf, x_leaf) evaluated in registers, lower Cholesky factor to L
// (m x m, ld m, strict upper triangle zero), the inverses of the 16 x 16
// diagonal blocks of L to linv (leaf_linv_size(m) doubles), sum log diag(L)
// to *part; a non-positive or NaN pivot sets *info to 1.
struct LeafChol {
    const double* px; const double* py;   // the leaf's points (SoA)
    int m;
    double* L;
    double* linv;
    double* part;
    int* info;
    KDev k;
};
void leaf_chol(Ctx& c, const std::vector<LeafChol>& ls);

// B <- L^{-1} B for the m rows of one leaf (L and linv from leaf_chol),
// ncols columns with leading dimension ldb, FP64.
struct LeafSolve {
    const double* L;
    const double* linv;
    int m;
    double* B; int ldb; int ncols;
};
void leaf_solve(Ctx& c, const std::vector<LeafSolve>& ls);

}  // namespace kmh
