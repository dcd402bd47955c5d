// Derivative HODLR matrices of a Nystrom HODLR (constructor.jl:66-121) as
// term lists over shared panels, and the stochastic second-order
// quantities built from them.
#pragma once
#include <vector>

#include "hodlr.h"

namespace kmh {

// Panels of one derivative kernel: its leaves (same layout as H.leaves),
// dK(x, P) (n x r, ld n) and dK(P, P) (r x r).
struct DPanels {
    double* leaf = nullptr;
    double* kxp = nullptr;
    double* kpp = nullptr;
};
DPanels dpanels_make(Hodlr& H, const kmh_kernel& dk);
void dpanels_free(Hodlr& H, DPanels& d);

// One block term: sign * Panel_left[target] S^-1 M_1 S^-1 ... M_m S^-1
//                 Panel_right[source]^T z
// panels: 0 = K(x,P), 1..3 = the caller's derivative panels;
// mats (chain entries): 1..3 = the corresponding dK(P,P).
struct DTerm {
    double sign;
    int left, right;
    std::vector<int> chain;   // outer to inner
};

// out = (leaf part + block terms) z for z (n x k, ld n): `fwd` terms map the
// first child's rows to the second child's, `bwd` the reverse
// (HODLR/src/baseoverloads.jl:216-234, HODLR/src/utils.jl:76-90, 110-177).
void dterm_matvec(Hodlr& H, const double* leaf, const std::vector<const double*>& panels,
                  const std::vector<const double*>& mats, const std::vector<DTerm>& fwd,
                  const std::vector<DTerm>& bwd, const double* Z, double* out, int k);

void hodlr_stoch_fisher(Hodlr& H, const kmh_kernel* dks, int p, const double* saa_host, int64_t nv, double* F);
void hodlr_stoch_hessian(Hodlr& H, const kmh_kernel* d1, int p, const kmh_kernel* d2, const double* y_host,
                         const double* saa_host, int64_t nv, bool exact_second, double* Hout);

}  // namespace kmh
