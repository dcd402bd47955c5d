// Kernel-matrix assembly and ACA launchers.
#pragma once
#include "common.h"
#include "kernelfn.h"

namespace kmh {

// Dense block dst(i, j) = k(row point i, col point j); points are SoA.
struct EvalDesc {
    const double* rx; const double* ry;
    const double* cx; const double* cy;
    int m, n;
    double* dst; int ld;
    int tile0, pad;
    KDev k;   // the kernel of this block (requests in one launch may differ)
};
void eval_blocks(Ctx& c, const kmh_kernel& kern, std::vector<EvalDesc> ds);   // sets every k
void eval_blocks(Ctx& c, std::vector<EvalDesc> ds);                           // per-descriptor k

// One off-diagonal block for the ACA kernel: rows a..a+nro (second child),
// cols c..c+ncl (first child).  U (nro x cap) and V (ncl x cap) share the
// leading dimension ld.  Outputs: rank, status (0 ok, 1 start-row error,
// 2 capacity reached before convergence).
struct AcaDesc {
    const double* px; const double* py;   // the request's points (SoA)
    int a, nro, c, ncl;
    double* U;
    double* V;
    int ld, maxsz, cap, pad;
    int* rank;
    int* status;
    double rtol;
    KDev k;
};
size_t aca_shared_bytes(int nro, int ncl, int cap);
void aca_launch(Ctx& c, const std::vector<AcaDesc>& ds);

}  // namespace kmh
