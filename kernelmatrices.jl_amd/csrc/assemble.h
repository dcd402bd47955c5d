// Kernel-matrix assembly and ACA launchers.
#pragma once
#include "common.h"
#include "kernelfn.h"

#include <memory>

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

// One off-diagonal block for ACA: rows a..a+nro of the row points, cols
// c..c+ncl of the column points (a symmetric HODLR passes its points for
// both).  U (nro x cap) and V (ncl x cap) share the leading dimension ld.
// Outputs: rank, status (0 ok, 1 start-row error, 2 capacity reached before
// convergence).
struct AcaDesc {
    const double* px; const double* py;   // row points (SoA)
    const double* qx; const double* qy;   // column points (SoA)
    int a, nro, c, ncl;
    double* U;
    double* V;
    int ld, maxsz, cap, pad;
    int* rank;
    int* status;
    double rtol;
    KDev k;
};

// ACA of a set of blocks.  The constructor queues the first cross of every
// block; run() steps all blocks until each has converged or stopped (status
// written to the descriptors' rank/status).  A block that stopped at its
// capacity (status 2) can be given wider panels holding the columns found so
// far with widen(), and the next run() continues it where it stopped.
class AcaRun {
public:
    AcaRun(Ctx& c, const std::vector<AcaDesc>& ds);
    ~AcaRun();
    AcaRun(const AcaRun&) = delete;
    AcaRun& operator=(const AcaRun&) = delete;
    void run();
    void widen(int block, double* U, double* V, int cap);
private:
    struct Impl;
    std::unique_ptr<Impl> m_;
};

}  // namespace kmh
