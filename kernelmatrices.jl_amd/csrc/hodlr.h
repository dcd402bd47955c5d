// HODLR matrix held on one device: tree, leaves, off-diagonal panels, factor.
#pragma once
#include <vector>

#include "common.h"

namespace kmh {

// Off-diagonal block of one level: rows a0..b1 (second child) x cols c0..d1
// (first child), 0-based half-open (indices.jl:40-51).
struct Block {
    int64_t a0, b1, c0, d1;
};

struct Hodlr {
    Ctx* ctx = nullptr;
    int64_t n = 0;
    int L = 0;                                   // number of off-diagonal levels
    std::vector<int64_t> leaf_off;               // nleaves + 1 boundaries
    std::vector<std::vector<Block>> blocks;      // [level][block]
    std::vector<std::vector<int>> ranks;         // [level][block]
    std::vector<int> rpad;                       // panel width per level
    std::vector<int64_t> coff;                   // first column of each level in P (L+1 entries)
    int64_t Rtot = 0;
    kmh_kernel kern{};
    kmh_build_opts opts{};

    DevBuf pts;                                  // x[0..n), y[0..n)
    DevBuf leaves;                               // leaf i at leaf_mat_off[i], ld = size
    std::vector<int64_t> leaf_mat_off;
    DevBuf P;                                    // n x Rtot, ld n: U/V (then Q) per level

    // Nystrom state (kept for the derivative HODLR)
    int nys_r = 0;
    std::vector<int64_t> nyind;
    DevBuf lm;                                   // landmark points, x[0..r), y[0..r)
    DevBuf kpp;                                  // chol(K(P,P)) lower, r x r
    DevBuf kxp;                                  // K(x, P), n x r, ld n
    DevBuf ynys;                                 // K(x, P) K(P,P)^{-1}, n x r, ld n

    // factor
    bool factorized = false;
    std::vector<std::vector<int64_t>> coff_blk;  // offset of block (2r x 2r) in X / G
    DevBuf X;                                    // chol(C) - I per block
    DevBuf G;                                    // I - chol(C)^{-1} per block
    double logdet = 0.0;

    bool nystrom() const { return opts.method == KMH_NYSTROM; }
    // U (rows a0..b1) and V (rows c0..d1) of block k at level l before
    // factorization, leading dimension n.  Nystrom blocks all share the two
    // panels K(x,P) and K(x,P)K(P,P)^{-1} (src/factorizations.jl:86-94);
    // ACA blocks live in P.
    const double* Uptr(int l, int k) const {
        const Block& b = blocks[l][k];
        return nystrom() ? kxp.as<double>() + b.a0 : P.as<double>() + (size_t)coff[l] * n + b.a0;
    }
    const double* Vptr(int l, int k) const {
        const Block& b = blocks[l][k];
        return nystrom() ? ynys.as<double>() + b.c0 : P.as<double>() + (size_t)coff[l] * n + b.c0;
    }
    int64_t leaf_size(int i) const { return leaf_off[i + 1] - leaf_off[i]; }
    int nleaves() const { return (int)leaf_off.size() - 1; }
    size_t device_bytes() const {
        return pts.bytes + leaves.bytes + P.bytes + lm.bytes + kpp.bytes + kxp.bytes + ynys.bytes + X.bytes + G.bytes;
    }
};

Hodlr* hodlr_build(Ctx& c, const double* pts, bool pts_on_device, int64_t n, int d, const kmh_kernel& k,
                   const kmh_build_opts& o);
void hodlr_factorize(Hodlr& H);
// y = K x (unfactorized: leaves + U V^T blocks; factorized: W W^T x)
void hodlr_matvec(Hodlr& H, const double* x, double* y, int nrhs);
// in-place on device data T (n x nrhs, ld n)
void hodlr_factor_apply(Hodlr& H, int op, double* T, int nrhs);
void hodlr_solve(Hodlr& H, double* T, int nrhs);
void hodlr_full_host(Hodlr& H, double* out);
void hodlr_block_uv_host(Hodlr& H, int level, int block, double* U, double* V);
void hodlr_stoch_gradient(Hodlr& H, const kmh_kernel* dks, int p, const double* y_host, const double* saa_host,
                          int64_t nv, double* grad);

}  // namespace kmh
