// Stochastic gradient of the HODLR Gaussian log-likelihood
// (stochasticgradient.jl:2-45) with the exact derivative of the Nystrom
// HODLR matrix (constructor.jl:66-95, HODLR/src/utils.jl:67-90).
//
// Every Nystrom block shares one landmark set P, so the derivative blocks
//   K1pd S^-1 Kp2 - K1p S^-1 Sj S^-1 Kp2 + K1p S^-1 Kp2d
// are built from two n x r panels, Kxp = K(x, P) and dKxp = dK(x, P), and the
// r x r matrices S = chol(K(P,P)) and Sj = dK(P,P).  For a block with source
// rows N and target rows T the product is
//   out_T += Kxp[T] (S^-1 dKxp[N]' z - S^-1 Sj S^-1 Kxp[N]' z) + dKxp[T] S^-1 Kxp[N]' z,
// evaluated for all SAA vectors at once (the reference loops them one by one).
// The trace term uses t = W^-T v for every Rademacher column v, then the
// column dots t' (dK t); the solve term y' K^-1 dK K^-1 y.
#include <algorithm>

#include "assemble.h"
#include "hodlr.h"
#include "la.h"

namespace kmh {

// out = dK z for z (n x k, ld n) given the derivative leaves dleaf and panel dkxp
static void deriv_matvec(Hodlr& H, const double* dleaf, const double* dkxp, const double* Sj, const double* Z,
                         double* out, int k) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    const int r = H.nys_r;
    const double* Kxp = H.kxp.as<double>();
    const double* S = H.kpp.as<double>();
    std::vector<GemmProb> g;
    for (int i = 0; i < H.nleaves(); ++i) {
        const int m = (int)H.leaf_size(i);
        const int64_t s = H.leaf_off[i];
        g.push_back({m, k, m, dleaf + H.leaf_mat_off[i], m, Z + s, (int)n, out + s, (int)n, 1.0, 0.0, 0});
    }
    gemm(c, false, false, g);
    if (H.L == 0) return;
    int nb = 0;
    for (int l = 0; l < H.L; ++l) nb += (int)H.blocks[l].size();
    // coefficient buffer per block: [p_cd | q_cd | p_ab | q_ab] (r x 4k), then E = [e_cd | e_ab] (r x 2k)
    const size_t cw = (size_t)4 * k, ew = (size_t)2 * k;
    double* C = static_cast<double*>(c.scratch((size_t)r * (cw + ew) * nb * sizeof(double)));
    double* E = C + (size_t)r * cw * nb;
    std::vector<GemmProb> tn, se;
    int bi = 0;
    for (int l = 0; l < H.L; ++l)
        for (const auto& b : H.blocks[l]) {
            double* cb = C + (size_t)r * cw * bi;
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
            tn.push_back({r, k, h1, Kxp + b.c0, (int)n, Z + b.c0, (int)n, cb, r, 1.0, 0.0, 0});
            tn.push_back({r, k, h1, dkxp + b.c0, (int)n, Z + b.c0, (int)n, cb + (size_t)r * k, r, 1.0, 0.0, 0});
            tn.push_back({r, k, h2, Kxp + b.a0, (int)n, Z + b.a0, (int)n, cb + (size_t)2 * r * k, r, 1.0, 0.0, 0});
            tn.push_back({r, k, h2, dkxp + b.a0, (int)n, Z + b.a0, (int)n, cb + (size_t)3 * r * k, r, 1.0, 0.0, 0});
            ++bi;
        }
    gemm(c, true, false, tn);
    // S^-1 on every coefficient column
    const int ncols = (int)(cw * nb);
    trsm_left(c, false, {{S, r, r, C, ncols, r}});
    trsm_left(c, true, {{S, r, r, C, ncols, r}});
    // E = S^-1 Sj a  for a = S^-1 p of both children
    bi = 0;
    for (int l = 0; l < H.L; ++l)
        for (size_t q = 0; q < H.blocks[l].size(); ++q) {
            double* cb = C + (size_t)r * cw * bi;
            double* eb = E + (size_t)r * ew * bi;
            se.push_back({r, k, r, Sj, r, cb, r, eb, r, 1.0, 0.0, 0});
            se.push_back({r, k, r, Sj, r, cb + (size_t)2 * r * k, r, eb + (size_t)r * k, r, 1.0, 0.0, 0});
            ++bi;
        }
    gemm(c, false, false, se);
    trsm_left(c, false, {{S, r, r, E, (int)(ew * nb), r}});
    trsm_left(c, true, {{S, r, r, E, (int)(ew * nb), r}});
    // b <- b - e  (coefficients of the Kxp terms)
    std::vector<Axpby2D> ax;
    bi = 0;
    for (int l = 0; l < H.L; ++l)
        for (size_t q = 0; q < H.blocks[l].size(); ++q) {
            double* cb = C + (size_t)r * cw * bi;
            double* eb = E + (size_t)r * ew * bi;
            ax.push_back({eb, r, cb + (size_t)r * k, r, r, k, 1.0, -1.0});
            ax.push_back({eb + (size_t)r * k, r, cb + (size_t)3 * r * k, r, r, k, 1.0, -1.0});
            ++bi;
        }
    axpby2d(c, ax);
    // accumulate, level by level (targets of one level are disjoint)
    bi = 0;
    for (int l = 0; l < H.L; ++l) {
        std::vector<GemmProb> g1, g2;
        for (const auto& b : H.blocks[l]) {
            const double* cb = C + (size_t)r * cw * bi;
            const double* a_cd = cb;
            const double* k_cd = cb + (size_t)r * k;       // S^-1 q_cd - S^-1 Sj S^-1 p_cd
            const double* a_ab = cb + (size_t)2 * r * k;
            const double* k_ab = cb + (size_t)3 * r * k;
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
            // DBlock_mul: rows a0..b1 from the first child's source
            g1.push_back({h2, k, r, Kxp + b.a0, (int)n, k_cd, r, out + b.a0, (int)n, 1.0, 1.0, 0});
            g2.push_back({h2, k, r, dkxp + b.a0, (int)n, a_cd, r, out + b.a0, (int)n, 1.0, 1.0, 0});
            // DBlock_mul_t: rows c0..d1 from the second child's source
            g1.push_back({h1, k, r, Kxp + b.c0, (int)n, k_ab, r, out + b.c0, (int)n, 1.0, 1.0, 0});
            g2.push_back({h1, k, r, dkxp + b.c0, (int)n, a_ab, r, out + b.c0, (int)n, 1.0, 1.0, 0});
            ++bi;
        }
        gemm(c, false, false, g1);
        gemm(c, false, false, g2);
    }
    c.scratch_free(C);
}

void hodlr_stoch_gradient(Hodlr& H, const kmh_kernel* dks, int p, const double* y_host, const double* saa_host,
                          int64_t nv, double* grad, bool profile) {
    KMH_CHECK(H.opts.method == KMH_NYSTROM, KMH_E_BAD_ARG, "This is only valid for Nystrom-block matrices.");
    KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized for this to work.");
    KMH_CHECK(H.L == 0 || H.nys_r > 0, KMH_E_STATE, "missing Nystrom state");
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    const int k = (int)(1 + nv);
    for (int j = 0; j < p; ++j) {
        KMH_CHECK(dks[j].family == H.kern.family, KMH_E_BAD_ARG, "derivative kernel family differs from the kernel");
        KMH_CHECK(dks[j].deriv != KMH_D0, KMH_E_BAD_ARG, "derivative kernels need a derivative code");
    }
    const size_t nk = (size_t)n * k;
    double* Z = static_cast<double*>(c.scratch(2 * nk * sizeof(double) + (size_t)(k + 1) * sizeof(double)));
    double* O = Z + nk;
    double* dots = O + nk;
    KMH_HIP(hipMemcpyAsync(Z, y_host, n * sizeof(double), hipMemcpyHostToDevice, c.q()));
    KMH_HIP(hipMemcpyAsync(Z + n, saa_host, (size_t)n * nv * sizeof(double), hipMemcpyHostToDevice, c.q()));
    hodlr_solve(H, Z, 1);                                  // K^-1 y
    hodlr_factor_apply(H, KMH_WTINV, Z + n, (int)nv);      // W^-T v for each SAA vector
    double yky = 0.0;                                      // y' K^-1 y (profile form only)
    if (profile) {
        double* yd = static_cast<double*>(c.scratch((size_t)(n + 1) * sizeof(double)));
        KMH_HIP(hipMemcpyAsync(yd, y_host, n * sizeof(double), hipMemcpyHostToDevice, c.q()));
        column_dots(c, yd, (int)n, Z, (int)n, (int)n, 1, yd + n);
        KMH_HIP(hipMemcpyAsync(&yky, yd + n, sizeof(double), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        c.scratch_free(yd);
    }
    const int r = H.nys_r;
    double* dleaf = static_cast<double*>(c.scratch(H.leaves.bytes));
    double* dkxp = r ? static_cast<double*>(c.scratch((size_t)n * r * sizeof(double))) : nullptr;
    double* Sj = r ? static_cast<double*>(c.scratch((size_t)r * r * sizeof(double))) : nullptr;
    double* prv = static_cast<double*>(c.scratch(n * sizeof(double)));
    const double* px = H.pts.as<double>();
    const double* py = px + n;
    std::vector<double> hd(k + 1);
    for (int j = 0; j < p; ++j) {
        std::vector<EvalDesc> ev;
        for (int i = 0; i < H.nleaves(); ++i) {
            const int64_t s = H.leaf_off[i];
            const int m = (int)H.leaf_size(i);
            ev.push_back({px + s, py + s, px + s, py + s, m, m, dleaf + H.leaf_mat_off[i], m, 0, 0});
        }
        if (r) {
            const double* lx = H.lm.as<double>();
            const double* ly = lx + r;
            ev.push_back({px, py, lx, ly, (int)n, r, dkxp, (int)n, 0, 0});
            ev.push_back({lx, ly, lx, ly, r, r, Sj, r, 0, 0});
        }
        eval_blocks(c, dks[j], ev);
        deriv_matvec(H, dleaf, dkxp, Sj, Z, O, k);
        // solve term: y' K^-1 dK K^-1 y   (stochasticgradient.jl:21)
        copy2d(c, {{O, (int)n, prv, (int)n, (int)n, 1, 1.0}});
        hodlr_solve(H, prv, 1);
        KMH_HIP(hipMemcpyAsync(O, prv, n * sizeof(double), hipMemcpyDeviceToDevice, c.q()));
        KMH_HIP(hipMemcpyAsync(prv, y_host, n * sizeof(double), hipMemcpyHostToDevice, c.q()));
        column_dots(c, prv, (int)n, O, (int)n, (int)n, 1, dots);
        // trace term: mean over SAA vectors of t' dK t  (stochasticgradient.jl:2-10, 24)
        column_dots(c, Z + n, (int)n, O + n, (int)n, (int)n, (int)nv, dots + 1);
        KMH_HIP(hipMemcpyAsync(hd.data(), dots, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        double tr = 0.0;
        for (int v = 1; v < k; ++v) tr += hd[v];
        tr /= (double)nv;
        // stochasticgradient.jl:24 ; profile form :60-62
        grad[j] = profile ? 0.5 * (tr - (double)n * hd[0] / yky) : 0.5 * (tr - hd[0]);
    }
    c.scratch_free(prv);
    if (Sj) c.scratch_free(Sj);
    if (dkxp) c.scratch_free(dkxp);
    c.scratch_free(dleaf);
    c.scratch_free(Z);
}

}  // namespace kmh
