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
// column dots t' (dK t); the solve term y' K^-1 dK K^-1 y.  Gradients of
// several handles run together: each step is one grouped launch across them,
// and all dot products are read back with one synchronization.
#include <algorithm>

#include "assemble.h"
#include "hodlr.h"
#include "la.h"

namespace kmh {

// One handle's part of a batched stochastic gradient.
struct GradWork {
    Hodlr* H;
    const kmh_kernel* dks;
    int p;
    int64_t nv;
    bool profile;
    double* grad;
    int k;            // 1 + nv right-hand sides
    double* Z;        // [K^-1 y | W^-T v_1 .. v_nv]   (n x k)
    double* O;        // dK Z                          (n x k)
    double* yd;       // y on the device
    double* dleaf;    // derivative leaves
    double* dkxp;     // dK(x, P)
    double* Sj;       // dK(P, P)
    double* dots;     // per parameter: [y' K^-1 dK K^-1 y, t_1' dK t_1, ...], then y' K^-1 y
};

// out = dK z for each work item's Z (n x k, ld n), given its derivative
// leaves and panel; every step is one grouped launch across the items.
//
// The products Kxp[T]' z and dKxp[T]' z over the rows T of every tree node
// are built once: leaves by one product, parents as the sums of their
// children.  The per-block coefficients (as DBlock_mul / DBlock_mul_t) are
// then summed down the tree, so each leaf's rows receive all levels' terms
// in one product:
//   out[leaf] = dleaf z[leaf] + Kxp[leaf] kappa(leaf) + dKxp[leaf] alpha(leaf).
// Per block (global index bi) the coefficient buffer holds four r x k slots
// [p_cd | q_cd | p_ab | q_ab]: the Kxp and dKxp products of the first child
// (rows c0..d1) and of the second child (rows a0..b1), then E.  Block t of
// level l spans node t of level l + 1, whose own slot pair lives in block
// t / 2 of level l + 1 (first pair for even t, second for odd t).
static void deriv_matvec_batch(Ctx& c, const std::vector<GradWork*>& ws) {
    std::vector<GemmProb> g;
    for (GradWork* w : ws) {
        Hodlr& H = *w->H;
        const int64_t n = H.n;
        for (int i = 0; i < H.nleaves(); ++i) {
            const int m = (int)H.leaf_size(i);
            const int64_t s = H.leaf_off[i];
            g.push_back({m, w->k, m, w->dleaf + H.leaf_mat_off[i], m, w->Z + s, (int)n, w->O + s, (int)n, 1.0, 0.0, 0});
        }
    }
    gemm(c, false, false, g);
    std::vector<double*> Cs(ws.size(), nullptr);
    std::vector<std::vector<int>> b0s(ws.size());   // first global block index of each level
    int maxL = 0;
    std::vector<GemmProb> tn;
    for (size_t wi = 0; wi < ws.size(); ++wi) {
        GradWork* w = ws[wi];
        Hodlr& H = *w->H;
        if (H.L == 0) continue;
        maxL = std::max(maxL, H.L);
        const int64_t n = H.n;
        const int r = H.nys_r, k = w->k;
        const size_t cw = (size_t)4 * k, ew = (size_t)2 * k;
        const int nb = H.nblocks();
        Cs[wi] = static_cast<double*>(c.scratch((size_t)r * (cw + ew) * nb * sizeof(double)));
        b0s[wi].assign(H.L + 1, 0);
        for (int l = 0; l < H.L; ++l) b0s[wi][l + 1] = b0s[wi][l] + (int)H.blocks[l].size();
        // leaf products into the level-0 blocks
        const double* Kxp = H.kxp.as<double>();
        for (size_t q = 0; q < H.blocks[0].size(); ++q) {
            const Block& b = H.blocks[0][q];
            double* cb = Cs[wi] + (size_t)r * cw * q;
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
            tn.push_back({r, k, h1, Kxp + b.c0, (int)n, w->Z + b.c0, (int)n, cb, r, 1.0, 0.0, 0});
            tn.push_back({r, k, h1, w->dkxp + b.c0, (int)n, w->Z + b.c0, (int)n, cb + (size_t)r * k, r, 1.0, 0.0, 0});
            tn.push_back({r, k, h2, Kxp + b.a0, (int)n, w->Z + b.a0, (int)n, cb + (size_t)2 * r * k, r, 1.0, 0.0, 0});
            tn.push_back({r, k, h2, w->dkxp + b.a0, (int)n, w->Z + b.a0, (int)n, cb + (size_t)3 * r * k, r, 1.0, 0.0, 0});
        }
    }
    gemm(c, true, false, tn);
    // up-sweep: node products of level l + 2 from the two child pairs of level l + 1
    for (int l = 0; l + 1 < maxL; ++l) {
        std::vector<Add2> up;
        for (size_t wi = 0; wi < ws.size(); ++wi) {
            Hodlr& H = *ws[wi]->H;
            if (l + 1 >= H.L) continue;
            const int r = H.nys_r, k = ws[wi]->k;
            const size_t cw = (size_t)4 * k, pair = (size_t)2 * r * k;
            for (size_t t = 0; t < H.blocks[l].size(); ++t) {
                const double* src = Cs[wi] + (size_t)r * cw * (b0s[wi][l] + t);
                double* dst = Cs[wi] + (size_t)r * cw * (b0s[wi][l + 1] + t / 2) + (t % 2) * pair;
                up.push_back({src, src + pair, dst, (int64_t)pair});
            }
        }
        add2(c, up);
    }
    // coefficients: a = S^-1 p, kappa = S^-1 q - S^-1 Sj S^-1 p  (HODLR/src/utils.jl:67-90)
    std::vector<GemmProb> se;
    std::vector<TrsmProb> sc, se_tr;
    std::vector<Axpby2D> ax;
    for (size_t wi = 0; wi < ws.size(); ++wi) {
        GradWork* w = ws[wi];
        Hodlr& H = *w->H;
        if (H.L == 0) continue;
        const int r = H.nys_r, k = w->k;
        const double* S = H.kpp.as<double>();
        const int nb = H.nblocks();
        const size_t cw = (size_t)4 * k, ew = (size_t)2 * k;
        double* C = Cs[wi];
        double* E = C + (size_t)r * cw * nb;
        for (int bi = 0; bi < nb; ++bi) {
            double* cb = C + (size_t)r * cw * bi;
            double* eb = E + (size_t)r * ew * bi;
            se.push_back({r, k, r, w->Sj, r, cb, r, eb, r, 1.0, 0.0, 0});
            se.push_back({r, k, r, w->Sj, r, cb + (size_t)2 * r * k, r, eb + (size_t)r * k, r, 1.0, 0.0, 0});
            ax.push_back({eb, r, cb + (size_t)r * k, r, r, k, 1.0, -1.0});
            ax.push_back({eb + (size_t)r * k, r, cb + (size_t)3 * r * k, r, r, k, 1.0, -1.0});
        }
        sc.push_back({S, r, r, C, (int)(cw * nb), r});
        se_tr.push_back({S, r, r, E, (int)(ew * nb), r});
    }
    trsm_left(c, false, sc);
    trsm_left(c, true, sc);
    gemm(c, false, false, se);
    trsm_left(c, false, se_tr);
    trsm_left(c, true, se_tr);
    axpby2d(c, ax);
    // Now a block's first pair [a_cd | kappa_cd] belongs to its second child
    // (rows a0..b1, DBlock_mul) and the second pair [a_ab | kappa_ab] to its
    // first child (rows c0..d1, DBlock_mul_t).  Down-sweep: a node's pair is
    // added to both pairs of the block one level below that it spans.
    for (int l = maxL - 1; l >= 1; --l) {
        std::vector<Axpby2D> dn;
        for (size_t wi = 0; wi < ws.size(); ++wi) {
            Hodlr& H = *ws[wi]->H;
            if (l >= H.L) continue;
            const int r = H.nys_r, k = ws[wi]->k;
            const size_t cw = (size_t)4 * k, pair = (size_t)2 * r * k;
            for (size_t t = 0; t < H.blocks[l - 1].size(); ++t) {
                // node t of level l is the first child (even t) or the second child (odd t) of block t / 2
                const double* np = Cs[wi] + (size_t)r * cw * (b0s[wi][l] + t / 2) + (t % 2 ? 0 : pair);
                double* cb = Cs[wi] + (size_t)r * cw * (b0s[wi][l - 1] + t);
                dn.push_back({np, (int)pair, cb, (int)pair, (int)pair, 1, 1.0, 1.0});
                dn.push_back({np, (int)pair, cb + pair, (int)pair, (int)pair, 1, 1.0, 1.0});
            }
        }
        axpby2d(c, dn);
    }
    // one product per leaf for all levels' terms
    std::vector<GemmProb> g1, g2;
    for (size_t wi = 0; wi < ws.size(); ++wi) {
        GradWork* w = ws[wi];
        Hodlr& H = *w->H;
        if (H.L == 0) continue;
        const int64_t n = H.n;
        const int r = H.nys_r, k = w->k;
        const double* Kxp = H.kxp.as<double>();
        const size_t cw = (size_t)4 * k;
        for (size_t q = 0; q < H.blocks[0].size(); ++q) {
            const Block& b = H.blocks[0][q];
            const double* cb = Cs[wi] + (size_t)r * cw * q;
            const int h1 = (int)(b.d1 - b.c0), h2 = (int)(b.b1 - b.a0);
            // second child (rows a0..b1): [a_cd | kappa_cd]
            g1.push_back({h2, k, r, Kxp + b.a0, (int)n, cb + (size_t)r * k, r, w->O + b.a0, (int)n, 1.0, 1.0, 0});
            g2.push_back({h2, k, r, w->dkxp + b.a0, (int)n, cb, r, w->O + b.a0, (int)n, 1.0, 1.0, 0});
            // first child (rows c0..d1): [a_ab | kappa_ab]
            g1.push_back({h1, k, r, Kxp + b.c0, (int)n, cb + (size_t)3 * r * k, r, w->O + b.c0, (int)n, 1.0, 1.0, 0});
            g2.push_back({h1, k, r, w->dkxp + b.c0, (int)n, cb + (size_t)2 * r * k, r, w->O + b.c0, (int)n, 1.0, 1.0, 0});
        }
    }
    gemm(c, false, false, g1);
    gemm(c, false, false, g2);
    for (double* C : Cs)
        if (C) c.scratch_free(C);
}

void check_grad_job(const GradJob& jb) {
    Hodlr& H = *jb.H;
    KMH_CHECK(H.opts.method == KMH_NYSTROM, KMH_E_BAD_ARG, "This is only valid for Nystrom-block matrices.");
    KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized for this to work.");
    KMH_CHECK(H.L == 0 || H.nys_r > 0, KMH_E_STATE, "missing Nystrom state");
    KMH_CHECK(jb.nv >= 1 && (jb.saa_host || jb.saa_dev) && (jb.y_host || jb.y_dev), KMH_E_BAD_ARG, "null argument");
    for (int j = 0; j < jb.p; ++j) {
        KMH_CHECK(jb.dks[j].family == H.kern.family, KMH_E_BAD_ARG, "derivative kernel family differs from the kernel");
        KMH_CHECK(jb.dks[j].deriv != KMH_D0, KMH_E_BAD_ARG, "derivative kernels need a derivative code");
    }
}

void hodlr_stoch_gradient_batch(Ctx& c, const std::vector<GradJob>& jobs) {
    if (jobs.empty()) return;
    std::vector<GradWork> ws(jobs.size());
    int maxp = 0;
    for (const GradJob& jb : jobs) {
        check_grad_job(jb);
        maxp = std::max(maxp, jb.p);
    }
    // buffers and right-hand sides
    std::vector<Hodlr*> need_solve, wt;
    std::vector<double*> need_T, wt_T;
    std::vector<int> need_k, wt_k;
    for (size_t i = 0; i < jobs.size(); ++i) {
        const GradJob& jb = jobs[i];
        GradWork& w = ws[i];
        Hodlr& H = *jb.H;
        const int64_t n = H.n;
        w.H = jb.H;
        w.dks = jb.dks;
        w.p = jb.p;
        w.nv = jb.nv;
        w.profile = jb.profile;
        w.grad = jb.grad;
        w.k = (int)(1 + jb.nv);
        const size_t nk = (size_t)n * w.k;
        const int r = H.nys_r;
        w.Z = static_cast<double*>(c.scratch((2 * nk + (size_t)n + (size_t)(jb.p * w.k + 1)) * sizeof(double)));
        w.O = w.Z + nk;
        w.yd = w.O + nk;
        w.dots = w.yd + n;
        w.dleaf = static_cast<double*>(c.scratch(H.leaves.bytes));
        w.dkxp = r ? static_cast<double*>(c.scratch((size_t)n * r * sizeof(double))) : nullptr;
        w.Sj = r ? static_cast<double*>(c.scratch((size_t)r * r * sizeof(double))) : nullptr;
        if (jb.y_dev) copy2d(c, {{jb.y_dev, (int)n, w.yd, (int)n, (int)n, 1, 1.0}});
        else KMH_HIP(hipMemcpyAsync(w.yd, jb.y_host, n * sizeof(double), hipMemcpyHostToDevice, c.q()));
        if (jb.kinv_y_dev) {
            copy2d(c, {{jb.kinv_y_dev, (int)n, w.Z, (int)n, (int)n, 1, 1.0}});
        } else {
            copy2d(c, {{w.yd, (int)n, w.Z, (int)n, (int)n, 1, 1.0}});
            need_solve.push_back(jb.H);
            need_T.push_back(w.Z);
            need_k.push_back(1);
        }
        if (jb.saa_dev)
            copy2d(c, {{jb.saa_dev, (int)n, w.Z + n, (int)n, (int)n, (int)jb.nv, 1.0}});
        else
            KMH_HIP(hipMemcpyAsync(w.Z + n, jb.saa_host, (size_t)n * jb.nv * sizeof(double), hipMemcpyHostToDevice,
                                   c.q()));
        wt.push_back(jb.H);
        wt_T.push_back(w.Z + n);
        wt_k.push_back((int)jb.nv);
    }
    if (!need_solve.empty()) hodlr_solve_batch(c, need_solve, need_T, need_k);   // K^-1 y
    hodlr_factor_apply_batch(c, KMH_WTINV, wt, wt_T, wt_k);                    // W^-T v per SAA vector
    std::vector<DotJob> dj;
    for (auto& w : ws)
        if (w.profile) dj.push_back({w.yd, w.Z, w.H->n, w.dots + (size_t)w.p * w.k});   // y' K^-1 y
    for (int j = 0; j < maxp; ++j) {
        std::vector<GradWork*> act;
        std::vector<EvalDesc> ev;
        for (auto& w : ws) {
            if (w.p <= j) continue;
            act.push_back(&w);
            Hodlr& H = *w.H;
            const int64_t n = H.n;
            const int r = H.nys_r;
            const KDev kd = to_kdev(w.dks[j]);
            const double* px = H.pts.as<double>();
            const double* py = px + n;
            for (int i = 0; i < H.nleaves(); ++i) {
                const int64_t s = H.leaf_off[i];
                const int m = (int)H.leaf_size(i);
                ev.push_back({px + s, py + s, px + s, py + s, m, m, w.dleaf + H.leaf_mat_off[i], m, 0, 0, kd});
            }
            if (r) {
                const double* lx = H.lm.as<double>();
                const double* ly = lx + r;
                ev.push_back({px, py, lx, ly, (int)n, r, w.dkxp, (int)n, 0, 0, kd});
                ev.push_back({lx, ly, lx, ly, r, r, w.Sj, r, 0, 0, kd});
            }
        }
        eval_blocks(c, ev);
        deriv_matvec_batch(c, act);
        // solve term y' K^-1 dK K^-1 y (stochasticgradient.jl:21) as
        // (K^-1 y)' (dK K^-1 y): the factored inverse W^-T W^-1 is symmetric,
        // so the second solve of the reference is not needed.  Trace term:
        // t' dK t for each SAA column (stochasticgradient.jl:2-10, 24).
        // Queued before the next parameter reuses O.
        for (GradWork* w : act) {
            const int64_t n = w->H->n;
            double* d = w->dots + (size_t)j * w->k;
            for (int v = 0; v < w->k; ++v) dj.push_back({w->Z + (size_t)v * n, w->O + (size_t)v * n, n, d + v});
        }
        dots_batch(c, dj);
        dj.clear();
    }
    dots_batch(c, dj);   // y' K^-1 y of profile jobs when no parameter ran
    std::vector<std::vector<double>> hd(ws.size());
    for (size_t i = 0; i < ws.size(); ++i) {
        hd[i].resize((size_t)ws[i].p * ws[i].k + 1);
        KMH_HIP(hipMemcpyAsync(hd[i].data(), ws[i].dots, hd[i].size() * sizeof(double), hipMemcpyDeviceToHost,
                               c.q()));
    }
    c.sync();
    for (size_t i = 0; i < ws.size(); ++i) {
        GradWork& w = ws[i];
        const double yky = w.profile ? hd[i][(size_t)w.p * w.k] : 0.0;
        for (int j = 0; j < w.p; ++j) {
            const double* d = hd[i].data() + (size_t)j * w.k;
            double tr = 0.0;
            for (int v = 1; v < w.k; ++v) tr += d[v];
            tr /= (double)w.nv;
            // stochasticgradient.jl:24 ; profile form :60-62
            w.grad[j] = w.profile ? 0.5 * (tr - (double)w.H->n * d[0] / yky) : 0.5 * (tr - d[0]);
        }
        if (w.Sj) c.scratch_free(w.Sj);
        if (w.dkxp) c.scratch_free(w.dkxp);
        c.scratch_free(w.dleaf);
        c.scratch_free(w.Z);
    }
}

void hodlr_stoch_gradient(Hodlr& H, const kmh_kernel* dks, int p, const double* y, const double* saa, int64_t nv,
                          double* grad, bool profile, bool on_device) {
    GradJob jb{};
    jb.H = &H;
    jb.kinv_y_dev = nullptr;
    jb.dks = dks;
    jb.p = p;
    jb.y_host = on_device ? nullptr : y;
    jb.y_dev = on_device ? y : nullptr;
    jb.saa_host = on_device ? nullptr : saa;
    jb.saa_dev = on_device ? saa : nullptr;
    jb.nv = nv;
    jb.grad = grad;
    jb.profile = profile;
    hodlr_stoch_gradient_batch(*H.ctx, {jb});
}

}  // namespace kmh
