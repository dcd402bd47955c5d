// Derivative HODLR products as term lists, and the stochastic expected
// Fisher matrix and Hessian of the HODLR Gaussian log-likelihood
// (stochastichessian.jl:3-147, HODLR/src/utils.jl:67-177).
//
// Every off-diagonal block of a Nystrom derivative HODLR is a sum of terms
//   sign * A[target] S^-1 M_1 S^-1 ... M_m S^-1 B[source]^T z,
// with A, B among the n x r panels K(x,P), dK_j(x,P), dK_k(x,P),
// d2K_jk(x,P) and M among dK_j(P,P), dK_k(P,P), d2K_jk(P,P); S is the
// Cholesky factor of K(P,P).  For all blocks of all levels at once the
// projections B[source]^T z, the S^-1 / M chains and the per-panel sums are
// grouped launches; the accumulation into the targets runs level by level.
#include <algorithm>
#include <map>

#include "assemble.h"
#include "deriv.h"
#include "la.h"

namespace kmh {

DPanels dpanels_make(Hodlr& H, const kmh_kernel& dk) {
    Ctx& c = *H.ctx;
    DPanels d;
    const int64_t n = H.n;
    const int r = H.nys_r;
    d.leaf = static_cast<double*>(c.scratch(H.leaves.bytes));
    std::vector<EvalDesc> ev;
    const double* px = H.pts.as<double>();
    const double* py = px + n;
    for (int i = 0; i < H.nleaves(); ++i) {
        const int64_t s = H.leaf_off[i];
        const int m = (int)H.leaf_size(i);
        ev.push_back({px + s, py + s, px + s, py + s, m, m, d.leaf + H.leaf_mat_off[i], m, 0, 0, {}});
    }
    if (r > 0) {
        d.kxp = static_cast<double*>(c.scratch((size_t)n * r * sizeof(double)));
        d.kpp = static_cast<double*>(c.scratch((size_t)r * r * sizeof(double)));
        const double* lx = H.lm.as<double>();
        const double* ly = lx + r;
        ev.push_back({px, py, lx, ly, (int)n, r, d.kxp, (int)n, 0, 0, {}});
        ev.push_back({lx, ly, lx, ly, r, r, d.kpp, r, 0, 0, {}});
    }
    eval_blocks(c, dk, ev);
    return d;
}

void dpanels_free(Hodlr& H, DPanels& d) {
    Ctx& c = *H.ctx;
    if (d.kpp) c.scratch_free(d.kpp);
    if (d.kxp) c.scratch_free(d.kxp);
    if (d.leaf) c.scratch_free(d.leaf);
    d = DPanels();
}

void dterm_matvec(Hodlr& H, const double* leaf, const std::vector<const double*>& panels,
                  const std::vector<const double*>& mats, const std::vector<DTerm>& fwd,
                  const std::vector<DTerm>& bwd, const double* Z, double* out, int k) {
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    std::vector<GemmProb> g;
    for (int i = 0; i < H.nleaves(); ++i) {
        const int m = (int)H.leaf_size(i);
        const int64_t s = H.leaf_off[i];
        g.push_back({m, k, m, leaf + H.leaf_mat_off[i], m, Z + s, (int)n, out + s, (int)n, 1.0, 0.0, 0});
    }
    gemm(c, false, false, g);
    if (H.L == 0) return;
    const int r = H.nys_r;
    const double* S = H.kpp.as<double>();
    const size_t rk = (size_t)r * k;
    // jobs: (level, block, direction)
    struct Job { int l; int b; int dir; };
    std::vector<Job> jobs;
    for (int l = 0; l < H.L; ++l)
        for (int b = 0; b < (int)H.blocks[l].size(); ++b)
            for (int dir = 0; dir < 2; ++dir) jobs.push_back({l, b, dir});
    // distinct right / left panels per direction
    auto panels_of = [](const std::vector<DTerm>& ts, bool left) {
        std::vector<int> v;
        for (const auto& t : ts) {
            int p = left ? t.left : t.right;
            if (std::find(v.begin(), v.end(), p) == v.end()) v.push_back(p);
        }
        return v;
    };
    const std::vector<int> rp[2] = {panels_of(fwd, false), panels_of(bwd, false)};
    const std::vector<int> lp[2] = {panels_of(fwd, true), panels_of(bwd, true)};
    const std::vector<DTerm>* terms[2] = {&fwd, &bwd};
    size_t nP = 0, nX = 0, nC = 0;
    for (const auto& j : jobs) {
        nP += rp[j.dir].size();
        nX += terms[j.dir]->size();
        nC += lp[j.dir].size();
    }
    double* buf = static_cast<double*>(c.scratch((nP + 2 * nX + nC) * rk * sizeof(double)));
    double* Pb = buf;
    double* Xb = Pb + nP * rk;
    double* Tb = Xb + nX * rk;
    double* Cb = Tb + nX * rk;
    std::vector<GemmProb> proj;
    std::vector<Copy2D> init;
    std::vector<TrsmProb> tr;
    size_t pi = 0, xi = 0;
    struct TermRef { double* X; double* T; const DTerm* t; };
    std::vector<TermRef> all;
    std::vector<size_t> job_x0(jobs.size()), job_c0(jobs.size());
    size_t ci = 0;
    for (size_t jb = 0; jb < jobs.size(); ++jb) {
        const auto& j = jobs[jb];
        const Block& b = H.blocks[j.l][j.b];
        const int64_t s0 = j.dir == 0 ? b.c0 : b.a0;
        const int hs = (int)(j.dir == 0 ? b.d1 - b.c0 : b.b1 - b.a0);
        std::map<int, double*> pmap;
        for (int p : rp[j.dir]) {
            double* dst = Pb + (pi++) * rk;
            pmap[p] = dst;
            proj.push_back({r, k, hs, panels[p] + s0, (int)n, Z + s0, (int)n, dst, r, 1.0, 0.0, 0});
        }
        job_x0[jb] = xi;
        for (const auto& t : *terms[j.dir]) {
            double* X = Xb + xi * rk;
            double* T = Tb + xi * rk;
            ++xi;
            init.push_back({pmap[t.right], r, X, r, r, k, 1.0});
            tr.push_back({S, r, r, X, k, r});
            all.push_back({X, T, &t});
        }
        job_c0[jb] = ci;
        ci += lp[j.dir].size();
    }
    gemm(c, true, false, proj);
    copy2d(c, init);
    trsm_left(c, false, tr);
    trsm_left(c, true, tr);
    // chains, innermost matrix first
    size_t maxlen = 0;
    for (const auto& t : fwd) maxlen = std::max(maxlen, t.chain.size());
    for (const auto& t : bwd) maxlen = std::max(maxlen, t.chain.size());
    for (size_t d = 1; d <= maxlen; ++d) {
        std::vector<GemmProb> mg;
        std::vector<TrsmProb> mt;
        std::vector<Copy2D> back;
        for (const auto& a : all) {
            const auto& ch = a.t->chain;
            if (ch.size() < d) continue;
            const double* M = mats[ch[ch.size() - d]];
            mg.push_back({r, k, r, M, r, a.X, r, a.T, r, 1.0, 0.0, 0});
            mt.push_back({S, r, r, a.T, k, r});
            back.push_back({a.T, r, a.X, r, r, k, 1.0});
        }
        gemm(c, false, false, mg);
        trsm_left(c, false, mt);
        trsm_left(c, true, mt);
        copy2d(c, back);
    }
    // combine per left panel: C_L = sum sign_t X_t, in term order.  The
    // i-th further term of every C goes into launch i, so no launch has two
    // problems updating the same C.
    std::vector<Copy2D> first;
    std::vector<std::vector<Axpby2D>> rest;
    for (size_t jb = 0; jb < jobs.size(); ++jb) {
        const auto& j = jobs[jb];
        const auto& L = lp[j.dir];
        for (size_t q = 0; q < L.size(); ++q) {
            double* C = Cb + (job_c0[jb] + q) * rk;
            size_t cnt = 0;
            size_t x = job_x0[jb];
            for (const auto& t : *terms[j.dir]) {
                const double* X = Xb + (x++) * rk;
                if (t.left != L[q]) continue;
                if (cnt == 0) {
                    first.push_back({X, r, C, r, r, k, t.sign});
                } else {
                    if (rest.size() < cnt) rest.resize(cnt);
                    rest[cnt - 1].push_back({X, r, C, r, r, k, 1.0, t.sign});
                }
                ++cnt;
            }
        }
    }
    copy2d(c, first);
    for (const auto& round : rest) axpby2d(c, round);
    // accumulate level by level, one launch per left panel (targets of one
    // level's forward and backward jobs are disjoint)
    for (int l = 0; l < H.L; ++l) {
        std::map<int, std::vector<GemmProb>> byp;
        for (size_t jb = 0; jb < jobs.size(); ++jb) {
            const auto& j = jobs[jb];
            if (j.l != l) continue;
            const Block& b = H.blocks[j.l][j.b];
            const int64_t t0 = j.dir == 0 ? b.a0 : b.c0;
            const int ht = (int)(j.dir == 0 ? b.b1 - b.a0 : b.d1 - b.c0);
            const auto& L = lp[j.dir];
            for (size_t q = 0; q < L.size(); ++q)
                byp[L[q]].push_back({ht, k, r, panels[L[q]] + t0, (int)n, Cb + (job_c0[jb] + q) * rk, r, out + t0,
                                     (int)n, 1.0, 1.0, 0});
        }
        for (auto& kv : byp) gemm(c, false, false, kv.second);
    }
    c.scratch_free(buf);
}

// first-derivative block terms (HODLR/src/utils.jl:76-90): panels 0 = K, 1 = dK;
// mats 1 = dK(P,P)
static const std::vector<DTerm> D1_FWD = {{+1.0, 1, 0, {}}, {-1.0, 0, 0, {1}}, {+1.0, 0, 1, {}}};
static const std::vector<DTerm> D1_BWD = {{+1.0, 0, 1, {}}, {-1.0, 0, 0, {1}}, {+1.0, 1, 0, {}}};

// second-derivative block terms: panels 0 = K, 1 = dK_j, 2 = dK_k, 3 = d2K_jk;
// mats 1 = dK_j(P,P), 2 = dK_k(P,P), 3 = d2K_jk(P,P).
// The reference's SDBlock_mul / SDBlock_mul_t (HODLR/src/utils.jl:110-155), term by term:
static const std::vector<DTerm> D2_REF_FWD = {
    {+1, 3, 0, {}},  {-1, 1, 0, {2}},    {+1, 1, 2, {}}, {+1, 2, 0, {1}}, {-1, 0, 0, {2, 1}}, {+1, 0, 0, {3}},
    {-1, 0, 0, {1, 2}}, {+1, 0, 1, {1}}, {+1, 2, 0, {}}, {-1, 0, 0, {2}}, {+1, 0, 3, {}}};
static const std::vector<DTerm> D2_REF_BWD = {
    {+1, 0, 3, {}},  {-1, 0, 1, {2}},    {+1, 2, 2, {}}, {+1, 0, 2, {1}}, {-1, 0, 0, {1, 2}}, {+1, 0, 0, {3}},
    {-1, 0, 0, {2, 1}}, {+1, 2, 0, {1}}, {+1, 1, 2, {}}, {-1, 1, 0, {2}}, {+1, 3, 0, {}}};
// The exact second derivative of K1 S^-1 K2 (the eleven terms of d_k d_j),
// and its transpose for the backward direction.
static const std::vector<DTerm> D2_EXACT_FWD = {
    {+1, 3, 0, {}},  {-1, 1, 0, {2}},    {+1, 1, 2, {}}, {-1, 2, 0, {1}}, {+1, 0, 0, {2, 1}}, {-1, 0, 0, {3}},
    {+1, 0, 0, {1, 2}}, {-1, 0, 2, {1}}, {+1, 2, 1, {}}, {-1, 0, 1, {2}}, {+1, 0, 3, {}}};
static const std::vector<DTerm> D2_EXACT_BWD = {
    {+1, 0, 3, {}},  {-1, 0, 1, {2}},    {+1, 2, 1, {}}, {-1, 0, 2, {1}}, {+1, 0, 0, {1, 2}}, {-1, 0, 0, {3}},
    {+1, 0, 0, {2, 1}}, {-1, 2, 0, {1}}, {+1, 1, 2, {}}, {-1, 1, 0, {2}}, {+1, 3, 0, {}}};

static void dk_matvec(Hodlr& H, const DPanels& d, const double* Z, double* out, int k) {
    dterm_matvec(H, d.leaf, {H.kxp.as<double>(), d.kxp}, {nullptr, d.kpp}, D1_FWD, D1_BWD, Z, out, k);
}

static void check_deriv_args(Hodlr& H, const kmh_kernel* dks, int p) {
    KMH_CHECK(H.opts.method == KMH_NYSTROM, KMH_E_BAD_ARG, "This is only valid for Nystrom-block matrices.");
    KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized.");
    KMH_CHECK(p >= 1, KMH_E_BAD_ARG, "You didn't supply the right number of gradient funs.");
    for (int j = 0; j < p; ++j) {
        check_kernel(dks[j], true);
        KMH_CHECK(dks[j].family == H.kern.family, KMH_E_BAD_ARG, "derivative kernel family differs from the kernel");
    }
}

// Z_j = W^{-1} dK_j W^{-T} V for each parameter (columns = SAA vectors)
static std::vector<double*> fisher_panels(Hodlr& H, const std::vector<DPanels>& dp, double* T, int nv) {
    Ctx& c = *H.ctx;
    std::vector<double*> Z;
    for (const auto& d : dp) {
        double* z = static_cast<double*>(c.scratch((size_t)H.n * nv * sizeof(double)));
        dk_matvec(H, d, T, z, nv);
        hodlr_factor_apply(H, KMH_WINV, z, nv);
        Z.push_back(z);
    }
    return Z;
}

static double sum_sq_cols(Ctx& c, const double* X, int64_t n, int nv, double* dots) {
    column_dots(c, X, (int)n, X, (int)n, (int)n, nv, dots);
    std::vector<double> h(nv);
    KMH_HIP(hipMemcpyAsync(h.data(), dots, nv * sizeof(double), hipMemcpyDeviceToHost, c.q()));
    c.sync();
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}

// stoch_fisher (stochastichessian.jl:91-115) on precomputed Z_j
static std::vector<double> fisher_from(Hodlr& H, const std::vector<double*>& Z, int nv, double* dots) {
    Ctx& c = *H.ctx;
    const int p = (int)Z.size();
    const int64_t n = H.n;
    std::vector<double> F((size_t)p * p, 0.0);
    for (int j = 0; j < p; ++j) F[j + j * p] = 0.5 * sum_sq_cols(c, Z[j], n, nv, dots) / nv;
    if (p > 1) {
        double* s = static_cast<double*>(c.scratch((size_t)n * nv * sizeof(double)));
        for (int j = 0; j < p; ++j)
            for (int k = j + 1; k < p; ++k) {
                copy2d(c, {{Z[j], (int)n, s, (int)n, (int)n, nv, 1.0}});
                axpby2d(c, {{Z[k], (int)n, s, (int)n, (int)n, nv, 1.0, 1.0}});
                const double tmp1 = 0.5 * sum_sq_cols(c, s, n, nv, dots) / nv;
                F[j + k * p] = 0.5 * (tmp1 - F[j + j * p] - F[k + k * p]);
                F[k + j * p] = F[j + k * p];
            }
        c.scratch_free(s);
    }
    return F;
}

void hodlr_stoch_fisher(Hodlr& H, const kmh_kernel* dks, int p, const double* saa_host, int64_t nv, double* Fout) {
    check_deriv_args(H, dks, p);
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    double* T = static_cast<double*>(c.scratch((size_t)n * nv * sizeof(double) + (size_t)(nv + 8) * sizeof(double)));
    double* dots = T + (size_t)n * nv;
    KMH_HIP(hipMemcpyAsync(T, saa_host, (size_t)n * nv * sizeof(double), hipMemcpyHostToDevice, c.q()));
    hodlr_factor_apply(H, KMH_WTINV, T, (int)nv);
    std::vector<DPanels> dp;
    for (int j = 0; j < p; ++j) dp.push_back(dpanels_make(H, dks[j]));
    auto Z = fisher_panels(H, dp, T, (int)nv);
    auto F = fisher_from(H, Z, (int)nv, dots);
    for (size_t i = 0; i < F.size(); ++i) Fout[i] = F[i];
    for (double* z : Z) c.scratch_free(z);
    for (auto& d : dp) dpanels_free(H, d);
    c.scratch_free(T);
}

void hodlr_stoch_hessian(Hodlr& H, const kmh_kernel* d1, int p, const kmh_kernel* d2, const double* y_host,
                         const double* saa_host, int64_t nv, bool exact_second, double* Hout) {
    check_deriv_args(H, d1, p);
    Ctx& c = *H.ctx;
    const int64_t n = H.n;
    const size_t nn = (size_t)n;
    // [y | V] -> [alpha = K^-1 y | T = W^-T V]
    const int k = (int)(1 + nv);
    double* Zb = static_cast<double*>(c.scratch((nn * k * 2 + nn * 2 + (size_t)(k + 8)) * sizeof(double)));
    double* O = Zb + nn * k;
    double* u = O + nn * k;
    double* w = u + nn;
    double* dots = w + nn;
    KMH_HIP(hipMemcpyAsync(Zb, y_host, nn * sizeof(double), hipMemcpyHostToDevice, c.q()));
    KMH_HIP(hipMemcpyAsync(Zb + nn, saa_host, nn * nv * sizeof(double), hipMemcpyHostToDevice, c.q()));
    hodlr_solve(H, Zb, 1);
    hodlr_factor_apply(H, KMH_WTINV, Zb + nn, (int)nv);
    double* alpha = Zb;
    double* T = Zb + nn;
    std::vector<DPanels> dp;
    for (int j = 0; j < p; ++j) dp.push_back(dpanels_make(H, d1[j]));
    // Out = -Fisher (stochastichessian.jl:130)
    auto Z = fisher_panels(H, dp, T, (int)nv);
    auto F = fisher_from(H, Z, (int)nv, dots);
    for (double* z : Z) c.scratch_free(z);
    std::vector<double> Hm((size_t)p * p);
    for (size_t i = 0; i < Hm.size(); ++i) Hm[i] = -F[i];
    // u_j = dK_j alpha, w_j = K^-1 u_j
    std::vector<double*> U, Wv;
    for (int j = 0; j < p; ++j) {
        double* uj = static_cast<double*>(c.scratch(2 * nn * sizeof(double)));
        double* wj = uj + nn;
        dk_matvec(H, dp[j], alpha, uj, 1);
        copy2d(c, {{uj, (int)n, wj, (int)n, (int)n, 1, 1.0}});
        hodlr_solve(H, wj, 1);
        U.push_back(uj);
        Wv.push_back(wj);
    }
    auto dot = [&](const double* a, const double* b) {
        column_dots(c, a, (int)n, b, (int)n, (int)n, 1, dots);
        double h = 0.0;
        KMH_HIP(hipMemcpyAsync(&h, dots, sizeof(double), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        return h;
    };
    int q = 0;
    for (int j = 0; j < p; ++j)
        for (int kk = j; kk < p; ++kk, ++q) {
            const kmh_kernel& djk = d2[q];
            const double base = -dot(U[kk], Wv[j]) - dot(U[j], Wv[kk]);   // stochastichessian.jl:9-10, 53-54
            if (djk.family == 0) {                                          // ZeroFunction (:139-142)
                Hm[j + kk * p] -= 0.5 * base;
                continue;
            }
            check_kernel(djk, true);
            DPanels d2p = dpanels_make(H, djk);
            dterm_matvec(H, d2p.leaf, {H.kxp.as<double>(), dp[j].kxp, dp[kk].kxp, d2p.kxp},
                         {nullptr, dp[j].kpp, dp[kk].kpp, d2p.kpp}, exact_second ? D2_EXACT_FWD : D2_REF_FWD,
                         exact_second ? D2_EXACT_BWD : D2_REF_BWD, Zb, O, k);
            // solve term + y' K^-1 d2K K^-1 y   (:11-12) and trace term (:79-89)
            const double sterm = base + dot(alpha, O);
            column_dots(c, T, (int)n, O + nn, (int)n, (int)n, (int)nv, dots);
            std::vector<double> hd(nv);
            KMH_HIP(hipMemcpyAsync(hd.data(), dots, nv * sizeof(double), hipMemcpyDeviceToHost, c.q()));
            c.sync();
            double t2 = 0.0;
            for (double v : hd) t2 += v;
            t2 /= (double)nv;
            Hm[j + kk * p] += 0.5 * t2 - 0.5 * sterm;                       // :138
            dpanels_free(H, d2p);
        }
    for (int j = 0; j < p; ++j)
        for (int kk = 0; kk < p; ++kk) Hout[j + kk * p] = (kk >= j) ? Hm[j + kk * p] : Hm[kk + j * p];
    for (double* x : U) c.scratch_free(x);
    for (auto& d : dp) dpanels_free(H, d);
    c.scratch_free(Zb);
}

}  // namespace kmh
