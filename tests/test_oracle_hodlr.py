"""Pins for the oracle's HODLR build / factorization / solve / logdet /
derivatives (CPU).  The pattern is the reference's own test
(test/runtests.jl:22-45: everything against dense FP64 LAPACK on full(HK)),
with tightened tolerances, plus closed forms that a dropped term, a wrong
sign or a transposed operand would break."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import hodlr as H
from oracle import kernels as kk
from oracle.kdsort import kd_sort


def make(n, fam=kk.MATERN32, params=(1.0, 0.1), nugget=1e-6, seed=0, sort=True):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 2))
    if sort:
        x = kd_sort(x)
    k = kk.Kernel(fam, params, nugget=nugget)
    return H.KernelMatrix(x, x, k), rng


@pytest.mark.parametrize("nystrom", [False, True])
@pytest.mark.parametrize("n,lvl", [(1024, H.LogLevel(8)), (1000, H.FixedLevel(3)), (515, H.FixedLevel(2))])
def test_runtests_equivalent(nystrom, n, lvl):
    """test/runtests.jl:22-45 with tightened tolerances."""
    K, rng = make(n, fam=kk.CAUCHY, params=(3.0, 2.0), nugget=1e-12, seed=n)
    K = H.KernelMatrix(K.x1 * 10.0, K.x1 * 10.0, K.kernel)   # [0,10]^2 as basicexample.jl
    HK = H.KernelHODLR(K, 1e-14 if not nystrom else np.finfo(float).eps, 72 if nystrom else 0,
                       lvl, nystrom=nystrom)
    HKf = HK.full()
    H.symmetricfactorize(HK)
    W = HK.W.full()
    tvec = rng.standard_normal(n)
    scale = np.abs(HKf).max()
    assert np.abs(W @ W.T - HKf).max() <= 1e-12 * scale                        # :28
    ld = np.linalg.slogdet(HKf)[1]
    assert abs(HK.logdet() - ld) <= 1e-10 * abs(ld)                            # :31
    assert np.allclose(HK.matvec(tvec), HKf @ tvec, rtol=1e-12, atol=0)       # :34
    x = HK.solve(tvec)
    xd = np.linalg.solve(HKf, tvec)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)                # :37 (cond ~1e12)
    assert np.linalg.norm(HKf @ x - tvec) <= 1e-9 * np.linalg.norm(tvec)
    t1 = HK.W.ldiv(tvec)
    assert np.allclose(t1, np.linalg.solve(W, tvec), rtol=1e-8)               # :40-41
    t2 = HK.W.ldiv_t(t1)
    assert np.allclose(t2, np.linalg.solve(W.T, t1), rtol=1e-6)               # :44-45
    assert np.allclose(HK.W.mul(HK.W.mul_t(tvec)), HKf @ tvec, rtol=1e-10, atol=1e-10 * scale)


def test_level0_is_dense_cholesky():
    K, rng = make(300)
    HK = H.KernelHODLR(K, 1e-10, 0, H.FixedLevel(0))
    F = HK.full()
    H.symmetricfactorize(HK)
    L = np.linalg.cholesky(F)
    assert np.allclose(HK.W.full(), L, rtol=0, atol=1e-14)
    assert np.isclose(HK.logdet(), 2 * np.sum(np.log(np.diag(L))), rtol=1e-14)


def test_block_det_closed_form():
    """det(W) = prod diag(chol([I R1R2'; R2R1' I])) (closed form of
    HODLR/src/baseoverloads.jl:26-28), and log det C = log det(I - G1 G2)."""
    rng = np.random.default_rng(3)
    V = rng.standard_normal((50, 6)) * 0.1
    U = rng.standard_normal((40, 6)) * 0.1
    W = H.lrsymfact(V.copy(), U.copy())
    C = np.block([[np.eye(50), V @ U.T], [U @ V.T, np.eye(40)]])
    assert np.allclose(W.full() @ W.full().T, C, atol=1e-14)
    G1, G2 = V.T @ V, U.T @ U
    assert np.isclose(2 * np.log(abs(W.det())), np.linalg.slogdet(np.eye(6) - G1 @ G2)[1], rtol=1e-12)
    assert np.isclose(2 * np.log(abs(W.det())), np.linalg.slogdet(C)[1], rtol=1e-12)


def test_aca_hodlr_matches_exact_kernel():
    K, rng = make(2048, nugget=1e-6)
    HK = H.KernelHODLR(K, 1e-10, 0, H.FixedLevel(3))
    F = HK.full()
    Kf = K.full()
    assert np.linalg.norm(F - Kf) <= 1e-9 * np.linalg.norm(Kf)
    H.symmetricfactorize(HK)
    ld_exact = np.linalg.slogdet(Kf)[1]
    assert abs(HK.logdet() - ld_exact) <= 1e-6 * abs(ld_exact)


def test_nystrom_block_properties():
    K, rng = make(600, params=(1.0, 0.3))
    HK = H.KernelHODLR(K, 0.0, 20, H.FixedLevel(2), nystrom=True)
    lm = HK.nyind
    F = HK.full()
    Kf = K.full()
    # each off-diagonal block reproduces K(rows, P) at landmark rows/cols
    for lev in range(HK.lvl):
        for (a, b, c, d) in HK.nonleafindices[lev]:
            cols_in = [j for j in lm if c <= j < d]
            rows_in = [i for i in lm if a <= i < b]
            if cols_in:
                assert np.allclose(F[a:b, cols_in], Kf[a:b, cols_in], atol=1e-10)
            if rows_in:
                assert np.allclose(F[rows_in, c:d], Kf[rows_in, c:d], atol=1e-10)
    assert np.allclose(F, F.T, atol=1e-14)
    with pytest.raises(ValueError):
        H.KernelHODLR(K, 0.0, 1, H.FixedLevel(2), nystrom=True)
    with pytest.raises(ValueError):
        H.KernelHODLR(K, 0.0, 149, H.FixedLevel(2), nystrom=True)   # leaves of 150 -> max 148


def test_not_factorized_errors():
    K, _ = make(256)
    HK = H.KernelHODLR(K, 1e-8, 0, H.FixedLevel(1))
    with pytest.raises(H.NotFactorizedError):
        HK.solve(np.ones(256))
    with pytest.raises(H.NotFactorizedError):
        HK.logdet()


def _dense_grad(Kf, dKs, y):
    Ki = np.linalg.inv(Kf)
    a = Ki @ y
    return np.array([0.5 * (np.trace(Ki @ dK) - a @ dK @ a) for dK in dKs])


@pytest.mark.parametrize("fam", [kk.MATERN12, kk.MATERN32])
def test_stoch_gradient_exact_trace(fam):
    """With vecs = sqrt(n) e_i the estimator's trace is exact, so the
    stochastic gradient equals 0.5 (tr(K^-1 dK) - y'K^-1 dK K^-1 y) on the
    HODLR matrices (stochasticgradient.jl:14-26, exactmaxlik.jl:42-48)."""
    n = 256
    K, rng = make(n, fam=fam, params=(1.3, 0.15), seed=4)
    HK = H.KernelHODLR(K, 0.0, 24, H.FixedLevel(2), nystrom=True)
    Kf = HK.full()
    H.symmetricfactorize(HK)
    y = rng.standard_normal(n)
    dks = [K.kernel.derivative(kk.D1), K.kernel.derivative(kk.D2)]
    vecs = [np.sqrt(n) * e for e in np.eye(n)]
    g = H.stoch_gradient(K, HK, y, dks, vecs, nparams=2)
    dKs = [H.DerivativeHODLR(K, dk, HK).full() for dk in dks]
    assert np.allclose(g, _dense_grad(Kf, dKs, y), rtol=1e-9)


@pytest.mark.parametrize("profile", [False, True])
def test_gradients_match_finite_differences_of_the_likelihood(profile):
    """With exact trace vectors, stoch_gradient / stoch_profile_gradient
    (stochasticgradient.jl:14-86) equal central differences of negloglik /
    profile_negloglik (maxlikfunctions.jl:2-19) of the Nystrom HODLR
    approximant rebuilt at perturbed parameters."""
    n = 256
    K, rng = make(n, params=(1.3, 0.15), seed=11)
    y = rng.standard_normal(n)
    x = K.x1

    def nll(params):
        Kq = H.KernelMatrix(x, x, kk.Kernel(K.kernel.family, params, nugget=K.kernel.nugget))
        HKq = H.KernelHODLR(Kq, 0.0, 24, H.FixedLevel(2), nystrom=True)
        H.symmetricfactorize(HKq)
        return H.profile_negloglik(HKq, y) if profile else H.negloglik(HKq, y)

    HK = H.KernelHODLR(K, 0.0, 24, H.FixedLevel(2), nystrom=True)
    H.symmetricfactorize(HK)
    dks = [K.kernel.derivative(kk.D1), K.kernel.derivative(kk.D2)]
    vecs = [np.sqrt(n) * e for e in np.eye(n)]
    g = (H.stoch_profile_gradient(K, HK, y, dks, vecs) if profile
         else H.stoch_gradient(K, HK, y, dks, vecs))
    p0 = list(K.kernel.params)
    for i in range(2):
        h = 1e-4 * abs(p0[i])   # truncation O(h^2) ~ 1e-8; rounding of the nll ~ 1e-11 / h
        up, dn = list(p0), list(p0)
        up[i] += h
        dn[i] -= h
        fd = (nll(up) - nll(dn)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-5 * max(1.0, abs(fd)), (i, g[i], fd)


def test_derivative_hodlr_is_exact_derivative():
    """full(DerivativeHODLR) is the parameter derivative of full(HK) (finite
    differences of the Nystrom approximant itself)."""
    n = 200
    K, rng = make(n, params=(1.1, 0.2), seed=7)
    x = K.x1
    for code, idx in ((kk.D1, 0), (kk.D2, 1)):
        h = 1e-6
        p = list(K.kernel.params)
        fulls = []
        for s in (+1, -1):
            q = list(p)
            q[idx] += s * h
            Kq = H.KernelMatrix(x, x, kk.Kernel(K.kernel.family, q, nugget=K.kernel.nugget))
            fulls.append(H.KernelHODLR(Kq, 0.0, 16, H.FixedLevel(2), nystrom=True).full())
        fd = (fulls[0] - fulls[1]) / (2 * h)
        HK = H.KernelHODLR(K, 0.0, 16, H.FixedLevel(2), nystrom=True)
        D = H.DerivativeHODLR(K, K.kernel.derivative(code), HK).full()
        assert np.abs(D - fd).max() <= 1e-5 * np.abs(D).max()
        v = rng.standard_normal(n)
        assert np.allclose(H.DerivativeHODLR(K, K.kernel.derivative(code), HK).matvec(v), D @ v,
                           rtol=1e-11, atol=1e-12)


def test_fisher_and_hessian_exact_trace():
    """stoch_fisher / stoch_hessian (stochastichessian.jl:91-147) with exact
    trace vectors vs the dense second-order formulas on the HODLR matrices."""
    n = 128
    K, rng = make(n, params=(1.2, 0.25), seed=9)
    HK = H.KernelHODLR(K, 0.0, 12, H.FixedLevel(2), nystrom=True)
    Kf = HK.full()
    H.symmetricfactorize(HK)
    y = rng.standard_normal(n)
    d1 = [K.kernel.derivative(kk.D1), K.kernel.derivative(kk.D2)]
    d2 = [[None, K.kernel.derivative(kk.D12)], [K.kernel.derivative(kk.D22)]]
    vecs = [np.sqrt(n) * e for e in np.eye(n)]
    DKs = [H.DerivativeHODLR(K, dk, HK) for dk in d1]
    Dj = [D.full() for D in DKs]
    Ki = np.linalg.inv(Kf)
    F = H.stoch_fisher(K, HK, DKs, vecs)
    Fd = np.array([[0.5 * np.trace(Ki @ A @ Ki @ B) for B in Dj] for A in Dj])
    assert np.allclose(F, Fd, rtol=1e-9)
    Hs = H.stoch_hessian(K, HK, y, d1, d2, vecs)
    a = Ki @ y
    Hd = np.zeros((2, 2))
    for j in range(2):
        for k in range(j, 2):
            djk = d2[j][k - j]
            if djk is None:
                Kjk = np.zeros((n, n))
            else:
                D2 = H.SecondDerivative(K, djk, HK)
                Kjk = np.stack([H.deriv2mul(DKs[j], DKs[k], D2, e) for e in np.eye(n)], axis=1)
            Hd[j, k] = (-0.5 * np.trace(Ki @ Dj[j] @ Ki @ Dj[k]) + 0.5 * np.trace(Ki @ Kjk)
                        + 0.5 * (a @ Dj[k] @ Ki @ Dj[j] @ a + a @ Dj[j] @ Ki @ Dj[k] @ a)
                        - 0.5 * a @ Kjk @ a)
            Hd[k, j] = Hd[j, k]
    assert np.allclose(Hs, Hd, rtol=1e-8)


def _second_derivative_dense(K, HK, j, k, code, exact):
    n = K.shape[0]
    DKj = H.DerivativeHODLR(K, K.kernel.derivative([kk.D1, kk.D2][j]), HK)
    DKk = H.DerivativeHODLR(K, K.kernel.derivative([kk.D1, kk.D2][k]), HK)
    D2 = H.SecondDerivative(K, K.kernel.derivative(code), HK)
    return np.stack([H.deriv2mul(DKj, DKk, D2, e, exact=exact) for e in np.eye(n)], axis=1)


@pytest.mark.parametrize("j,k,code", [(0, 1, kk.D12), (1, 1, kk.D22)])
def test_second_derivative_blocks_vs_finite_difference(j, k, code):
    """The exact second-derivative HODLR equals the finite difference of the
    derivative HODLR; the reference's SDBlock_mul (HODLR/src/utils.jl:110-132)
    does not on its off-diagonal blocks (its leaves agree), which is why the
    restatement keeps both and labels the reference one as such."""
    n = 128
    K, rng = make(n, params=(1.2, 0.25), seed=3)
    HK = H.KernelHODLR(K, 0.0, 12, H.FixedLevel(2), nystrom=True)
    h = 1e-5
    fulls = []
    for s in (+1, -1):
        q = list(K.kernel.params)
        q[k] += s * h
        Kq = H.KernelMatrix(K.x1, K.x1, kk.Kernel(K.kernel.family, q, nugget=K.kernel.nugget))
        HKq = H.KernelHODLR(Kq, 0.0, 12, H.FixedLevel(2), nystrom=True)
        fulls.append(H.DerivativeHODLR(Kq, Kq.kernel.derivative([kk.D1, kk.D2][j]), HKq).full())
    fd = (fulls[0] - fulls[1]) / (2 * h)
    exact = _second_derivative_dense(K, HK, j, k, code, True)
    ref = _second_derivative_dense(K, HK, j, k, code, False)
    scale = np.abs(fd).max()
    assert np.abs(exact - fd).max() <= 1e-6 * scale
    leaf = np.zeros((n, n), bool)
    for (a0, b1, c0, d1) in HK.leafindices:
        leaf[a0:b1, c0:d1] = True
    assert np.abs((ref - fd)[leaf]).max() <= 1e-6 * scale
    assert np.abs((ref - fd)[~leaf]).max() > 1e-2 * np.abs(fd[~leaf]).max()


@pytest.mark.parametrize("j,k,code", [(0, 1, kk.D12), (1, 0, kk.D12), (1, 1, kk.D22)])
def test_reference_second_derivative_blocks_are_exact_plus_stated_delta(j, k, code):
    """Pins the reference-form block products SDBlock_mul / SDBlock_mul_t
    (HODLR/src/utils.jl:110-155) through the exact form (pinned by finite
    differences above): block by block, reference - exact equals the
    difference read off the reference's term list, written out here as
    separate dense products.  With A = K(P,P)^-1, K1 = K(rows,P),
    K2 = K(P,cols), a subscript j or k the parameter derivative and
    S_j, S_k, S_jk the derivatives of K(P,P):
      SDBlock_mul - exact =
          2 K1k A Sj A K2 - 2 K1 A Sk A Sj A K2 + 2 K1 A Sjk A K2
        - 2 K1 A Sj A Sk A K2 + K1 A Sj A (K2j + K2k)
        + K1k A (K2 - K2j) - K1 A Sk A (K2 - K2j)
      SDBlock_mul_t - exact^T = ( (K1k - K1j) A K2k + 2 K1k A Sj A K2
        - 2 K1 A Sk A Sj A K2 + 2 K1 A Sjk A K2 - 2 K1 A Sj A Sk A K2
        + 2 K1 A Sj A K2k )^T
    (sign flips of the d_k(-K1 A Sj A K2) terms, K2 in place of K2j in
    two terms of SDBlock_mul, K1k in place of K1j and the flipped K2k term
    in SDBlock_mul_t).  A dropped term, a wrong sign or a swapped operand in
    the restatement of either reference function breaks one of the two."""
    n = 128
    K, rng = make(n, params=(1.2, 0.25), seed=5)
    HK = H.KernelHODLR(K, 0.0, 12, H.FixedLevel(2), nystrom=True)
    d1 = [K.kernel.derivative(kk.D1), K.kernel.derivative(kk.D2)]
    DKj = H.DerivativeHODLR(K, d1[j], HK)
    DKk = H.DerivativeHODLR(K, d1[k], HK)
    D2 = H.SecondDerivative(K, K.kernel.derivative(code), HK)
    lm = K.x1[H.nystrom_landmarks(n, HK.mrnk)]
    Sfull = H.kernel_matrix(K.kernel, lm, lm)
    A = np.linalg.inv(Sfull)
    Sj, Sk, Sjk = DKj.Sj, DKk.Sj, D2.Sjk
    nchk = 0
    for lev in range(len(DKj.B)):
        for b, (a0, b1, c0, d1_) in enumerate(DKj.nonleafindices[lev]):
            Bj, Bk, Bjk = DKj.B[lev][b], DKk.B[lev][b], D2.B[lev][b]
            K1, K2 = Bj.K1p, Bj.Kp2
            K1j, K1k, K2j, K2k = Bj.K1pd, Bk.K1pd, Bj.Kp2d, Bk.Kp2d
            args = (Bj, Bk, Bjk)
            mats = (DKj.S, Sj, Sk, Sjk)
            Mref = np.stack([H.sdblock_mul(*args, e, *mats) for e in np.eye(d1_ - c0)], axis=1)
            Nref = np.stack([H.sdblock_mul_t(*args, e, *mats) for e in np.eye(b1 - a0)], axis=1)
            E = H.sdblock_full_exact(*args, *mats)
            delta = (2 * K1k @ A @ Sj @ A @ K2 - 2 * K1 @ A @ Sk @ A @ Sj @ A @ K2
                     + 2 * K1 @ A @ Sjk @ A @ K2 - 2 * K1 @ A @ Sj @ A @ Sk @ A @ K2
                     + K1 @ A @ Sj @ A @ (K2j + K2k)
                     + K1k @ A @ (K2 - K2j) - K1 @ A @ Sk @ A @ (K2 - K2j))
            delta_t = ((K1k - K1j) @ A @ K2k + 2 * K1k @ A @ Sj @ A @ K2
                       - 2 * K1 @ A @ Sk @ A @ Sj @ A @ K2 + 2 * K1 @ A @ Sjk @ A @ K2
                       - 2 * K1 @ A @ Sj @ A @ Sk @ A @ K2 + 2 * K1 @ A @ Sj @ A @ K2k).T
            scale = max(np.abs(Mref).max(), np.abs(E).max(), np.abs(delta).max())
            assert np.abs(delta).max() > 1e-3 * scale          # the forms do differ
            assert np.abs((Mref - E) - delta).max() <= 1e-9 * scale
            scale_t = max(np.abs(Nref).max(), np.abs(E).max(), np.abs(delta_t).max())
            assert np.abs((Nref - E.T) - delta_t).max() <= 1e-9 * scale_t
            nchk += 1
    assert nchk == 3
