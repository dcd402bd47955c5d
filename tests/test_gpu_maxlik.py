"""GPU parity of the likelihood drivers (HODLR/src/maxlikfunctions.jl) and
the profile gradient (stochasticgradient.jl:48-86) against the oracle, on
the same KD-ordered points, data and SAA vectors; called through the C ABI.
Tolerances: likelihoods 1e-10 relative, gradients / Fisher / Hessian 1e-8
relative (stochastic estimators share their vectors, so the comparison is
deterministic)."""
import numpy as np
import pytest

from oracle import hodlr as H
from oracle import kernels as kk
from oracle.kdsort import kd_sort

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def km(gpu_lib):
    import kmhodlr
    return kmhodlr


def setup(n=2048, params=(1.3, 0.12), nv=16, seed=31):
    rng = np.random.default_rng(seed)
    x = kd_sort(rng.random((n, 2)))
    y = rng.standard_normal(n)
    vecs = [rng.choice([-1.0, 1.0], size=n) for _ in range(nv)]
    Ko = H.KernelMatrix(x, x, kk.Kernel(kk.MATERN32, params, nugget=1e-6))
    HKo = H.KernelHODLR(Ko, 0.0, 40, H.FixedLevel(3), nystrom=True)
    H.symmetricfactorize(HKo)
    return x, y, vecs, Ko, HKo


def opts_for(km, vecs, d2=False):
    k = km.matern32(1e-6)
    return km.maxlik.maxlikopts(kernfun=k, level=km.FixedLevel(3), rank=40, saavecs=vecs,
                                dfuns=[k.d1, k.d2], d2funs=[[None, k.d12], [k.d22]] if d2 else None)


def test_profile_gradient_parity(km):
    x, y, vecs, Ko, HKo = setup()
    go = H.stoch_profile_gradient(Ko, HKo, y, [Ko.kernel.derivative(kk.D1), Ko.kernel.derivative(kk.D2)], vecs)
    Kg = km.KernelMatrix(x, x, (1.3, 0.12), km.matern32(1e-6))
    HKg = km.KernelHODLR(Kg, 0.0, 40, km.FixedLevel(3), nystrom=True)
    km.symmetricfactorize(HKg)
    k = Kg.kernel
    gg = km.stoch_profile_gradient(Kg, HKg, y, [k.d1, k.d2], vecs)
    assert np.allclose(gg, go, rtol=1e-8, atol=0), (gg, go)
    assert abs(km.profile_negloglik(HKg, y) - H.profile_negloglik(HKo, y)) <= 1e-10 * abs(H.profile_negloglik(HKo, y))


def test_objectives_match_oracle(km):
    """nll_objective / nlpl_objective / scaleparm_mle / nll_gradient /
    fisher_matrix (maxlikfunctions.jl:21-146)."""
    x, y, vecs, Ko, HKo = setup()
    data = km.maxlik.Maxlikdata(x, y)
    opts = opts_for(km, vecs)
    prms = (1.3, 0.12)
    d1o = [Ko.kernel.derivative(kk.D1), Ko.kernel.derivative(kk.D2)]
    grad = np.zeros(2)
    nll = km.maxlik.nll_objective(prms, grad, data, opts)
    nllo = H.negloglik(HKo, y)
    assert abs(nll - nllo) <= 1e-10 * abs(nllo)
    assert np.allclose(grad, H.stoch_gradient(Ko, HKo, y, d1o, vecs), rtol=1e-8, atol=0)
    # an empty grad skips the gradient (length(grad) > 0 test, :41)
    assert km.maxlik.nll_objective(prms, np.zeros(0), data, opts) == nll
    pgrad = np.zeros(2)
    pnll = km.maxlik.nlpl_objective(prms, pgrad, data, opts)
    assert abs(pnll - H.profile_negloglik(HKo, y)) <= 1e-10 * abs(pnll)
    assert np.allclose(pgrad, H.stoch_profile_gradient(Ko, HKo, y, d1o, vecs), rtol=1e-8, atol=0)
    s = km.maxlik.scaleparm_mle(prms, data, opts)
    assert abs(s - float(np.dot(y, HKo.solve(y))) / len(y)) <= 1e-10 * s
    assert np.allclose(km.maxlik.nll_gradient(prms, data, opts), grad, rtol=0, atol=0)
    DKs = [H.DerivativeHODLR(Ko, dk, HKo) for dk in d1o]
    assert np.allclose(km.maxlik.fisher_matrix(prms, data, opts), H.stoch_fisher(Ko, HKo, DKs, vecs), rtol=1e-8)


def test_hessian_driver_and_errors(km):
    x, y, vecs, Ko, HKo = setup(n=1024, nv=8, seed=32)
    data = km.maxlik.Maxlikdata(x, y)
    prms = (1.3, 0.12)
    with pytest.raises(km.ErrorException, match="second derivatives"):
        km.maxlik.nll_hessian(prms, data, opts_for(km, vecs))
    Hg = km.maxlik.nll_hessian(prms, data, opts_for(km, vecs, d2=True))
    ko = Ko.kernel
    Ho = H.stoch_hessian(Ko, HKo, y, [ko.derivative(kk.D1), ko.derivative(kk.D2)],
                         [[None, ko.derivative(kk.D12)], [ko.derivative(kk.D22)]], vecs)
    assert np.allclose(Hg, Ho, rtol=1e-8, atol=1e-10 * np.abs(Ho).max())


def test_gpsimulate_exact_is_dense_cholesky(km):
    """gpsimulate(...; exact=true) (maxlikfunctions.jl:148-173) draws
    chol(K)' z: the Cholesky factor is unique, so the draw equals numpy's."""
    rng = np.random.default_rng(4)
    locs = rng.random((700, 2))
    z = rng.standard_normal(700)
    d = km.maxlik.gpsimulate(km.matern12(1e-6), locs, (1.5, 0.2), exact=True, kdtreesort=True, z=z)
    xs = d.pts_s
    assert np.array_equal(xs, locs[km.kd_order(locs)])
    Kd = H.KernelMatrix(xs, xs, kk.Kernel(kk.MATERN12, (1.5, 0.2), nugget=1e-6)).full()
    ref = np.linalg.cholesky(Kd) @ z
    assert np.linalg.norm(d.dat_s - ref) <= 1e-10 * np.linalg.norm(ref)


def test_gpsimulate_hodlr_uses_the_symmetric_factor(km):
    """The HODLR draw is W z for the factor W (W W' = full(HK))."""
    rng = np.random.default_rng(5)
    locs = kd_sort(rng.random((2048, 2)))
    z = rng.standard_normal(2048)
    k = km.matern32(1e-6)
    opts = km.maxlik.maxlikopts(kernfun=k, level=km.FixedLevel(3), rank=40)
    with pytest.warns(UserWarning):
        d = km.maxlik.gpsimulate(k, locs, (1.0, 0.1), opts, z=z)
    Kg = km.KernelMatrix(locs, locs, (1.0, 0.1), k)
    HKg = km.KernelHODLR(Kg, 0.0, 40, km.FixedLevel(3), nystrom=True)
    Fo = HKg.full()
    km.symmetricfactorize(HKg)
    assert np.array_equal(d.dat_s, HKg.W.mul(z))
    Wf = HKg.W.mul(np.eye(2048))
    assert np.abs(Wf @ Wf.T - Fo).max() <= 1e-11 * np.abs(Fo).max()
    with pytest.raises(km.ErrorException, match="Supply HODLR options"):
        km.maxlik.gpsimulate(k, locs, (1.0, 0.1), None, kdtreesort=True, z=z)
