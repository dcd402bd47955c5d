"""Pins for the oracle's covariance functions (CPU)."""
import mpmath
import numpy as np
import pytest

from oracle import kernels as kk


def stein_matern_cor(nu, x):
    """mtn_cor via the general Bessel form (src/covariancefunctions.jl:141-212),
    evaluated with mpmath's besselk."""
    if x <= 1e-8:
        return 1.0
    z = mpmath.sqrt(2 * nu) * x
    return float(z ** nu * mpmath.besselk(nu, z) / (mpmath.gamma(nu) * 2 ** (nu - 1)))


@pytest.mark.parametrize("fam,nu", [(kk.MATERN12, 0.5), (kk.MATERN32, 1.5)])
def test_matern_closed_form_matches_bessel(fam, nu):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(1e-6, 6.0, 64), [1e-9, 5e-9, 1.0, 2.5]])
    nu2 = 1 if fam == kk.MATERN12 else 3
    for x in xs:
        ref = stein_matern_cor(mpmath.mpf(nu), mpmath.mpf(float(x)))
        got = float(kk.matern_cor(nu2, x))
        assert abs(got - ref) <= 2e-15 * max(1.0, abs(ref)), (x, got, ref)


@pytest.mark.parametrize("fam", [kk.MATERN12, kk.MATERN32])
def test_matern_derivatives_finite_difference(fam):
    rng = np.random.default_rng(1)
    X = rng.random((40, 2))
    t0, rho = 1.3, 0.21
    k = kk.Kernel(fam, (t0, rho))
    base = k(X, X)
    h = 1e-6
    d1 = k.derivative(kk.D1)(X, X)
    fd1 = (kk.Kernel(fam, (t0 + h, rho))(X, X) - kk.Kernel(fam, (t0 - h, rho))(X, X)) / (2 * h)
    assert np.allclose(d1, fd1, atol=1e-8)
    d2 = k.derivative(kk.D2)(X, X)
    fd2 = (kk.Kernel(fam, (t0, rho + h))(X, X) - kk.Kernel(fam, (t0, rho - h))(X, X)) / (2 * h)
    assert np.allclose(d2, fd2, atol=1e-6)
    d12 = k.derivative(kk.D12)(X, X)
    fd12 = (kk.Kernel(fam, (t0, rho + h)).derivative(kk.D1)(X, X)
            - kk.Kernel(fam, (t0, rho - h)).derivative(kk.D1)(X, X)) / (2 * h)
    assert np.allclose(d12, fd12, atol=1e-6)
    d22 = k.derivative(kk.D22)(X, X)
    fd22 = (kk.Kernel(fam, (t0, rho + h)).derivative(kk.D2)(X, X)
            - kk.Kernel(fam, (t0, rho - h)).derivative(kk.D2)(X, X)) / (2 * h)
    assert np.allclose(d22, fd22, atol=1e-4 * np.abs(d22).max())
    assert np.all(k.derivative(kk.D11)(X, X) == 0.0)
    # r == 0 branches (src/covariancefunctions.jl:258-286)
    assert np.all(np.diag(base) == t0)
    assert np.all(np.diag(d1) == 1.0)
    assert np.all(np.diag(d2) == 0.0)


def test_nugget_by_point_equality():
    X = np.array([[0.0, 0.0], [1.0, 2.0], [0.0, 0.0], [1.0, 2.5]])
    k = kk.Kernel(kk.MATERN32, (1.0, 0.1), nugget=0.5)
    K = k(X, X)
    # duplicates at rows 0 and 2 get the nugget on both the diagonal and the
    # off-diagonal pair, as `x == y` does (README.md:44-46)
    assert K[0, 2] == 1.5 and K[2, 0] == 1.5 and K[0, 0] == 1.5
    assert K[1, 3] < 1.0
    kd = k.derivative(kk.D2)
    assert kd(X, X)[0, 0] == 0.0


def test_rational_kernels():
    X = np.array([[0.0, 0.0], [0.3, 0.4]])
    k = kk.Kernel(kk.RATIONAL2, (2.0, 0.5), nugget=1.0)
    K = k(X, X)
    r = 0.5
    assert np.isclose(K[0, 1], 4.0 / (1.0 + (r / 0.5) ** 2) ** 2, rtol=1e-15)
    assert K[0, 0] == 4.0 + 1.0
    c = kk.Kernel(kk.CAUCHY, (3.0, 2.0), nugget=1e-12)
    C = c(X, X)
    assert np.isclose(C[0, 1], 9.0 / (1.0 + (2.0 * r) ** 2), rtol=1e-15)
