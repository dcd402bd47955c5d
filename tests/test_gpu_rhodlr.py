"""RKernelHODLR on the GPU (constructor.jl:130-161, baseoverloads.jl:291-305)
through the C ABI, against the oracle's restatement on the same inputs:
per-block ACA ranks equal, R*x equal to the oracle's R*x to 1e-12 of
||K||_F ||x||, and both close to the dense K(x1, x2) x at the ACA tolerance.
Covers x1 != x2 of different sizes with coincident points (nugget off the
diagonal), several right-hand sides, FixedLevel / LogLevel, level 0 and the
reference's maxrank error."""
import numpy as np
import pytest

from oracle import hodlr as H
from oracle import kernels as kk
from oracle.kdsort import kd_sort

pytestmark = pytest.mark.gpu

FAM = {"matern32": kk.MATERN32, "matern12": kk.MATERN12, "rational": kk.RATIONAL2}


@pytest.fixture(scope="module")
def km(gpu_lib):
    import kmhodlr
    return kmhodlr


def _points(n, rng):
    return kd_sort(rng.random((n, 2)))


def _pair(km, n1, n2, fam, params, nugget, shared, seed):
    rng = np.random.default_rng(seed)
    x1, x2 = _points(n1, rng), _points(n2, rng)
    if shared:
        x2[:shared] = x1[-shared:]
    kerns = {"matern32": km.matern32, "matern12": km.matern12, "rational": km.rational_kernel}
    Kg = km.KernelMatrix(x1, x2, params, kerns[fam](nugget))
    Ko = H.KernelMatrix(x1, x2, kk.Kernel(FAM[fam], params, nugget=nugget))
    return Kg, Ko, rng


def _lvl(km, kind, lv):
    return (km.FixedLevel(lv), H.FixedLevel(lv)) if kind == "fixed" else (km.LogLevel(lv), H.LogLevel(lv))


CASES = [
    # n1, n2, family, params, nugget, shared, tol, maxrank, level
    (3000, 2500, "matern32", (1.0, 0.1), 1e-6, 7, 1e-10, 0, ("fixed", 4)),
    (4096, 4096, "matern12", (1.3, 0.2), 1e-6, 0, 1e-8, 0, ("log", 7)),
    (1777, 2900, "rational", (1.0, 1.0), 1.0, 3, 1e-12, 40, ("fixed", 3)),
    (513, 257, "matern32", (2.0, 0.3), 0.0, 0, 1e-9, 0, ("fixed", 5)),
]


@pytest.mark.parametrize("case", CASES, ids=["m32_fixed4", "m12_log7", "rational_maxrank40", "small_deep"])
def test_rhodlr_matches_oracle(km, case):
    n1, n2, fam, params, nugget, shared, tol, mr, (kind, lv) = case
    print("rhodlr", case, flush=True)
    Kg, Ko, rng = _pair(km, n1, n2, fam, params, nugget, shared, seed=n1 + n2)
    lg, lo = _lvl(km, kind, lv)
    R = km.RKernelHODLR(Kg, tol, mr, lg)
    Ro = H.rkernelhodlr(Ko, tol, mr, lo)
    assert R.shape == (n1, n2)
    assert R.ranks() == Ro.ranks()
    D = Ko.full()
    nrm = np.linalg.norm(D)   # Frobenius: a cheap upper bound of ||K||_2
    X = rng.standard_normal((n2, 3))
    Yg = R * X
    Yo = np.stack([Ro.mul(X[:, j]) for j in range(3)], axis=1)
    for j in range(3):
        assert np.linalg.norm(Yg[:, j] - Yo[:, j]) <= 1e-12 * nrm * np.linalg.norm(X[:, j])
        # the approximation itself: within a (generous) multiple of the ACA tolerance
        assert np.linalg.norm(Yg[:, j] - D @ X[:, j]) <= max(1e-6, 1e4 * tol) * nrm * np.linalg.norm(X[:, j])
    v = X[:, 0]
    assert np.allclose(R * v, Yg[:, 0], rtol=0, atol=0)   # vector and matrix paths agree bitwise


def test_rhodlr_level_zero_is_dense(km):
    Kg, Ko, rng = _pair(km, 300, 200, "matern32", (1.0, 0.1), 1e-6, 2, seed=11)
    R = km.RKernelHODLR(Kg, 1e-10, 0, km.FixedLevel(0))
    assert R.levels == 0 and R.ranks() == []
    v = rng.standard_normal(200)
    D = Ko.full()
    assert np.linalg.norm(R * v - D @ v) <= 1e-13 * np.linalg.norm(D) * np.linalg.norm(v)


def test_rhodlr_maxrank_one_is_the_reference_error(km):
    # maxsz = 1 makes the reference's start-row loop end at strt == maxsz
    # (src/factorizations.jl:26): "Incompatible maxrank or matrix dimensions."
    Kg, _, _ = _pair(km, 256, 256, "matern32", (1.0, 0.1), 1e-6, 0, seed=5)
    with pytest.raises(km.ErrorException, match="Incompatible maxrank or matrix dimensions"):
        km.RKernelHODLR(Kg, 1e-10, 1, km.FixedLevel(2))
