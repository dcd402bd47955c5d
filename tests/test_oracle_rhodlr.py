"""CPU pins of the oracle's RKernelHODLR (constructor.jl:130-161,
baseoverloads.jl:291-305): the recursive floor splits, dense leaves at the
last level, level 0 = the dense matrix, LogLevel conversion, and the matvec
against the dense kernel matrix (exact where every off-diagonal ACA reaches
full rank, tolerance-bounded otherwise), for x1 != x2 of different sizes."""
import numpy as np
import pytest

from oracle import hodlr as H
from oracle import kernels as kk


def _K(n1, n2, seed=0, fam=kk.MATERN32, params=(1.0, 0.1), nugget=1e-6, shared=0):
    rng = np.random.default_rng(seed)
    x1 = rng.random((n1, 2))
    x2 = rng.random((n2, 2))
    if shared:   # a few coincident points: the nugget lands off the diagonal
        x2[:shared] = x1[-shared:]
    return H.KernelMatrix(x1, x2, kk.Kernel(fam, params, nugget=nugget)), rng


def _tree_shapes(R, out, depth=0):
    if isinstance(R, np.ndarray):
        out.setdefault(depth, []).append(("leaf", R.shape))
        return
    out.setdefault(depth, []).append(("node", R.shape, R.A12.U.shape[0], R.A12.V.shape[0]))
    _tree_shapes(R.A11, out, depth + 1)
    _tree_shapes(R.A22, out, depth + 1)


def test_floor_splits_and_leaves():
    K, _ = _K(301, 250)
    R = H.rkernelhodlr(K, 1e-10, 0, H.FixedLevel(3))
    shapes = {}
    _tree_shapes(R, shapes)
    assert shapes[0] == [("node", (301, 250), 150, 125)]          # K12: rows 1:150, cols 126:250
    assert [s[1] for s in shapes[1]] == [(150, 125), (151, 125)]
    assert all(s[0] == "leaf" for s in shapes[3]) and len(shapes[3]) == 8
    assert sum(s[1][0] for s in shapes[3]) == 301 and sum(s[1][1] for s in shapes[3]) == 250
    ranks = R.ranks()
    assert [len(r) for r in ranks] == [1, 2, 4]


def test_level_zero_is_dense_and_loglevel_converts():
    K, _ = _K(64, 50)
    D = H.rkernelhodlr(K, 1e-12, 0, H.FixedLevel(0))
    assert isinstance(D, np.ndarray) and np.array_equal(D, K.full())
    # LogLevel(k) -> FixedLevel(floor(log2(min size)) - k): min(64, 50) -> 5 - 3 = 2
    R1 = H.rkernelhodlr(K, 1e-12, 0, H.LogLevel(3))
    R2 = H.rkernelhodlr(K, 1e-12, 0, H.FixedLevel(2))
    assert R1.ranks() == R2.ranks()


@pytest.mark.parametrize("fam,params", [(kk.MATERN32, (1.0, 0.1)), (kk.MATERN12, (1.3, 0.2)),
                                        (kk.RATIONAL2, (1.0, 1.0))])
def test_matvec_against_dense(fam, params):
    K, rng = _K(400, 333, seed=3, fam=fam, params=params, shared=5)
    R = H.rkernelhodlr(K, 1e-12, 0, H.FixedLevel(3))
    D = K.full()
    v = rng.standard_normal(333)
    err = np.linalg.norm(R.mul(v) - D @ v) / (np.linalg.norm(D, 2) * np.linalg.norm(v))
    assert err < 1e-9, err


def test_full_rank_blocks_are_exact():
    # maxrank large enough and tol 0: every ACA stops at the exact rank, so
    # the matvec reproduces the dense product to rounding
    K, rng = _K(40, 37, seed=5, nugget=0.0)
    R = H.rkernelhodlr(K, 0.0, 0, H.FixedLevel(2))
    v = rng.standard_normal(37)
    D = K.full()
    assert np.allclose(R.mul(v), D @ v, rtol=1e-10, atol=1e-10 * np.abs(D).max())
