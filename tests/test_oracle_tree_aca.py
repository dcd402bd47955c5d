"""Pins for the oracle's tree indices and ACA (CPU)."""
import numpy as np
import pytest

from oracle import hodlr as H
from oracle import kernels as kk
from oracle.kdsort import kd_sort


@pytest.mark.parametrize("n", [1, 2, 7, 64, 100, 1023, 1024, 1025, 4099])
@pytest.mark.parametrize("lvl", [H.FixedLevel(0), H.FixedLevel(3), H.FixedLevel(20),
                                 H.LogLevel(0), H.LogLevel(2), H.LogLevel(8)])
def test_tree_tiles_matrix_exactly(n, lvl):
    lv, lf, nonlf = H.hodlr_indices(n, lvl)
    # level count (indices.jl:5-24)
    fl = int(np.floor(np.log2(n)))
    want = min(max(0, lvl.lv), fl) if isinstance(lvl, H.FixedLevel) else min(max(0, fl - lvl.lv), fl)
    assert lv == want and len(lf) == 2 ** lv and len(nonlf) == lv
    szs = [b - a for (a, b, _, _) in lf]
    assert sum(szs) == n and max(szs) - min(szs) <= 1
    assert szs == sorted(szs, reverse=True)          # remainder goes to the first leaves
    cover = np.zeros((n, n), dtype=np.int64)
    for (a, b, c, d) in lf:
        cover[a:b, c:d] += 1
    for j, lev in enumerate(nonlf, start=1):
        assert len(lev) == 2 ** (lv - j)
        for k, (a, b, c, d) in enumerate(lev):
            # rows = second child, cols = first child, children adjacent
            assert d == a and c < d < b
            first = lf[k * 2 ** j: k * 2 ** j + 2 ** (j - 1)]
            second = lf[k * 2 ** j + 2 ** (j - 1): (k + 1) * 2 ** j]
            assert (c, d) == (first[0][0], first[-1][1])
            assert (a, b) == (second[0][0], second[-1][1])
            cover[a:b, c:d] += 1
            cover[c:d, a:b] += 1
    assert np.all(cover == 1)


def test_restrictedmaxabs_rules():
    # first strict maximum wins, signed value returned (src/utils.jl:2-15)
    x = np.array([1.0, -3.0, 3.0, -3.0, 2.0])
    assert H.restrictedmaxabs(x, set()) == (-3.0, 1)
    assert H.restrictedmaxabs(x, {1}) == (3.0, 2)
    assert H.restrictedmaxabs(x, {1, 2}) == (-3.0, 3)
    # all-zero candidates: (0.0, first free index)
    z = np.array([5.0, 0.0, 0.0, 0.0])
    assert H.restrictedmaxabs(z, {0}) == (0.0, 1)
    assert H.restrictedmaxabs(z, {0, 1}) == (0.0, 2)
    # every index excluded: start index stops at the last one
    assert H.restrictedmaxabs(np.array([1.0, 2.0]), {0, 1}) == (0.0, 1)


def test_aca_known_answer():
    """A full-rank 3x3 matrix whose pivot sequence is fixed by hand."""
    M = np.array([[1.0, 2.0, 3.0],
                  [2.0, 4.0, 6.5],
                  [0.0, 1.0, 1.0]])
    tr = []
    U, V = H.aca(H.DenseSource(M), 1e-14, 0, trace=tr)
    # row 0 -> col 2 (|3| max); residual col 2 -> row 1 (|6.5 - ...|):
    # col 2 of M = (3, 6.5, 1); after step 1 U[:,0] = (3, 6.5, 1) and the next
    # row is the largest |U| entry outside row 0 -> row 1.
    assert tr[0] == (0, 2)
    # residual row 1 = (2,4,6.5) - 6.5*(1/3,2/3,1) = (-1/6,-1/3,0) -> col 1
    assert tr[1] == (1, 1)
    assert tr[2] == (2, 0)
    assert np.allclose(U @ V.T, M, atol=1e-13)
    assert U.shape[1] == 3


@pytest.mark.parametrize("rank", [1, 3, 7])
def test_aca_recovers_exact_rank(rank):
    rng = np.random.default_rng(rank)
    A = rng.standard_normal((40, rank)) @ rng.standard_normal((rank, 33))
    U, V = H.aca(H.DenseSource(A), 1e-12, 0)
    assert U.shape[1] in (rank, rank + 1)   # the stop test counts the pair that met tol
    assert np.allclose(U @ V.T, A, atol=1e-10 * np.abs(A).max())


def test_aca_interpolation_property():
    """The cross approximation reproduces every pivot row and column."""
    rng = np.random.default_rng(5)
    x = kd_sort(rng.random((300, 2)))
    k = kk.Kernel(kk.MATERN32, (1.0, 0.2), nugget=1e-6)
    Kb = H.KernelMatrix(x[150:], x[:150], k)
    tr = []
    U, V = H.aca(Kb, 1e-6, 0, trace=tr)
    full = Kb.full()
    R = full - U @ V.T
    rows = [p[0] for p in tr]
    cols = [p[1] for p in tr]
    scale = np.abs(full).max()
    assert np.abs(R[rows, :]).max() <= 1e-13 * scale
    assert np.abs(R[:, cols]).max() <= 1e-13 * scale
    assert len(set(rows)) == len(rows) and len(set(cols)) == len(cols)
    # tolerance-scaled accuracy on a smooth kernel
    assert np.linalg.norm(R) <= 1e-4 * np.linalg.norm(full)


def test_aca_maxrank_and_quirks():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((20, 20))
    U, V = H.aca(H.DenseSource(A), 1e-16, 5)
    assert U.shape == (20, 5) and V.shape == (20, 5)
    # maxrank = 1 trips the start-row check (src/factorizations.jl:27)
    with pytest.raises(H.ACAError):
        H.aca(H.DenseSource(A), 1e-16, 1)
    # zero first rows are skipped (src/factorizations.jl:22-26)
    B = A.copy()
    B[0, :] = 0.0
    tr = []
    H.aca(H.DenseSource(B), 1e-10, 0, trace=tr)
    assert tr[0][0] == 1


def test_nystrom_landmarks_match_linrange():
    # n = 65536, r = 72 starts 1, 924, 1847, 2770 (1-based)
    lm = H.nystrom_landmarks(65536, 72) + 1
    assert list(lm[:4]) == [1, 924, 1847, 2770]
    assert lm[-1] == 65536 and len(set(lm)) == 72
    # exact fractional parts: every value is far from a .5 tie at these sizes
    for n in (1024, 4096, 65536, 131072, 262144):
        j = np.arange(72)
        frac = np.array([(1 * (71 - jj) + n * jj) % 71 for jj in j]) / 71.0
        assert np.all(np.abs(frac - 0.5) > 0.007)
