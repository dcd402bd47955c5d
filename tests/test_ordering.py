"""KD ordering of the points (maxlikdata(sortmethod=:KDTree),
HODLR/src/constructor.jl:171-181) in the library's host code, against the
oracle's restatement; host-only, so it runs without a GPU."""
import numpy as np
import pytest

from oracle.kdsort import kd_order as oracle_kd_order


@pytest.fixture(scope="module")
def km():
    import kmhodlr
    try:
        kmhodlr.load()
    except kmhodlr.LibraryMissing as e:
        pytest.skip(str(e))
    return kmhodlr


@pytest.mark.parametrize("n,d,seed", [(1, 2, 0), (9, 2, 1), (1000, 2, 2), (4099, 2, 3), (3000, 3, 4)])
def test_kd_order_matches_oracle(km, n, d, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, d))
    assert np.array_equal(km.kd_order(x), oracle_kd_order(x))


def test_kd_order_ties_and_cells(km):
    # duplicated coordinates: stable order on ties
    rng = np.random.default_rng(5)
    x = np.round(rng.random((2000, 2)) * 8) / 8
    p = km.kd_order(x)
    assert np.array_equal(p, oracle_kd_order(x))
    assert np.array_equal(np.sort(p), np.arange(2000))
    # the two halves of the first split are separated along the wider axis
    y = np.column_stack([rng.random(1024) * 10.0, rng.random(1024)])
    q = km.kd_order(y)
    assert y[q[:512], 0].max() <= y[q[512:], 0].min()


def test_maxlikdata(km):
    rng = np.random.default_rng(6)
    x = rng.random((500, 2))
    data = np.arange(500.0)
    xs, ds = km.maxlikdata(x, data)
    perm = oracle_kd_order(x)
    assert np.array_equal(xs, x[perm]) and np.array_equal(ds, data[perm])
    with pytest.warns(UserWarning):
        xs2, ds2 = km.maxlikdata(x, data, sortmethod="Nothing")
    assert xs2 is not None and np.array_equal(ds2, data)
    with pytest.raises(km.ErrorException, match="The options are"):
        km.maxlikdata(x, data, sortmethod="Hilbert")
