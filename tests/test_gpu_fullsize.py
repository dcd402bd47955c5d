"""Config-2 size (n = 65536, Matern-3/2, FixedLevel(8), 256 leaves of 256):
the GPU against the oracle's results for the same inputs, committed in
tests/golden/fullsize_n65536.json by tests/golden/make_fullsize.py (the
oracle needs minutes at this size).  Bars: per-block ranks equal, logdet and
negloglik within 1e-10 relative, the solve K^-1 y within 1e-8 relative in
norm and in its dot with y, sampled entries within 1e-7 of the largest (a
forward error: with the 1e-6 nugget cond(K) is about 4e9 here, so two
backward-stable solvers may differ by up to cond(K) u ~ 4e-7 entrywise), and
the normwise backward error of the
solve through the factorized matvec W W^T, ||W W^T s - y|| / ||s||, below
1e-12 (K has unit diagonal, so ||K|| >= 1; with the 1e-6 nugget ||s|| is
about 1e5 ||y||, so a bar relative to ||y|| would test the conditioning,
not the solve)."""
import json
import os

import numpy as np
import pytest

from tests.golden.make_fullsize import inputs

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_n65536.json")


@pytest.fixture(scope="module")
def km(gpu_lib):
    import kmhodlr
    return kmhodlr


@pytest.fixture(scope="module")
def golden():
    if not os.path.exists(FIX):
        pytest.skip("fixture missing: run tests/golden/make_fullsize.py")
    with open(FIX) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["nystrom72", "aca1e-10"])
def test_config2_full_size(km, golden, name):
    g = golden[name]
    print("full size", name, flush=True)
    x, y = inputs(golden["n"], golden["seed"])
    K = km.KernelMatrix(x, x, tuple(golden["params"]), km.matern32(golden["nugget"]))
    HK = km.KernelHODLR(K, g["ep"], g["maxrank"], km.FixedLevel(golden["level"]), nystrom=(name == "nystrom72"))
    assert HK.ranks() == g["ranks"]
    km.symmetricfactorize(HK)
    ld = HK.logdet()
    assert abs(ld - g["logdet"]) <= 1e-10 * abs(g["logdet"]), (ld, g["logdet"])
    nll = km.negloglik(HK, y)
    assert abs(nll - g["negloglik"]) <= 1e-10 * abs(g["negloglik"])
    s = HK.solve(y)
    assert abs(np.linalg.norm(s) - g["solve_norm"]) <= 1e-8 * g["solve_norm"]
    assert abs(np.dot(s, y) - g["solve_dot_y"]) <= 1e-8 * abs(g["solve_dot_y"])
    ref = np.asarray(g["solve_every_512"])
    assert np.abs(s[::512] - ref).max() <= 1e-7 * np.abs(ref).max()
    r = HK.matvec(s) - y
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(s)
