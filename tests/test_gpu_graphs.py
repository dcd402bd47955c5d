"""Follow-up solves on a factorized HODLR matrix through captured hipGraphs
(kmh_hodlr_solve_dev): every replay gives bitwise the uncaptured solve's
result, with the program's own eager work and synchronizations between the
replays and new right-hand sides written into the same buffers."""
import ctypes as C

import numpy as np
import pytest

from oracle.kdsort import kd_sort
from tests.devbuf import DevArray

pytestmark = pytest.mark.gpu


def _factorized(lib, L, ctx, n, level, rank, method, seed):
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(kd_sort(rng.random((n, 2))))
    kern = L.kmh_kernel()
    kern.family = L.KMH_MATERN32
    kern.params[0], kern.params[1], kern.nugget = 1.1, 0.1, 1e-6
    opts = L.kmh_build_opts()
    opts.method = method
    opts.maxrank = rank
    opts.tol = 1e-10
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = level
    h = C.c_void_p()
    assert lib.kmh_hodlr_build(ctx.ptr, x.ctypes.data, n, 2, C.byref(kern), C.byref(opts), C.byref(h)) == 0
    assert lib.kmh_hodlr_factorize(h) == 0, lib.last_error()
    return h


def _stats(lib, h):
    cap, rep = C.c_int64(), C.c_int64()
    assert lib.kmh_hodlr_solve_graphs(h, -1, C.byref(cap), C.byref(rep)) == 0
    return cap.value, rep.value


@pytest.mark.parametrize("n,level,rank,method,nrhs", [(16384, 6, 72, "nys", 1), (8192, 5, 0, "aca", 3),
                                                      (65536, 8, 72, "nys", 2)])
def test_graph_replays_match_eager_solves(gpu_lib, n, level, rank, method, nrhs):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    m = L.KMH_NYSTROM if method == "nys" else L.KMH_ACA
    h = _factorized(lib, L, ctx, n, level, rank, m, n)
    he = _factorized(lib, L, ctx, n, level, rank, m, n)           # the same matrix, graphs off
    assert lib.kmh_hodlr_solve_graphs(he, 0, None, None) == 0
    rng = np.random.default_rng(1)
    bh = rng.standard_normal((nrhs, n))                            # column-major n x nrhs
    b, x = DevArray(bh), DevArray(np.zeros_like(bh))
    be, xe = DevArray(bh), DevArray(np.zeros_like(bh))
    scratch = DevArray(np.zeros(n))
    for it in range(5):
        if it == 3:                                                # new right-hand sides, same buffers
            bh = rng.standard_normal((nrhs, n))
            b.put(bh)
            be.put(bh)
        assert lib.kmh_hodlr_solve_dev(h, b.addr, x.addr, nrhs) == 0, lib.last_error()
        # eager work of the program between replays
        assert lib.kmh_hodlr_matvec_dev(he, scratch.addr, scratch.addr, 1) == 0
        assert lib.kmh_hodlr_solve_dev(he, be.addr, xe.addr, nrhs) == 0
        assert lib.kmh_ctx_synchronize(ctx.ptr) == 0
        got, ref = x.get(), xe.get()
        assert np.isfinite(ref).all() and np.array_equal(got, ref), it
    assert _stats(lib, h) == (1, 3)     # solve 1 eager, 2 captured, 3..5 replayed
    assert _stats(lib, he) == (0, 0)
    # in-place solves of one buffer (b == x) are graphs of their own
    y = DevArray(bh)
    for it in range(3):
        y.put(bh)
        assert lib.kmh_hodlr_solve_dev(h, y.addr, y.addr, nrhs) == 0
        assert lib.kmh_ctx_synchronize(ctx.ptr) == 0
        assert np.array_equal(y.get(), xe.get())
    assert _stats(lib, h) == (2, 4)
    assert lib.kmh_hodlr_destroy(h) == 0
    assert lib.kmh_hodlr_destroy(he) == 0
