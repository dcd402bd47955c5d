"""Batched execution: independent problems of different sizes, kernels and
compression methods share every grouped launch.  Each problem's results
must be the same alone and in the mix (bitwise: every reduction has a fixed
order that does not depend on the other problems of a launch)."""
import ctypes as C

import numpy as np
import pytest

from oracle.kdsort import kd_sort

pytestmark = pytest.mark.gpu


def _req(L, x, fam, params, nugget, method, level, maxrank, tol):
    r = L.kmh_build_req()
    r.pts = x.ctypes.data
    r.n = x.shape[0]
    r.d = 2
    r.on_device = 0
    r.kernel.family = fam
    r.kernel.deriv = 0
    r.kernel.params[0], r.kernel.params[1] = params
    r.kernel.nugget = nugget
    r.opts.method = method
    r.opts.level_kind = L.KMH_FIXED_LEVEL
    r.opts.level = level
    r.opts.maxrank = maxrank
    r.opts.tol = tol
    r.opts.precision = L.KMH_FP64
    return r


def test_alone_equals_in_the_mix(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    rng = np.random.default_rng(42)
    specs = [(3000, L.KMH_MATERN32, (1.0, 0.1), L.KMH_NYSTROM, 3, 48, 0.0),
             (1500, L.KMH_MATERN12, (0.7, 0.2), L.KMH_ACA, 2, 0, 1e-8),
             (5000, L.KMH_MATERN32, (2.0, 0.05), L.KMH_NYSTROM, 4, 72, 0.0),
             (800, L.KMH_MATERN12, (1.0, 0.3), L.KMH_ACA, 1, 0, 1e-10),
             # leaves of 250 and 150 rows (fused leaf kernels) beside the larger ones
             (2000, L.KMH_MATERN32, (1.0, 0.1), L.KMH_NYSTROM, 3, 40, 0.0),
             (1200, L.KMH_MATERN12, (1.0, 0.2), L.KMH_ACA, 3, 0, 1e-9)]
    xs = [np.ascontiguousarray(kd_sort(rng.random((n, 2)))) for (n, *_) in specs]
    ys = [rng.standard_normal(n) for (n, *_) in specs]
    reqs = [_req(L, x, f, p, 1e-6, m, lv, mr, tol) for x, (n, f, p, m, lv, mr, tol) in zip(xs, specs)]

    def run(idx):
        arr = (L.kmh_build_req * len(idx))(*[reqs[i] for i in idx])
        hs = (C.c_void_p * len(idx))()
        st = (C.c_int32 * len(idx))()
        assert lib.kmh_hodlr_build_batch(ctx.ptr, len(idx), arr, hs, st) == 0, lib.last_error()
        assert lib.kmh_hodlr_factorize_batch(ctx.ptr, len(idx), hs, st) == 0, lib.last_error()
        out = {}
        for j, i in enumerate(idx):
            ld = C.c_double()
            assert lib.kmh_hodlr_logdet(hs[j], C.byref(ld)) == 0
            x = np.zeros_like(ys[i])
            assert lib.kmh_hodlr_solve(hs[j], ys[i].ctypes.data, x.ctypes.data, 1) == 0
            nb = 2 ** specs[i][4] - 1
            rk = (C.c_int32 * nb)()
            lib.kmh_hodlr_ranks(hs[j], rk)
            out[i] = (ld.value, x, list(rk))
            lib.kmh_hodlr_destroy(hs[j])
        return out

    alone = {}
    for i in range(len(specs)):
        alone.update(run([i]))
    mixed = run([4, 2, 0, 5, 3, 1])
    for i in range(len(specs)):
        print("problem", i, "logdet", alone[i][0])
        assert alone[i][2] == mixed[i][2]
        assert alone[i][0] == mixed[i][0]
        assert np.array_equal(alone[i][1], mixed[i][1])


def test_batch_reports_per_request_failures(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    rng = np.random.default_rng(1)
    x = np.ascontiguousarray(kd_sort(rng.random((1200, 2))))
    good = _req(L, x, L.KMH_MATERN32, (1.0, 0.1), 1e-6, L.KMH_NYSTROM, 2, 40, 0.0)
    bad_rank = _req(L, x, L.KMH_MATERN32, (1.0, 0.1), 1e-6, L.KMH_NYSTROM, 2, 400, 0.0)
    not_pd = _req(L, x, L.KMH_MATERN32, (1.0, 0.1), -5.0, L.KMH_ACA, 2, 0, 1e-8)
    arr = (L.kmh_build_req * 3)(good, bad_rank, not_pd)
    hs = (C.c_void_p * 3)()
    st = (C.c_int32 * 3)()
    rc = lib.kmh_hodlr_build_batch(ctx.ptr, 3, arr, hs, st)
    assert rc == L.KMH_E_BAD_ARG and list(st) == [0, L.KMH_E_BAD_ARG, 0]
    assert hs[0] and not hs[1] and hs[2]
    two = (C.c_void_p * 2)(hs[0], hs[2])
    st2 = (C.c_int32 * 2)()
    rc = lib.kmh_hodlr_factorize_batch(ctx.ptr, 2, two, st2)
    assert rc == L.KMH_E_NOT_PD and list(st2) == [0, L.KMH_E_NOT_PD]
    ld = C.c_double()
    assert lib.kmh_hodlr_logdet(two[0], C.byref(ld)) == 0 and np.isfinite(ld.value)
    lib.kmh_hodlr_destroy(two[0])
    lib.kmh_hodlr_destroy(two[1])
