"""The allocation-free solve (kmh_hodlr_solve_dev_noalloc): the same
results as kmh_hodlr_solve_dev, bitwise, on two calls in a row (the second
on a smaller problem with fewer right-hand sides, so it runs over what the
first left in the fixed array), without a block from the context's scratch
pool; calls whose temporaries could exceed the fixed array fail before any
work is queued."""
import ctypes as C

import numpy as np
import pytest

from oracle.kdsort import kd_sort

pytestmark = pytest.mark.gpu


def _factorized(lib, L, ctx, n, level, rank, seed):
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(kd_sort(rng.random((n, 2))))
    kern = L.kmh_kernel()
    kern.family = L.KMH_MATERN32
    kern.params[0], kern.params[1], kern.nugget = 1.3, 0.12, 1e-6
    opts = L.kmh_build_opts()
    opts.method = L.KMH_NYSTROM
    opts.maxrank = rank
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = level
    h = C.c_void_p()
    assert lib.kmh_hodlr_build(ctx.ptr, x.ctypes.data, n, 2, C.byref(kern), C.byref(opts), C.byref(h)) == 0, \
        lib.last_error()
    assert lib.kmh_hodlr_factorize(h) == 0, lib.last_error()
    return h


class _Dev:
    """A device buffer of float64 through the HIP runtime the library itself
    links (libamdhip64.so.7), with host copies in and out."""
    _hip = None

    def __init__(self, host):
        if _Dev._hip is None:
            _Dev._hip = C.CDLL("libamdhip64.so.7")
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert _Dev._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        self.put(host)

    def put(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.nbytes == self.nbytes
        assert _Dev._hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def get(self):
        out = np.empty(self.shape)
        assert _Dev._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            _Dev._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _count(lib, ctx):
    c = C.c_uint64()
    assert lib.kmh_test_pool_allocs(ctx.ptr, C.byref(c)) == 0
    return c.value


def test_noalloc_solve_matches_solve_dev(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    cases = [(16384, 6, 72, 4, 1), (3000, 3, 40, 1, 2)]     # n, level, rank, nrhs, seed: second smaller
    for n, level, rank, nrhs, seed in cases:
        print("case n=%d level=%d rank=%d nrhs=%d" % (n, level, rank, nrhs), flush=True)
        h = _factorized(lib, L, ctx, n, level, rank, seed)
        bh = np.random.default_rng(seed + 10).standard_normal((nrhs, n))   # column-major n x nrhs
        b, x_ref, x_na = _Dev(bh), _Dev(np.zeros_like(bh)), _Dev(np.full_like(bh, np.nan))
        assert lib.kmh_hodlr_solve_dev(h, b.addr, x_ref.addr, nrhs) == 0, lib.last_error()
        assert lib.kmh_ctx_synchronize(ctx.ptr) == 0
        before = _count(lib, ctx)
        assert lib.kmh_hodlr_solve_dev_noalloc(h, b.addr, x_na.addr, nrhs) == 0, lib.last_error()
        assert lib.kmh_ctx_synchronize(ctx.ptr) == 0
        assert _count(lib, ctx) == before          # nothing from the pool
        ref = x_ref.get()
        assert np.isfinite(ref).all() and np.array_equal(x_na.get(), ref)
        # in place (b == x), as the _dev entry allows
        x_in = _Dev(bh)
        assert lib.kmh_hodlr_solve_dev_noalloc(h, x_in.addr, x_in.addr, nrhs) == 0, lib.last_error()
        assert lib.kmh_ctx_synchronize(ctx.ptr) == 0
        assert np.array_equal(x_in.get(), ref)
        for d in (b, x_ref, x_na, x_in):
            d.free()
        assert lib.kmh_hodlr_destroy(h) == 0


def test_noalloc_solve_rejects_what_does_not_fit(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    n = 16384
    h = _factorized(lib, L, ctx, n, 6, 72, 3)
    nrhs = 4096                                        # temporaries far beyond 32 MiB
    b = _Dev(np.zeros((nrhs, n)))
    x = _Dev(np.full((nrhs, n), 7.0))
    rc = lib.kmh_hodlr_solve_dev_noalloc(h, b.addr, x.addr, nrhs)
    assert rc == L.KMH_E_BAD_ARG and "fixed" in lib.last_error()
    assert lib.kmh_ctx_synchronize(ctx.ptr) == 0
    assert bool((x.get() == 7.0).all())                # nothing ran
    # unfactorized handles are refused as by the solve
    h2 = C.c_void_p()
    rng = np.random.default_rng(4)
    pts = np.ascontiguousarray(rng.random((2000, 2)))
    kern = L.kmh_kernel()
    kern.family = L.KMH_MATERN32
    kern.params[0], kern.params[1], kern.nugget = 1.0, 0.1, 1e-6
    opts = L.kmh_build_opts()
    opts.method = L.KMH_NYSTROM
    opts.maxrank = 30
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = 2
    assert lib.kmh_hodlr_build(ctx.ptr, pts.ctypes.data, 2000, 2, C.byref(kern), C.byref(opts), C.byref(h2)) == 0
    assert lib.kmh_hodlr_solve_dev_noalloc(h2, b.addr, x.addr, 1) == L.KMH_E_NOT_FACTORIZED
    b.free()
    x.free()
    assert lib.kmh_hodlr_destroy(h2) == 0
    assert lib.kmh_hodlr_destroy(h) == 0
