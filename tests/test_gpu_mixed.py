"""Mixed forms: KMH_MIXED (FP32 off-diagonal panels, Nystrom cascade panels
and Q factors) and KMH_MIXED_Q (FP32 Q factors and panel QR, FP64 cascade
panels); FP64 leaves / capacitance / logdet in both, solves refined to FP64
against the FP64 HODLR operator.  Checked against the FP64 form of the same
problem (the error is stated per case) and against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import hodlr as H
from oracle import kernels as kk
from oracle.kdsort import kd_sort

pytestmark = pytest.mark.gpu


def _build(lib, L, ctx, x, fam, params, nugget, method, level, rank, tol, precision):
    kern = L.kmh_kernel()
    kern.family = fam
    kern.params[0], kern.params[1], kern.nugget = params[0], params[1], nugget
    opts = L.kmh_build_opts()
    opts.method = method
    opts.maxrank = rank
    opts.tol = tol
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = level
    opts.precision = precision
    h = C.c_void_p()
    n = x.shape[0]
    x = np.ascontiguousarray(x)   # held across the call
    assert lib.kmh_hodlr_build(ctx.ptr, x.ctypes.data, n, 2, C.byref(kern), C.byref(opts),
                               C.byref(h)) == 0, lib.last_error()
    return h


def _ranks(lib, h, L_):
    r = (C.c_int32 * (2 ** L_ - 1))()
    assert lib.kmh_hodlr_ranks(h, r) == 0
    return list(r)


def _solve(lib, h, b):
    x = np.zeros_like(b)
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    assert lib.kmh_hodlr_solve(h, b.ctypes.data, x.ctypes.data, nrhs) == 0, lib.last_error()
    return x


def _logdet(lib, h):
    ld = C.c_double()
    assert lib.kmh_hodlr_logdet(h, C.byref(ld)) == 0, lib.last_error()
    return ld.value


@pytest.mark.parametrize("mode", ["q", "panels"])
@pytest.mark.parametrize("n,level,params", [(4096, 4, (1.0, 0.1)), (16384, 6, (1.0, 0.1)),
                                            (16384, 6, (2.0, 0.2)), (65536, 8, (1.0, 0.1)),
                                            (65536, 8, (0.25, 0.05))])
def test_mixed_nystrom_vs_fp64(gpu_lib, n, level, params, mode):
    """Nystrom rank 72, Matern-3/2, nugget 1e-6: logdet of the mixed form
    within 1e-7 of the FP64 form's (relative; the error of each case is
    printed: FP32 panels perturb the R factors at the 1e-8 level, and the
    capacitance blocks with singular values near 1 amplify it, so neither
    mode holds 1e-8 on every problem); the refined solve's residual against
    the FP64 operator within 3x the FP64 solve's own residual (or <= 1e-10)
    and its solution within the FP64 solve's forward error (cond(K) u) of
    the FP64 solve's."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    rng = np.random.default_rng(n + level)
    x = kd_sort(rng.random((n, 2)))
    b = np.asfortranarray(rng.standard_normal((n, 2)))
    out = {}
    mixed = L.KMH_MIXED_Q if mode == "q" else L.KMH_MIXED
    quad = {}
    for prec in (L.KMH_FP64, mixed):
        h = _build(lib, L, ctx, x, L.KMH_MATERN32, params, 1e-6, L.KMH_NYSTROM, level, 72, 0.0, prec)
        assert lib.kmh_hodlr_factorize(h) == 0, lib.last_error()
        z = np.zeros(n)
        b0 = np.ascontiguousarray(b[:, 0])
        assert lib.kmh_hodlr_factor_apply(h, L.KMH_WINV, b0.ctypes.data, z.ctypes.data, 1) == 0
        quad[prec] = float(z @ z)          # b' (W W^T)^-1 b from the factor alone
        xs = _solve(lib, h, b)
        it, rr = C.c_int32(), C.c_double()
        assert lib.kmh_hodlr_refinement(h, C.byref(it), C.byref(rr)) == 0
        # residual against the FP64 operator (the mixed handle's matvec is it)
        kx = np.zeros_like(b)
        assert lib.kmh_hodlr_matvec(h, xs.ctypes.data, kx.ctypes.data, 2) == 0
        res = np.linalg.norm(b - kx, axis=0) / np.linalg.norm(b, axis=0)
        out[prec] = (_logdet(lib, h), xs, res.max(), it.value, rr.value)
        assert lib.kmh_hodlr_destroy(h) == 0
    ld64, x64, res64, _, _ = out[L.KMH_FP64]
    ld32, x32, res32, iters, rr = out[mixed]
    rel_ld = abs(ld32 - ld64) / abs(ld64)
    fwd = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    rel_q = abs(quad[mixed] - quad[L.KMH_FP64]) / abs(quad[L.KMH_FP64])
    print("mode=%s n=%d params=%s: logdet rel err %.2e, factor-only quad form rel err %.2e, IR %d corrections, "
          "residual %.2e (fp64 solve %.2e), solution vs fp64 %.2e"
          % (mode, n, params, rel_ld, rel_q, iters, res32, res64, fwd))
    assert rel_ld <= 1e-7
    assert iters >= 1 and rr == pytest.approx(res32, rel=1e-3, abs=1e-15)
    assert res32 <= max(1e-10, 3 * res64)
    assert fwd <= 1e-5


def test_mixed_nystrom_matches_oracle_logdet(gpu_lib):
    """The mixed form's logdet against the oracle's FP64 restatement (same
    points, parameters and landmarks) at n = 8192, to 1e-7, and its refined
    y'K^-1 y to 1e-8."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    n, level = 8192, 5
    rng = np.random.default_rng(3)
    x = kd_sort(rng.random((n, 2)))
    y = rng.standard_normal(n)
    h = _build(lib, L, ctx, x, L.KMH_MATERN32, (1.3, 0.12), 1e-6, L.KMH_NYSTROM, level, 72, 0.0, L.KMH_MIXED)
    assert lib.kmh_hodlr_factorize(h) == 0
    ld = _logdet(lib, h)
    xs = _solve(lib, h, y)
    K = H.KernelMatrix(x, x, kk.Kernel(kk.MATERN32, (1.3, 0.12), nugget=1e-6))
    HK = H.KernelHODLR(K, 0.0, 72, H.FixedLevel(level), nystrom=True)
    H.symmetricfactorize(HK)
    ldo = HK.logdet()
    xo = HK.solve(y)
    print("mixed logdet %.15e oracle %.15e rel %.2e" % (ld, ldo, abs(ld - ldo) / abs(ldo)))
    assert abs(ld - ldo) <= 1e-7 * abs(ldo)
    assert abs(y @ xs - y @ xo) <= 1e-8 * abs(y @ xo)
    assert lib.kmh_hodlr_destroy(h) == 0


@pytest.mark.parametrize("n,level,rho", [(4096, 4, 0.1), (9000, 5, 0.25)])
def test_mixed_aca(gpu_lib, n, level, rho):
    """ACA at tol 1e-6 (Matern-1/2, the batch workload's kernel): the mixed
    form keeps the ACA panels in FP32, so its operator is the FP64 one with
    U, V rounded to FP32.  Ranks equal the FP64 form's; logdet within 1e-7
    of the FP64 form's; the refined solve reaches a residual <= 1e-10 against
    the mixed form's own FP64 operator (its matvec) and the dense matrix of
    that operator (full(HK) before factorization)."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    rng = np.random.default_rng(n)
    x = kd_sort(rng.random((n, 2)))
    b = rng.standard_normal(n)
    res = {}
    for prec in (L.KMH_FP64, L.KMH_MIXED):
        h = _build(lib, L, ctx, x, L.KMH_MATERN12, (1.0, rho), 1e-6, L.KMH_ACA, level, 0, 1e-6, prec)
        ranks = _ranks(lib, h, level)
        full = None
        if prec == L.KMH_MIXED and n <= 4096:
            full = np.zeros((n, n), order="F")
            assert lib.kmh_hodlr_full(h, full.ctypes.data) == 0
        assert lib.kmh_hodlr_factorize(h) == 0, lib.last_error()
        xs = _solve(lib, h, b)
        kx = np.zeros(n)
        assert lib.kmh_hodlr_matvec(h, xs.ctypes.data, kx.ctypes.data, 1) == 0
        res[prec] = (ranks, _logdet(lib, h), xs, np.linalg.norm(b - kx) / np.linalg.norm(b), full)
        assert lib.kmh_hodlr_destroy(h) == 0
    r64, ld64, x64, _, _ = res[L.KMH_FP64]
    r32, ld32, x32, rel32, full = res[L.KMH_MIXED]
    assert r64 == r32
    print("ACA n=%d: mixed logdet rel err %.2e, residual %.2e" % (n, abs(ld32 - ld64) / abs(ld64), rel32))
    assert abs(ld32 - ld64) <= 1e-7 * abs(ld64)
    assert rel32 <= 1e-10
    if full is not None:
        assert np.linalg.norm(b - full @ x32) <= 1e-9 * np.linalg.norm(b)
        assert abs(ld32 - np.linalg.slogdet(full)[1]) <= 1e-9 * abs(ld32)


def test_mixed_and_fp64_in_one_batch(gpu_lib):
    """Mixed and FP64 problems share the batched build and factorization;
    each result is bitwise the one it gets alone."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    rng = np.random.default_rng(11)
    probs = []
    for i, (n, lvl, prec, method) in enumerate([(3000, 3, L.KMH_MIXED, L.KMH_NYSTROM), (2500, 3, L.KMH_FP64, L.KMH_NYSTROM),
                                                (4000, 4, L.KMH_MIXED, L.KMH_ACA), (1800, 2, L.KMH_MIXED, L.KMH_NYSTROM)]):
        x = np.ascontiguousarray(kd_sort(rng.random((n, 2))))
        probs.append((x, lvl, prec, method, rng.standard_normal(n)))

    def req(x, lvl, prec, method):
        r = L.kmh_build_req()
        r.pts = x.ctypes.data
        r.n = x.shape[0]
        r.d = 2
        r.kernel.family = L.KMH_MATERN32 if method == L.KMH_NYSTROM else L.KMH_MATERN12
        r.kernel.params[0], r.kernel.params[1], r.kernel.nugget = 1.0, 0.1, 1e-6
        r.opts.method = method
        r.opts.maxrank = 40 if method == L.KMH_NYSTROM else 0
        r.opts.tol = 1e-6
        r.opts.level_kind = L.KMH_FIXED_LEVEL
        r.opts.level = lvl
        r.opts.precision = prec
        return r

    def run(sel):
        reqs = (L.kmh_build_req * len(sel))(*[req(*probs[i][:4]) for i in sel])
        hs = (C.c_void_p * len(sel))()
        st = (C.c_int32 * len(sel))()
        assert lib.kmh_hodlr_build_batch(ctx.ptr, len(sel), reqs, hs, st) == 0, lib.last_error()
        assert all(v == 0 for v in st)
        assert lib.kmh_hodlr_factorize_batch(ctx.ptr, len(sel), hs, st) == 0
        assert all(v == 0 for v in st)
        out = {}
        for j, i in enumerate(sel):
            h = C.c_void_p(hs[j])
            out[i] = (_logdet(lib, h), _solve(lib, h, probs[i][4]))
            assert lib.kmh_hodlr_destroy(h) == 0
        return out

    mixed = run([0, 1, 2, 3])
    for i in range(4):
        alone = run([i])[i]
        assert alone[0] == mixed[i][0] and np.array_equal(alone[1], mixed[i][1])


def test_mixed_served_equals_direct(gpu_lib):
    """Served mixed-form requests (nll = 0.5 logdet + 0.5 y'x with x the
    refined K^-1 y) give bitwise the direct path's logdet and y'K^-1 y, and
    their nll is within 1e-7 of the FP64 form's."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    rng = np.random.default_rng(13)
    probs = []
    for n in (4096, 3000, 5000):
        x = np.ascontiguousarray(kd_sort(rng.random((n, 2))))
        probs.append((x, rng.standard_normal(n), (rng.uniform(0.8, 1.25), rng.uniform(0.08, 0.12)), 4))
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, 8, 0, C.byref(srv)) == 0
    keep, ids = [], []
    for x, y, params, lvl in probs:
        r = L.kmh_serve_req()
        r.pts = x.ctypes.data
        r.n = x.shape[0]
        r.d = 2
        r.kernel.family = L.KMH_MATERN32
        r.kernel.params[0], r.kernel.params[1], r.kernel.nugget = params[0], params[1], 1e-6
        r.opts.method = L.KMH_NYSTROM
        r.opts.maxrank = 72
        r.opts.level_kind = L.KMH_FIXED_LEVEL
        r.opts.level = lvl
        r.opts.precision = L.KMH_MIXED
        r.y = y.ctypes.data
        keep.append(r)
        rid = C.c_int64()
        assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0
        ids.append(rid.value)
    done = C.c_int32()
    for _ in range(200):
        assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()
        w, f, rb = C.c_int32(), C.c_int32(), C.c_uint64()
        lib.kmh_server_pending(srv, C.byref(w), C.byref(f), C.byref(rb))
        if w.value == 0 and f.value == 0:
            break
    for (x, y, params, lvl), rid, r in zip(probs, ids, keep):
        o = L.kmh_serve_result()
        assert lib.kmh_server_result(srv, rid, C.byref(o)) == 0 and o.status == 0, lib.last_error()
        nll = {}
        for prec in (L.KMH_MIXED, L.KMH_FP64):
            h = _build(lib, L, ctx, x, L.KMH_MATERN32, params, 1e-6, L.KMH_NYSTROM, lvl, 72, 0.0, prec)
            assert lib.kmh_hodlr_factorize(h) == 0
            ld = _logdet(lib, h)
            xs = _solve(lib, h, y)
            nll[prec] = 0.5 * ld + 0.5 * float(y @ xs)
            if prec == L.KMH_MIXED:
                assert ld == o.logdet
            assert lib.kmh_hodlr_destroy(h) == 0
        assert abs(o.nll - nll[L.KMH_MIXED]) <= 1e-13 * abs(o.nll)
        print("served mixed nll rel err vs FP64 %.2e" % (abs(o.nll - nll[L.KMH_FP64]) / abs(nll[L.KMH_FP64])))
        assert abs(o.nll - nll[L.KMH_FP64]) <= 1e-7 * abs(nll[L.KMH_FP64])
    assert lib.kmh_server_destroy(srv) == 0
