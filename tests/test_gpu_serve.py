"""Serving loop: continuous batching of likelihood(+gradient) requests.
Each result must be the same alone and in the mix, admission respects the
HBM budget, and results match the oracle's nll_objective pieces."""
import ctypes as C

import numpy as np
import pytest

from oracle.kdsort import kd_sort

pytestmark = pytest.mark.gpu


def _mkreq(L, x, y, saa, params, level, rank, grad=True):
    r = L.kmh_serve_req()
    r.pts = x.ctypes.data
    r.n = x.shape[0]
    r.d = 2
    r.kernel.family = L.KMH_MATERN32
    r.kernel.params[0], r.kernel.params[1] = params
    r.kernel.nugget = 1e-6
    r.opts.method = L.KMH_NYSTROM
    r.opts.level_kind = L.KMH_FIXED_LEVEL
    r.opts.level = level
    r.opts.maxrank = rank
    r.opts.precision = L.KMH_FP64
    r.y = y.ctypes.data
    if grad:
        r.ndk = 2
        r.saa = saa.ctypes.data
        r.nv = saa.shape[1]
        for j, code in enumerate((L.KMH_D1, L.KMH_D2)):
            r.dks[j].family = L.KMH_MATERN32
            r.dks[j].deriv = code
            r.dks[j].params[0], r.dks[j].params[1] = params
    return r


def _problems(n_list, seed=0, nv=8):
    rng = np.random.default_rng(seed)
    out = []
    for n in n_list:
        x = np.ascontiguousarray(kd_sort(rng.random((n, 2))))
        y = rng.standard_normal(n)
        saa = np.asfortranarray(rng.choice([-1.0, 1.0], size=(n, nv)))
        params = (rng.uniform(0.5, 2.0), rng.uniform(0.05, 0.3))
        level = max(1, int(np.log2(n / 256)))
        out.append((x, y, saa, params, level))
    return out


def _serve(lib, L, ctx, probs, max_inflight, budget=0, order=None, lanes=None, reserve=None):
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, max_inflight, budget, C.byref(srv)) == 0
    if lanes is not None:
        assert lib.kmh_server_set_lanes(srv, lanes) == 0, lib.last_error()
    if reserve is not None:   # (n, opts, nv, requests)
        n, opts, nv, count = reserve
        assert lib.kmh_server_reserve(srv, n, C.byref(opts), nv, count) == 0, lib.last_error()
    ids = {}
    keep = []
    order = order or list(range(len(probs)))
    for i in order:
        x, y, saa, params, level = probs[i]
        r = _mkreq(L, x, y, saa, params, level, 40)
        keep.append(r)
        rid = C.c_int64()
        assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0, lib.last_error()
        ids[i] = rid.value
    peak = 0
    for _ in range(1000):
        done = C.c_int32()
        assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()
        w, f, rb = C.c_int32(), C.c_int32(), C.c_uint64()
        lib.kmh_server_pending(srv, C.byref(w), C.byref(f), C.byref(rb))
        peak = max(peak, f.value)
        if w.value == 0 and f.value == 0:
            break
    res = {}
    for i, rid in ids.items():
        out = L.kmh_serve_result()
        rc = lib.kmh_server_result(srv, rid, C.byref(out))
        res[i] = (rc, out.status, out.nll, out.logdet, list(out.grad[:out.ngrad]), out.latency_us)
    lib.kmh_server_destroy(srv)
    return res, peak


def test_serve_alone_equals_mixed(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([1200, 4096, 2500, 700, 3000])
    alone, peak1 = _serve(lib, L, ctx, probs, 1)
    mixed, peakm = _serve(lib, L, ctx, probs, 8, order=[3, 1, 4, 0, 2])
    assert peak1 == 1 and peakm > 1
    for i in range(len(probs)):
        assert alone[i][0] == 0 and alone[i][1] == 0
        assert alone[i][2] == mixed[i][2] and alone[i][3] == mixed[i][3]
        assert alone[i][4] == mixed[i][4]
        print("request", i, "nll", alone[i][2], "grad", alone[i][4])


def test_serve_lanes_do_not_change_results(gpu_lib):
    """The same requests through 1, 2 and 3 lanes (private contexts stepped
    on parallel host threads): bitwise the same nll, logdet and gradient."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([1500, 2600, 900, 3100, 2048, 1200], seed=11)
    ref, _ = _serve(lib, L, ctx, probs, 6, lanes=1)
    for lanes in (2, 3):
        res, peak = _serve(lib, L, ctx, probs, 6, order=[5, 2, 0, 4, 1, 3], lanes=lanes)
        assert peak > 1
        for i in range(len(probs)):
            assert res[i][0] == 0 and res[i][1] == 0, lib.last_error()
            assert res[i][2] == ref[i][2] and res[i][3] == ref[i][3] and res[i][4] == ref[i][4]


def test_serve_set_lanes_errors(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, 4, 0, C.byref(srv)) == 0
    assert lib.kmh_server_set_lanes(srv, 0) == L.KMH_E_BAD_ARG
    assert lib.kmh_server_set_lanes(srv, 5) == L.KMH_E_BAD_ARG
    assert lib.kmh_server_set_lanes(srv, 4) == 0
    (x, y, saa, params, level), = _problems([1024], seed=2)
    r = _mkreq(L, x, y, saa, params, level, 40, grad=False)
    rid = C.c_int64()
    assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0
    assert lib.kmh_server_set_lanes(srv, 1) == L.KMH_E_STATE     # a request is queued
    done = C.c_int32()
    for _ in range(100):
        assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()
        if done.value:
            break
    out = L.kmh_serve_result()
    assert lib.kmh_server_result(srv, rid.value, C.byref(out)) == 0 and out.status == 0
    assert lib.kmh_server_set_lanes(srv, 1) == 0
    assert lib.kmh_server_destroy(srv) == 0


def test_serve_matches_direct_path(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([2000, 1500], seed=3)
    res, _ = _serve(lib, L, ctx, probs, 4)
    for i, (x, y, saa, params, level) in enumerate(probs):
        K = km.KernelMatrix(x, x, params, km.matern32(1e-6))
        HK = km.KernelHODLR(K, 0.0, 40, km.FixedLevel(level), nystrom=True)
        km.symmetricfactorize(HK)
        nll = km.negloglik(HK, y)
        g = km.stoch_gradient(K, HK, y, [K.kernel.d1, K.kernel.d2], saa)
        assert abs(res[i][2] - nll) <= 1e-12 * abs(nll)
        assert np.allclose(res[i][4], g, rtol=1e-12, atol=0)


def test_serve_admission_budget(gpu_lib):
    import kmhodlr as km
    from kmhodlr import _lib as L
    from kmhodlr import _lib
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([2048, 2048, 2048, 2048], seed=5)
    # budget for two of these requests at a time; per-request footprint as
    # the server counts it (n=2048, 8 leaves of 256, r=40, 3 levels, nv=8)
    n, m, r, L3, nv = 2048, 256, 40, 3, 8
    fp = (16 * n + 8 * n * m + 8 * n * r * L3 + 8 * 7 * 2 * 4 * r * r + 8 * n * (r * 4 + 3) + 24 * r * r
          + 8 * n * (2 + nv) * 2 + 8 * n * 16)
    res, peak = _serve(lib, L, ctx, probs, 8, budget=2 * fp + 1)
    assert peak <= 2
    assert all(r[1] == 0 for r in res.values())
    # a request larger than the whole budget fails alone, without blocking
    res2, _ = _serve(lib, L, ctx, probs[:2], 8, budget=1000)
    assert all(r[1] == L.KMH_E_BUDGET for r in res2.values())


def test_context_destroyed_before_its_handles(gpu_lib):
    """A context destroyed while a HODLR handle and a server still hold it is
    torn down only when the last of them is destroyed; the handle stays
    usable until then."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = C.c_void_p()
    assert lib.kmh_ctx_create(0, 0, C.byref(ctx)) == 0
    rng = np.random.default_rng(7)
    n = 1500
    x = rng.random((n, 2))
    kern = L.kmh_kernel()
    kern.family = L.KMH_MATERN32
    kern.params[0], kern.params[1], kern.nugget = 1.0, 0.1, 1e-6
    opts = L.kmh_build_opts()
    opts.method = L.KMH_NYSTROM
    opts.maxrank = 40
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = 2
    h = C.c_void_p()
    x = np.ascontiguousarray(x)   # held across the call
    assert lib.kmh_hodlr_build(ctx, x.ctypes.data, n, 2, C.byref(kern), C.byref(opts),
                               C.byref(h)) == 0, lib.last_error()
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx, 4, 0, C.byref(srv)) == 0
    assert lib.kmh_ctx_destroy(ctx) == 0          # deferred: two holders left
    assert lib.kmh_hodlr_factorize(h) == 0, lib.last_error()
    ld = C.c_double()
    assert lib.kmh_hodlr_logdet(h, C.byref(ld)) == 0 and np.isfinite(ld.value)
    assert lib.kmh_server_destroy(srv) == 0
    assert lib.kmh_hodlr_destroy(h) == 0          # last holder: the context goes with it


def test_server_destroyed_with_requests_in_flight(gpu_lib):
    """A server with three lanes destroyed after two steps (requests built,
    factorization under way): its private lane contexts go after the
    requests' handles have been returned to their pools, and the caller's
    context stays usable for another server."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([1500, 2600, 900, 3100, 2048], seed=13)
    live0 = _live(lib, ctx)
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, 8, 0, C.byref(srv)) == 0
    assert lib.kmh_server_set_lanes(srv, 3) == 0
    keep = []
    for x, y, saa, params, level in probs:
        r = _mkreq(L, x, y, saa, params, level, 40)
        keep.append(r)
        rid = C.c_int64()
        assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0, lib.last_error()
    done = C.c_int32()
    for _ in range(2):
        assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()
    w, f, rb = C.c_int32(), C.c_int32(), C.c_uint64()
    assert lib.kmh_server_pending(srv, C.byref(w), C.byref(f), C.byref(rb)) == 0
    assert f.value == len(probs) and done.value == 0
    assert lib.kmh_server_destroy(srv) == 0
    assert _live(lib, ctx) == live0       # the transient factorization blocks went with the handles
    res, _ = _serve(lib, L, ctx, probs[:2], 4)
    assert all(r[0] == 0 and r[1] == 0 for r in res.values())


def test_serve_failed_request_leaves_the_others(gpu_lib):
    """A request whose kernel matrix is not positive definite (nugget -5)
    fails with KMH_E_NOT_PD; the requests beside it, in its lane and in the
    other lane, finish with the same results as alone.  The server's build
    leaves the landmark Cholesky's flag to the factorization's final check:
    the failure still names the landmark matrix K(P,P), whose diagonal is
    sigma^2 - 5 < 0."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([1500, 2600, 900, 3100], seed=17)
    alone, _ = _serve(lib, L, ctx, probs, 1, lanes=1)
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, 8, 0, C.byref(srv)) == 0
    assert lib.kmh_server_set_lanes(srv, 2) == 0
    keep, ids = [], []
    for i, (x, y, saa, params, level) in enumerate(probs + [probs[1]]):
        r = _mkreq(L, x, y, saa, params, level, 40)
        if i == len(probs):
            r.kernel.nugget = -5.0
        keep.append(r)
        rid = C.c_int64()
        assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0, lib.last_error()
        ids.append(rid.value)
    done = C.c_int32()
    for _ in range(100):
        assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()
        w, f, rb = C.c_int32(), C.c_int32(), C.c_uint64()
        lib.kmh_server_pending(srv, C.byref(w), C.byref(f), C.byref(rb))
        if w.value == 0 and f.value == 0:
            break
    for i, rid in enumerate(ids):
        out = L.kmh_serve_result()
        rc = lib.kmh_server_result(srv, rid, C.byref(out))
        if i == len(probs):
            assert rc == 0 and out.status == L.KMH_E_NOT_PD
            assert "landmark matrix" in lib.last_error(), lib.last_error()
        else:
            assert rc == 0 and out.status == 0, lib.last_error()
            assert out.nll == alone[i][2] and out.logdet == alone[i][3]
    assert lib.kmh_server_destroy(srv) == 0


def _run_all(lib, srv, limit=200):
    done = C.c_int32()
    for _ in range(limit):
        assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()
        w, f, rb = C.c_int32(), C.c_int32(), C.c_uint64()
        lib.kmh_server_pending(srv, C.byref(w), C.byref(f), C.byref(rb))
        if w.value == 0 and f.value == 0:
            return rb.value
    raise AssertionError("server did not drain")


def _live(lib, ctx):
    b = C.c_uint64()
    assert lib.kmh_test_scratch_live(ctx.ptr, C.byref(b)) == 0
    return b.value


@pytest.mark.parametrize("lanes", [1, 2])
def test_resident_factor_follow_up_solve(gpu_lib, lanes):
    """keep_factor: a served request's factor stays resident after its
    nll_objective; a follow-up HK \\ b on it (kmh_server_factor +
    kmh_hodlr_solve, HODLR/src/baseoverloads.jl:181-192) equals the direct
    path's build + symmetricfactorize! + solve bitwise, and so does logdet.
    The factor's bytes stay reserved until kmh_server_release."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([3000, 1800, 2500], seed=21)
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, 8, 0, C.byref(srv)) == 0
    assert lib.kmh_server_set_lanes(srv, lanes) == 0
    keep, ids = [], []
    for i, (x, y, saa, params, level) in enumerate(probs):
        r = _mkreq(L, x, y, saa, params, level, 40, grad=(i != 1))
        r.keep_factor = 1
        keep.append(r)
        rid = C.c_int64()
        assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0, lib.last_error()
        ids.append(rid.value)
    reserved = _run_all(lib, srv)
    assert reserved > 0
    rng = np.random.default_rng(5)
    total = 0
    for i, rid in enumerate(ids):
        x, y, saa, params, level = probs[i]
        n = x.shape[0]
        out = L.kmh_serve_result()
        assert lib.kmh_server_result(srv, rid, C.byref(out)) == 0 and out.status == 0, lib.last_error()
        assert out.resident == 1 and out.factor_bytes > 0
        total += out.factor_bytes
        h = C.c_void_p()
        assert lib.kmh_server_factor(srv, rid, C.byref(h)) == 0, lib.last_error()
        b = np.asfortranarray(rng.standard_normal((n, 2)))
        xs = np.zeros_like(b)
        assert lib.kmh_hodlr_solve(h, b.ctypes.data, xs.ctypes.data, 2) == 0, lib.last_error()
        ld = C.c_double()
        assert lib.kmh_hodlr_logdet(h, C.byref(ld)) == 0 and ld.value == out.logdet
        # the direct path on the same inputs
        hd = C.c_void_p()
        assert lib.kmh_hodlr_build(ctx.ptr, x.ctypes.data, n, 2, C.byref(keep[i].kernel), C.byref(keep[i].opts),
                                   C.byref(hd)) == 0, lib.last_error()
        assert lib.kmh_hodlr_factorize(hd) == 0, lib.last_error()
        xd = np.zeros_like(b)
        assert lib.kmh_hodlr_solve(hd, b.ctypes.data, xd.ctypes.data, 2) == 0
        ldd = C.c_double()
        assert lib.kmh_hodlr_logdet(hd, C.byref(ldd)) == 0
        assert np.array_equal(xs, xd) and ldd.value == ld.value
        assert lib.kmh_hodlr_destroy(hd) == 0
        assert lib.kmh_hodlr_destroy(h) == L.KMH_E_STATE        # lent by the server
    assert total == reserved
    for rid in ids:
        assert lib.kmh_server_release(srv, rid) == 0, lib.last_error()
    w, f, rb = C.c_int32(), C.c_int32(), C.c_uint64()
    assert lib.kmh_server_pending(srv, C.byref(w), C.byref(f), C.byref(rb)) == 0 and rb.value == 0
    h = C.c_void_p()
    assert lib.kmh_server_factor(srv, ids[0], C.byref(h)) == L.KMH_E_BAD_ARG
    assert lib.kmh_server_destroy(srv) == 0


def test_resident_factor_survives_server_destroy_order(gpu_lib):
    """A server destroyed while factors are resident frees them with it (its
    private lane contexts go after them); the caller's context's scratch
    returns to its value before the server."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([1500, 2200, 1800], seed=27)
    live0 = _live(lib, ctx)
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, 8, 0, C.byref(srv)) == 0
    assert lib.kmh_server_set_lanes(srv, 3) == 0
    keep = []
    for x, y, saa, params, level in probs:
        r = _mkreq(L, x, y, saa, params, level, 40, grad=False)
        r.keep_factor = 1
        keep.append(r)
        rid = C.c_int64()
        assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0
    _run_all(lib, srv)
    assert lib.kmh_server_set_lanes(srv, 1) == L.KMH_E_STATE    # factors resident
    assert lib.kmh_server_destroy(srv) == 0
    assert _live(lib, ctx) == live0


def test_serve_step_failure_fails_its_lane_only(gpu_lib):
    """An allocation failure inside a step (injected on the host, no device
    fault) fails every request of that lane with KMH_E_OOM; their device
    memory, the failed stage's blocks and their reservations are released,
    the step returns KMH_OK, and new requests afterwards get the same
    results as before."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([1500, 2600, 900], seed=23)
    alone, _ = _serve(lib, L, ctx, probs, 1, lanes=1)
    live0 = _live(lib, ctx)
    for fail_after in (0, 3, 9):       # in the build, the leaf stage, a level
        srv = C.c_void_p()
        assert lib.kmh_server_create(ctx.ptr, 8, 0, C.byref(srv)) == 0
        assert lib.kmh_server_set_lanes(srv, 1) == 0
        keep, ids = [], []
        for x, y, saa, params, level in probs:
            r = _mkreq(L, x, y, saa, params, level, 40)
            keep.append(r)
            rid = C.c_int64()
            assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0
            ids.append(rid.value)
        done = C.c_int32()
        if fail_after > 0:
            assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()   # build
        assert lib.kmh_test_fail_scratch(ctx.ptr, fail_after) == 0
        assert lib.kmh_server_step(srv, C.byref(done)) == 0, lib.last_error()
        assert lib.kmh_test_fail_scratch(ctx.ptr, -1) == 0
        if fail_after > 0:
            # a stage shared by the lane's requests: all of them fail
            assert done.value == len(probs)
            w, f, rb = C.c_int32(), C.c_int32(), C.c_uint64()
            assert lib.kmh_server_pending(srv, C.byref(w), C.byref(f), C.byref(rb)) == 0
            assert f.value == 0 and rb.value == 0
        else:
            # the build reports per request: the one whose allocation failed
            # fails alone, the others go on
            assert done.value == 1
            assert _run_all(lib, srv) == 0
        for i, rid in enumerate(ids):
            out = L.kmh_serve_result()
            assert lib.kmh_server_result(srv, rid, C.byref(out)) == 0
            if fail_after > 0 or i == 0:
                assert out.status == L.KMH_E_OOM
            else:
                assert out.status == 0 and out.nll == alone[i][2], lib.last_error()
        assert _live(lib, ctx) == live0
        # the same server serves the next requests
        ids2 = []
        for r in keep:
            rid = C.c_int64()
            assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0
            ids2.append(rid.value)
        _run_all(lib, srv)
        for i, rid in enumerate(ids2):
            out = L.kmh_serve_result()
            assert lib.kmh_server_result(srv, rid, C.byref(out)) == 0 and out.status == 0, lib.last_error()
            assert out.nll == alone[i][2] and out.logdet == alone[i][3]
            assert list(out.grad[:out.ngrad]) == alone[i][4]
        assert lib.kmh_server_destroy(srv) == 0
        assert _live(lib, ctx) == live0


def test_serve_aca_request_under_budget(gpu_lib):
    """ACA requests are admitted on a rank estimate; after the build the
    reservation is the actual footprint.  Under a budget that fits the
    estimate but not a much larger actual, the request fails alone with
    KMH_E_BUDGET; under a roomy budget it succeeds and the reservation
    returns to zero."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    (x, y, saa, params, level), = _problems([2048], seed=29)
    r = _mkreq(L, x, y, saa, params, level, 0, grad=False)
    r.opts.method = L.KMH_ACA
    r.opts.tol = 1e-10
    for budget, ok in ((1 << 34, True), (1 << 20, False)):
        srv = C.c_void_p()
        assert lib.kmh_server_create(ctx.ptr, 4, budget, C.byref(srv)) == 0
        rid = C.c_int64()
        assert lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)) == 0
        assert _run_all(lib, srv) == 0
        out = L.kmh_serve_result()
        assert lib.kmh_server_result(srv, rid.value, C.byref(out)) == 0
        assert (out.status == 0) == ok, lib.last_error()
        if not ok:
            assert out.status == L.KMH_E_BUDGET
        else:
            res = out
        assert lib.kmh_server_destroy(srv) == 0
    out = res
    # the direct path's nll on the same ACA matrix
    K = km.KernelMatrix(x, x, params, km.matern32(1e-6))
    HK = km.KernelHODLR(K, 1e-10, 0, km.FixedLevel(level), nystrom=False)
    km.symmetricfactorize(HK)
    assert abs(out.nll - km.negloglik(HK, y)) <= 1e-10 * abs(out.nll)


def test_reserved_pools_serve_the_same(gpu_lib):
    """kmh_server_reserve sizes every lane's pool up front (and kmh_ctx_reserve
    a context's): results are bitwise those of a server that grows its pools
    while serving; bad arguments are refused."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0)
    probs = _problems([2048, 3000, 2048, 4096], seed=11)
    base, _ = _serve(lib, L, ctx, probs, 4, lanes=2)
    opts = L.kmh_build_opts()
    opts.method = L.KMH_NYSTROM
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = 4
    opts.maxrank = 40
    res, _ = _serve(lib, L, ctx, probs, 4, lanes=2, reserve=(4096, opts, 8, 4))
    for i in range(len(probs)):
        assert res[i][:5] == base[i][:5], i
    assert lib.kmh_ctx_reserve(ctx.ptr, 64 << 20) == 0, lib.last_error()
    assert lib.kmh_ctx_reserve(ctx.ptr, 0) == 0
    assert lib.kmh_ctx_reserve(None, 1) == L.KMH_E_BAD_ARG
    srv = C.c_void_p()
    assert lib.kmh_server_create(ctx.ptr, 2, 0, C.byref(srv)) == 0
    assert lib.kmh_server_reserve(srv, 0, C.byref(opts), 0, 1) == L.KMH_E_BAD_ARG
    assert lib.kmh_server_reserve(srv, 1024, None, 0, 1) == L.KMH_E_BAD_ARG
    assert lib.kmh_server_reserve(srv, 1024, C.byref(opts), 0, 0) == 0
    assert lib.kmh_server_destroy(srv) == 0
