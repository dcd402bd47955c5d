"""GPU checks of the batched FP64 primitives (GEMM on MFMA, Cholesky,
triangular solve, panel QR) against numpy, through the test hooks."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu_lib):
    import kmhodlr
    return kmhodlr.Context(0)


def _gemm(ctx, ta, tb, A, B, Cm, alpha=1.0, beta=0.0, lower=0):
    lib = ctx.lib
    A = np.asfortranarray(A)
    B = np.asfortranarray(B)
    Cm = np.asfortranarray(Cm.copy())
    m, n = Cm.shape
    k = A.shape[0] if ta else A.shape[1]
    rc = lib.kmh_test_gemm(ctx.ptr, ta, tb, m, n, k, A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0],
                           Cm.ctypes.data, m, alpha, beta, lower)
    assert rc == 0, lib.last_error()
    return Cm


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 5, 3), (64, 64, 16), (130, 70, 33), (72, 65, 5000),
                                   (300, 144, 72), (72, 144, 4096), (65, 160, 40), (129, 161, 7),
                                   # tall A^T B (k split in 256-row slabs) and wide A B shapes
                                   (72, 72, 256), (72, 72, 257), (80, 80, 4097), (32, 8, 32768),
                                   (80, 1, 513), (1, 80, 255), (81, 80, 300),
                                   (4097, 144, 72), (8192, 8, 64), (4096, 16, 600), (63, 32, 17),
                                   # row-per-thread narrow A B shapes (k up to 2048 / 8, 16, 32;
                                   # 128-row workgroups when m <= 128)
                                   (300, 32, 64), (5000, 16, 33), (77, 8, 8), (257, 9, 1),
                                   (1000, 1, 72), (700, 8, 256), (300, 16, 128), (129, 12, 100),
                                   (128, 32, 32), (100, 5, 200), (128, 8, 257)])
def test_gemm_matches_numpy(ctx, ta, tb, m, n, k):
    print("gemm", ta, tb, m, n, k, flush=True)
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    got = _gemm(ctx, ta, tb, A, B, C0, alpha=-0.75, beta=0.5)
    ref = -0.75 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.5 * C0
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * max(1, k) ** 0.5


def test_gemm_beta_zero_ignores_nan(ctx):
    A = np.ones((8, 4))
    B = np.ones((4, 8))
    C0 = np.full((8, 8), np.nan)
    got = _gemm(ctx, 0, 0, A, B, C0, 1.0, 0.0)
    assert np.all(got == 4.0)


def test_gemm_lower_skips_upper_tiles(ctx):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((200, 7))
    C0 = np.zeros((200, 200))
    got = _gemm(ctx, 0, 1, A, A, C0, 1.0, 0.0, lower=1)
    ref = A @ A.T
    assert np.allclose(np.tril(got), np.tril(ref), atol=1e-13)
    assert np.all(got[:64, 128:] == 0.0)


@pytest.mark.parametrize("n", [1, 5, 64, 65, 200, 257])
def test_potrf(ctx, n):
    print("potrf", n, flush=True)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    A = np.asfortranarray(X @ X.T + n * np.eye(n))
    B = A.copy(order="F")
    rc = ctx.lib.kmh_test_potrf(ctx.ptr, n, B.ctypes.data, n)
    assert rc == 0, ctx.lib.last_error()
    Lr = np.linalg.cholesky(A)
    assert np.allclose(B, Lr, rtol=0, atol=1e-12 * np.abs(Lr).max())
    assert np.all(np.triu(B, 1) == 0.0)


def test_potrf_not_pd(ctx):
    A = np.asfortranarray(np.diag([1.0, -1.0, 2.0]))
    rc = ctx.lib.kmh_test_potrf(ctx.ptr, 3, A.ctypes.data, 3)
    assert rc == 2 and "PosDef" in ctx.lib.last_error()
    # the flag is cleared: a good matrix factors again
    B = np.asfortranarray(np.eye(3))
    assert ctx.lib.kmh_test_potrf(ctx.ptr, 3, B.ctypes.data, 3) == 0


def _capacitance(ctx, R1, R2, ld=None):
    r = R1.shape[0]
    ld = ld or 2 * r
    X = np.asfortranarray(np.full((ld, 2 * r), 7.0))
    G = np.asfortranarray(np.full((ld, 2 * r), 7.0))
    ls = C.c_double()
    R1 = np.asfortranarray(R1)   # held: a temporary's address would be freed before the call
    R2 = np.asfortranarray(R2)
    rc = ctx.lib.kmh_test_capacitance(ctx.ptr, r, R1.ctypes.data, R2.ctypes.data, r, X.ctypes.data,
                                      G.ctypes.data, ld, C.byref(ls))
    return rc, X, G, ls.value


@pytest.mark.parametrize("r", [1, 5, 16, 17, 40, 72, 80])
def test_capacitance_matches_numpy(ctx, r):
    """The fused capacitance kernel against the dense closed forms of
    lrsymfact (HODLR/src/utils.jl:29-35): with Bm = R1 R2^T and
    L22 = chol(I - Bm^T Bm), the last r rows of X = chol(C) - I are
    [Bm^T, L22 - I] and of G = I - chol(C)^-1 are [L22^-1 Bm^T, I - L22^-1];
    the first r rows are not written; the log-sum is sum log diag L22.
    Sizes cover one block, a partial block and the padded 16 x 16 tiling."""
    print("capacitance", r, flush=True)
    rng = np.random.default_rng(100 + r)
    R1 = np.triu(rng.standard_normal((r, r)))
    R2 = np.triu(rng.standard_normal((r, r)))
    Bm = R1 @ R2.T
    s = np.linalg.norm(Bm, 2)
    R1 *= 0.97 / s                      # singular values of Bm up to 0.97: S well defined, condition ~ 17
    Bm = R1 @ R2.T
    rc, X, G, ls = _capacitance(ctx, R1, R2, ld=2 * r + 3)
    assert rc == 0, ctx.lib.last_error()
    L = np.linalg.cholesky(np.eye(r) - Bm.T @ Bm)
    Li = np.linalg.inv(L)
    Xr = np.hstack([Bm.T, L - np.eye(r)])
    Gr = np.hstack([Li @ Bm.T, np.eye(r) - Li])
    scale = max(1.0, np.abs(Li).max())
    assert np.abs(X[r:2 * r] - Xr).max() <= 1e-13 * scale
    assert np.abs(G[r:2 * r] - Gr).max() <= 1e-12 * scale
    assert np.all(X[:r] == 7.0) and np.all(G[:r] == 7.0) and np.all(X[2 * r:] == 7.0)
    assert abs(ls - np.log(np.diag(L)).sum()) <= 1e-12 * max(1.0, abs(ls))


def test_capacitance_not_pd(ctx):
    """|Bm| > 1 makes I - Bm^T Bm indefinite: the kernel's flag turns into
    KMH_E_NOT_PD (the reference's PosDefException), and the next block factors."""
    r = 20
    R1 = 2.0 * np.eye(r)
    R2 = np.eye(r)
    rc, _, _, _ = _capacitance(ctx, R1, R2)
    assert rc == 2 and "PosDef" in ctx.lib.last_error()
    rc, _, _, ls = _capacitance(ctx, 0.5 * np.eye(r), np.eye(r))
    assert rc == 0 and abs(ls - r * 0.5 * np.log(0.75)) <= 1e-13 * r


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("n,ncols", [(3, 2), (64, 300), (257, 70)])
def test_trsm(ctx, trans, n, ncols):
    print("trsm", trans, n, ncols, flush=True)
    rng = np.random.default_rng(n + ncols)
    # well-conditioned triangle (a random unit-scale triangle of order 257 is
    # exponentially ill-conditioned)
    Lm = np.asfortranarray(np.tril(rng.standard_normal((n, n))) / np.sqrt(n) + 2 * np.eye(n))
    B = np.asfortranarray(rng.standard_normal((n, ncols)))
    X = B.copy(order="F")
    rc = ctx.lib.kmh_test_trsm(ctx.ptr, trans, n, Lm.ctypes.data, n, ncols, X.ctypes.data, n)
    assert rc == 0, ctx.lib.last_error()
    ref = np.linalg.solve(Lm.T if trans else Lm, B)
    assert np.allclose(X, ref, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("m,r", [(5, 3), (64, 64), (256, 72), (300, 64), (1000, 130), (5000, 72), (33000, 70)])
def test_qr(ctx, m, r):
    print("qr", m, r, flush=True)
    rng = np.random.default_rng(m + r)
    A0 = rng.standard_normal((m, r)) * np.logspace(0, -8, r)    # graded columns
    A = np.asfortranarray(A0.copy())
    R = np.zeros((r, r), order="F")
    rc = ctx.lib.kmh_test_qr(ctx.ptr, m, r, A.ctypes.data, m, R.ctypes.data, r)
    assert rc == 0, ctx.lib.last_error()
    assert np.abs(A.T @ A - np.eye(r)).max() < 1e-12
    assert np.abs(A @ R - A0).max() <= 1e-13 * np.abs(A0).max()
    assert np.all(np.tril(R, -1) == 0.0)


def test_qr_tile_widths(ctx):
    """Block widths on both sides of every 8-column phase boundary of the
    tile QR (a short last phase rotates its Householder vectors into place
    before the cross-phase V^T V entries), on one tile (Q = E - V T V1^T
    straight from the root) and on a three-tile tree (inner node, lone
    pass-through nodes of a 5-tile tree)."""
    rng = np.random.default_rng(11)
    for m in (100, 300, 600):
        for r in (1, 2, 7, 8, 9, 15, 16, 17, 23, 24):
            print("qr widths", m, r, flush=True)
            A0 = rng.standard_normal((m, r)) * np.logspace(0, -6, r)
            A = np.asfortranarray(A0.copy())
            R = np.zeros((r, r), order="F")
            rc = ctx.lib.kmh_test_qr(ctx.ptr, m, r, A.ctypes.data, m, R.ctypes.data, r)
            assert rc == 0, ctx.lib.last_error()
            assert np.abs(A.T @ A - np.eye(r)).max() < 1e-12, (m, r)
            assert np.abs(A @ R - A0).max() <= 1e-13 * np.abs(A0).max(), (m, r)
            assert np.all(np.tril(R, -1) == 0.0), (m, r)


def test_qr_rank_deficient(ctx):
    """Zero and repeated columns inside a 64-column block (Householder
    tiles) and numerically dependent columns in a later block still give
    orthonormal Q."""
    rng = np.random.default_rng(7)
    m, r = 700, 80
    A0 = rng.standard_normal((m, r))
    A0[:, 10] = 0.0
    A0[:, 50] = A0[:, 3]
    A0[:, 70:] = A0[:, :10] @ rng.standard_normal((10, 10)) + 1e-15 * rng.standard_normal((m, 10))
    A = np.asfortranarray(A0.copy())
    R = np.zeros((r, r), order="F")
    assert ctx.lib.kmh_test_qr(ctx.ptr, m, r, A.ctypes.data, m, R.ctypes.data, r) == 0
    assert np.abs(A.T @ A - np.eye(r)).max() < 1e-10
    assert np.abs(A @ R - A0).max() <= 1e-13 * np.abs(A0).max()


def test_descriptor_arena_ring(ctx):
    """Uploads cycle several times through the descriptor arena's segment
    ring (64 MiB); every kernel must still read its own upload."""
    lib = ctx.lib
    bad = C.c_int64(-1)
    # 2600 uploads of 128 KiB = 325 MiB: about five passes around the ring,
    # segment ends hit exactly (128 KiB divides the 16 MiB segments)
    rc = lib.kmh_test_arena(ctx.ptr, 2600, 16384, C.byref(bad))
    assert rc == 0, lib.last_error()
    assert bad.value == 0
    # uneven sizes (not dividing the segments)
    rc = lib.kmh_test_arena(ctx.ptr, 3000, 12345, C.byref(bad))
    assert rc == 0, lib.last_error()
    assert bad.value == 0


def test_profile_hooks_time_sampled_launches(gpu_lib):
    """kmh_ctx_profile_config: only the masked tag is timed, at every
    stride-th launch; the GEMM totals then cover those launches.  The wide NN
    shape (4097 x 144, k = 72) routes to k_gemm_nn_wide."""
    import kmhodlr
    from kmhodlr import _lib as L
    ctx = kmhodlr.Context(0)
    lib = ctx.lib
    rng = np.random.default_rng(7)
    m, n, k = 4097, 144, 72
    A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
    byt = 8.0 * (m * k + k * n + m * n)
    fl = 2.0 * m * n * k

    def window(calls):
        assert lib.kmh_ctx_profile_begin(ctx.ptr) == 0
        for _ in range(calls):
            out = _gemm(ctx, 0, 0, A, B, np.zeros((m, n)))
        gfl, gms, gl = C.c_double(), C.c_double(), C.c_int64()
        assert lib.kmh_ctx_profile_end(ctx.ptr, C.byref(gfl), C.byref(gms), C.byref(gl)) == 0
        wb, wms, wl = C.c_double(), C.c_double(), C.c_int64()
        assert lib.kmh_ctx_profile_query(ctx.ptr, L.KMH_PROF_NN_WIDE, C.byref(wb), C.byref(wms), C.byref(wl)) == 0
        np.testing.assert_allclose(out, A @ B, rtol=1e-12, atol=1e-12 * np.abs(A @ B).max())
        return gfl.value, gms.value, gl.value, wb.value, wms.value, wl.value

    gfl, gms, gl, wb, wms, wl = window(5)          # defaults: every launch of every tag
    assert gl == 5 and wl == 5 and gfl == 5 * fl and wb == 5 * byt and wms > 0 and gms >= wms
    assert lib.kmh_ctx_profile_config(ctx.ptr, 1 << L.KMH_PROF_NN_WIDE, 2) == 0
    gfl, gms, gl, wb, wms, wl = window(5)          # launches 0, 2, 4
    assert wl == 3 and gl == 3 and wb == 3 * byt and gfl == 3 * fl and wms > 0
    assert lib.kmh_ctx_profile_config(ctx.ptr, 1 << L.KMH_PROF_TN_TALL, 1) == 0
    assert window(2)[5] == 0                       # NN_WIDE not in the mask
    assert lib.kmh_ctx_profile_config(ctx.ptr, 0xF, 1) == 0
    assert lib.kmh_ctx_profile_begin(ctx.ptr) == 0   # per-launch spans (testing hook)
    _gemm(ctx, 0, 0, A, B, np.zeros((m, n)))
    _gemm(ctx, 0, 0, A, B, np.zeros((m, n)))
    cap = 4
    tg, gr, cnt = (C.c_int32 * cap)(), (C.c_int32 * cap)(), C.c_int32()
    am, fp, ms = (C.c_double * cap)(), (C.c_double * cap)(), (C.c_double * cap)()
    assert lib.kmh_test_profile_spans(ctx.ptr, cap, tg, am, fp, ms, gr, C.byref(cnt)) == 0
    assert cnt.value == 2
    for i in range(2):
        assert tg[i] == L.KMH_PROF_NN_WIDE and am[i] == byt and fp[i] == fl and ms[i] > 0 and gr[i] == (m + 63) // 64
    d = C.c_double()
    assert lib.kmh_ctx_profile_end(ctx.ptr, C.byref(d), C.byref(d), C.byref(C.c_int64())) == 0
    assert lib.kmh_test_profile_spans(ctx.ptr, cap, tg, am, fp, ms, gr, C.byref(cnt)) == L.KMH_E_STATE
    assert lib.kmh_ctx_profile_config(ctx.ptr, 0, 1) == L.KMH_E_BAD_ARG
    assert lib.kmh_ctx_profile_config(ctx.ptr, 1, 0) == L.KMH_E_BAD_ARG
    assert lib.kmh_ctx_profile_begin(ctx.ptr) == 0
    assert lib.kmh_ctx_profile_config(ctx.ptr, 1, 1) == L.KMH_E_STATE
    d = C.c_double()
    assert lib.kmh_ctx_profile_end(ctx.ptr, C.byref(d), C.byref(d), C.byref(C.c_int64())) == 0


def test_profile_hooks_time_tile_qr(gpu_lib):
    """KMH_PROF_TILE_QR: the 24-column tile QR launches of a panel QR are
    timed with their algorithmic bytes (tiles read, Householder vectors and R
    written) and stay out of the GEMM totals.  A 300 x 24 panel: two leaf
    tiles (150 rows each, four rows per lane) and one inner node stacking
    their R factors."""
    import kmhodlr
    from kmhodlr import _lib as L
    ctx = kmhodlr.Context(0)
    lib = ctx.lib
    m, r = 300, 24
    A0 = np.random.default_rng(3).standard_normal((m, r))
    A = np.asfortranarray(A0.copy())
    R = np.zeros((r, r), order="F")
    assert lib.kmh_ctx_profile_config(ctx.ptr, 1 << L.KMH_PROF_TILE_QR, 1) == 0
    assert lib.kmh_ctx_profile_begin(ctx.ptr) == 0
    assert lib.kmh_test_qr(ctx.ptr, m, r, A.ctypes.data, m, R.ctypes.data, r) == 0, lib.last_error()
    gfl, gms, gl = C.c_double(), C.c_double(), C.c_int64()
    assert lib.kmh_ctx_profile_end(ctx.ptr, C.byref(gfl), C.byref(gms), C.byref(gl)) == 0
    b, ms, n = C.c_double(), C.c_double(), C.c_int64()
    assert lib.kmh_ctx_profile_query(ctx.ptr, L.KMH_PROF_TILE_QR, C.byref(b), C.byref(ms), C.byref(n)) == 0
    assert gl.value == 0 and gfl.value == 0.0           # the mask times the tile QR only
    assert n.value == 2 and ms.value > 0                # the leaf launch and the inner level
    leaves = 2 * 8.0 * (2 * 150 * r + r * r)
    inner = 8.0 * (2 * 2 * r * r + r * r)
    assert b.value == leaves + inner
    assert np.abs(A @ R - A0).max() <= 1e-13 * np.abs(A0).max()


def test_arena_growth_between_uploads_of_one_launch(gpu_lib):
    """Descriptor-arena growth in the middle of a launch's uploads (arena
    shrunk through a test hook, so the HODLR stages' descriptor lists make it
    grow at different points): the arena replaced by a grow stays valid for
    the launch that uploaded into it, and build + factorize + solve give
    bitwise the results of a context with the default arena."""
    import ctypes as C
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    rng = np.random.default_rng(31)
    n = 3000
    x = np.ascontiguousarray(rng.random((n, 2)))
    b = rng.standard_normal(n)
    kern = L.kmh_kernel()
    kern.family = L.KMH_MATERN32
    kern.params[0], kern.params[1], kern.nugget = 1.0, 0.1, 1e-6
    opts = L.kmh_build_opts()
    opts.method = L.KMH_NYSTROM
    opts.maxrank = 40
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = 4

    def run(shrink):
        ctx = km.Context(0)
        if shrink:
            assert lib.kmh_test_arena_shrink(ctx.ptr, shrink) == 0, lib.last_error()
        h = C.c_void_p()
        assert lib.kmh_hodlr_build(ctx.ptr, x.ctypes.data, n, 2, C.byref(kern), C.byref(opts), C.byref(h)) == 0
        assert lib.kmh_hodlr_factorize(h) == 0, lib.last_error()
        out = np.zeros(n)
        assert lib.kmh_hodlr_solve(h, b.ctypes.data, out.ctypes.data, 1) == 0
        ld = C.c_double()
        assert lib.kmh_hodlr_logdet(h, C.byref(ld)) == 0
        assert lib.kmh_hodlr_destroy(h) == 0
        return out, ld.value

    ref = run(0)
    for shrink in (1024, 2048, 4096, 8192, 16384):
        got = run(shrink)
        assert np.array_equal(got[0], ref[0]) and got[1] == ref[1], shrink
