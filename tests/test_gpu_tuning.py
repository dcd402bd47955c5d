"""The GEMM autotuner and its table: a tuned table is written and read
back; a table that routes the path's products to other kernel classes
changes results only at rounding level; malformed tables are refused."""
import ctypes as C

import numpy as np
import pytest

from oracle.kdsort import kd_sort

pytestmark = pytest.mark.gpu


def _wide_launches(lib, L, ctx, x, b):
    """launches of the wide NN kernel during build + factorize + solve"""
    assert lib.kmh_ctx_profile_config(ctx.ptr, 1 << L.KMH_PROF_NN_WIDE, 1) == 0
    assert lib.kmh_ctx_profile_begin(ctx.ptr) == 0
    out = _run(lib, L, ctx, x, b)
    f, ms, cnt = C.c_double(), C.c_double(), C.c_int64()
    assert lib.kmh_ctx_profile_end(ctx.ptr, C.byref(f), C.byref(ms), C.byref(cnt)) == 0
    a, t, spans = C.c_double(), C.c_double(), C.c_int64()
    assert lib.kmh_ctx_profile_query(ctx.ptr, L.KMH_PROF_NN_WIDE, C.byref(a), C.byref(t), C.byref(spans)) == 0
    return out, spans.value


def _run(lib, L, ctx, x, b):
    kern = L.kmh_kernel()
    kern.family = L.KMH_MATERN32
    kern.params[0], kern.params[1], kern.nugget = 1.0, 0.1, 1e-6
    opts = L.kmh_build_opts()
    opts.method = L.KMH_NYSTROM
    opts.maxrank = 72
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = 5
    h = C.c_void_p()
    assert lib.kmh_hodlr_build(ctx.ptr, x.ctypes.data, x.shape[0], 2, C.byref(kern), C.byref(opts), C.byref(h)) == 0
    assert lib.kmh_hodlr_factorize(h) == 0, lib.last_error()
    ld = C.c_double()
    assert lib.kmh_hodlr_logdet(h, C.byref(ld)) == 0
    s = np.zeros_like(b)
    assert lib.kmh_hodlr_solve(h, b.ctypes.data, s.ctypes.data, 1) == 0
    assert lib.kmh_hodlr_destroy(h) == 0
    return ld.value, s


def test_autotune_writes_a_loadable_table(gpu_lib, tmp_path):
    import kmhodlr as km
    lib = km.load()
    ctx = km.Context(0, tuning=None)
    sh = [(0, 0, 1024, 144, 72, 64), (1, 0, 72, 144, 1024, 64), (0, 0, 1024, 24, 48, 64)]
    arr = (C.c_int32 * 18)(*[v for s in sh for v in s])
    path = str(tmp_path / "t.txt")
    assert lib.kmh_gemm_autotune(ctx.ptr, 3, arr, 3, path.encode()) == 0, lib.last_error()
    n = C.c_int32()
    assert lib.kmh_gemm_tuning_size(ctx.ptr, C.byref(n)) == 0 and n.value == 3
    text = open(path).read()
    print(text)
    assert text.count("\ngemm ") == 3
    ctx2 = km.Context(0, tuning=None)
    assert lib.kmh_gemm_tuning_load(ctx2.ptr, path.encode()) == 0
    assert lib.kmh_gemm_tuning_size(ctx2.ptr, C.byref(n)) == 0 and n.value == 3
    assert lib.kmh_gemm_tuning_load(ctx2.ptr, None) == 0
    assert lib.kmh_gemm_tuning_size(ctx2.ptr, C.byref(n)) == 0 and n.value == 0


def test_table_routes_change_only_rounding(gpu_lib, tmp_path):
    """Every Woodbury update, QR update and leaf-solve update of an
    n = 8192 factorization sent to the generic MFMA kernel instead of the
    wide and row kernels (fewer wide launches): logdet within 1e-12 and the
    solve within 1e-9 of the built-in routing's (both kernels accumulate the
    same k order on the FP64 MFMA, so they agree bitwise on these shapes)."""
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    rng = np.random.default_rng(2)
    x = np.ascontiguousarray(kd_sort(rng.random((8192, 2))))
    b = rng.standard_normal(8192)
    ref, wide_ref = _wide_launches(lib, L, km.Context(0, tuning=None), x, b)
    lines = []
    for lm in range(7, 13):                  # m = 128 .. 4096
        for ln, lk in ((8, 7), (5, 5), (5, 6), (8, 6)):
            lines.append("gemm 0 0 %d %d %d generic" % (lm, ln, lk))
    path = str(tmp_path / "force.txt")
    open(path, "w").write("\n".join(lines) + "\n")
    ctx = km.Context(0, tuning=path)
    n = C.c_int32()
    assert lib.kmh_gemm_tuning_size(ctx.ptr, C.byref(n)) == 0 and n.value == len(lines)
    got, wide_got = _wide_launches(lib, L, ctx, x, b)
    print("wide NN launches: built-in rules %d, table %d" % (wide_ref, wide_got))
    assert wide_got < wide_ref
    assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0])
    assert np.linalg.norm(got[1] - ref[1]) <= 1e-9 * np.linalg.norm(ref[1])


def test_malformed_table_is_refused(gpu_lib, tmp_path):
    import kmhodlr as km
    from kmhodlr import _lib as L
    lib = km.load()
    ctx = km.Context(0, tuning=None)
    path = str(tmp_path / "bad.txt")
    open(path, "w").write("gemm 0 0 7 8 7 fastest\n")
    assert lib.kmh_gemm_tuning_load(ctx.ptr, path.encode()) == L.KMH_E_BAD_ARG
    assert lib.kmh_gemm_tuning_load(ctx.ptr, str(tmp_path / "missing.txt").encode()) == L.KMH_E_BAD_ARG
