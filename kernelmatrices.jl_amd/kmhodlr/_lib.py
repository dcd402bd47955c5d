"""ctypes binding of libkmhodlr.so (the C ABI in include/kmhodlr.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("KMHODLR_LIB", os.path.join(HERE, "libkmhodlr.so"))

KMH_OK = 0
KMH_E_BAD_ARG = 1
KMH_E_NOT_PD = 2
KMH_E_NOT_FACTORIZED = 3
KMH_E_OOM = 4
KMH_E_BUDGET = 5
KMH_E_DEVICE = 6
KMH_E_ACA = 7
KMH_E_STATE = 8

KMH_RATIONAL2, KMH_CAUCHY, KMH_MATERN12, KMH_MATERN32 = 1, 2, 3, 4
KMH_D0, KMH_D1, KMH_D2, KMH_D11, KMH_D12, KMH_D22 = 0, 1, 2, 11, 12, 22
KMH_ACA, KMH_NYSTROM = 0, 1
KMH_FIXED_LEVEL, KMH_LOG_LEVEL = 0, 1
KMH_FP64 = 0
KMH_MIXED = 1
KMH_MIXED_Q = 2
KMH_W, KMH_WT, KMH_WINV, KMH_WTINV = 0, 1, 2, 3
KMH_PROF_GEMM, KMH_PROF_NN_WIDE, KMH_PROF_TN_TALL, KMH_PROF_NN_ROWS, KMH_PROF_TILE_QR = 0, 1, 2, 3, 4


class kmh_kernel(C.Structure):
    _fields_ = [("family", C.c_int32), ("deriv", C.c_int32),
                ("params", C.c_double * 4), ("nugget", C.c_double)]


class kmh_build_opts(C.Structure):
    _fields_ = [("method", C.c_int32), ("level_kind", C.c_int32), ("level", C.c_int32),
                ("maxrank", C.c_int32), ("tol", C.c_double), ("precision", C.c_int32),
                ("reserved", C.c_int32 * 3)]


class kmh_build_req(C.Structure):
    _fields_ = [("pts", C.c_void_p), ("n", C.c_int64), ("d", C.c_int32), ("on_device", C.c_int32),
                ("kernel", kmh_kernel), ("opts", kmh_build_opts)]


KMH_SERVE_MAX_PARAMS = 4


class kmh_serve_req(C.Structure):
    _fields_ = [("pts", C.c_void_p), ("n", C.c_int64), ("d", C.c_int32), ("ndk", C.c_int32),
                ("on_device", C.c_int32), ("keep_factor", C.c_int32),
                ("kernel", kmh_kernel), ("opts", kmh_build_opts), ("y", C.c_void_p), ("saa", C.c_void_p),
                ("nv", C.c_int64), ("dks", kmh_kernel * KMH_SERVE_MAX_PARAMS)]


class kmh_serve_result(C.Structure):
    _fields_ = [("status", C.c_int32), ("ngrad", C.c_int32), ("nll", C.c_double), ("logdet", C.c_double),
                ("grad", C.c_double * KMH_SERVE_MAX_PARAMS), ("queue_us", C.c_double),
                ("latency_us", C.c_double), ("resident", C.c_int32), ("reserved", C.c_int32),
                ("factor_bytes", C.c_int64)]


class kmh_info(C.Structure):
    _fields_ = [("n", C.c_int64), ("levels", C.c_int32), ("nleaves", C.c_int32),
                ("factorized", C.c_int32), ("method", C.c_int32), ("device_bytes", C.c_int64)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.c_int64

# name -> (argtypes); every function returns int status except kmh_last_error
SIGNATURES = {
    "kmh_last_error": [],
    "kmh_device_count": [C.POINTER(C.c_int)],
    "kmh_ctx_create": [C.c_int, C.c_uint64, C.POINTER(_P)],
    "kmh_ctx_destroy": [_P],
    "kmh_ctx_synchronize": [_P],
    "kmh_ctx_bytes_in_use": [_P, C.POINTER(C.c_uint64)],
    "kmh_ctx_reserve": [_P, C.c_uint64],
    "kmh_ctx_profile_begin": [_P],
    "kmh_ctx_profile_end": [_P, _D, _D, C.POINTER(C.c_int64)],
    "kmh_ctx_profile_query": [_P, C.c_int32, _D, _D, C.POINTER(C.c_int64)],
    "kmh_ctx_profile_config": [_P, C.c_uint32, C.c_int32],
    "kmh_gemm_autotune": [_P, C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_char_p],
    "kmh_gemm_tuning_load": [_P, C.c_char_p],
    "kmh_gemm_tuning_size": [_P, C.POINTER(C.c_int32)],
    "kmh_hodlr_build": [_P, _P, _I64, C.c_int32, C.POINTER(kmh_kernel), C.POINTER(kmh_build_opts), C.POINTER(_P)],
    "kmh_hodlr_build_dev": [_P, _P, _I64, C.c_int32, C.POINTER(kmh_kernel), C.POINTER(kmh_build_opts), C.POINTER(_P)],
    "kmh_hodlr_destroy": [_P],
    "kmh_hodlr_info": [_P, C.POINTER(kmh_info)],
    "kmh_hodlr_leaf_offsets": [_P, C.POINTER(C.c_int64)],
    "kmh_hodlr_ranks": [_P, C.POINTER(C.c_int32)],
    "kmh_hodlr_full": [_P, _P],
    "kmh_hodlr_block_uv": [_P, C.c_int32, C.c_int32, _P, _P],
    "kmh_hodlr_matvec": [_P, _P, _P, _I64],
    "kmh_hodlr_matvec_dev": [_P, _P, _P, _I64],
    "kmh_hodlr_factorize": [_P],
    "kmh_hodlr_factor_apply": [_P, C.c_int32, _P, _P, _I64],
    "kmh_hodlr_factor_apply_dev": [_P, C.c_int32, _P, _P, _I64],
    "kmh_hodlr_solve": [_P, _P, _P, _I64],
    "kmh_hodlr_solve_dev": [_P, _P, _P, _I64],
    "kmh_hodlr_solve_dev_noalloc": [_P, _P, _P, _I64],
    "kmh_hodlr_solve_graphs": [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "kmh_hodlr_logdet": [_P, _D],
    "kmh_hodlr_refinement": [_P, C.POINTER(C.c_int32), _D],
    "kmh_stoch_gradient": [_P, C.POINTER(kmh_kernel), C.c_int32, _P, _P, _I64, _P],
    "kmh_stoch_gradient_dev": [_P, C.POINTER(kmh_kernel), C.c_int32, _P, _P, _I64, _P],
    "kmh_hodlr_build_batch": [_P, C.c_int32, C.POINTER(kmh_build_req), C.POINTER(_P), C.POINTER(C.c_int32)],
    "kmh_hodlr_factorize_batch": [_P, C.c_int32, C.POINTER(_P), C.POINTER(C.c_int32)],
    "kmh_hodlr_solve_batch_dev": [_P, C.c_int32, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_int64)],
    "kmh_stoch_profile_gradient": [_P, C.POINTER(kmh_kernel), C.c_int32, _P, _P, _I64, _P],
    "kmh_stoch_fisher": [_P, C.POINTER(kmh_kernel), C.c_int32, _P, _I64, _P],
    "kmh_stoch_hessian": [_P, C.POINTER(kmh_kernel), C.c_int32, C.POINTER(kmh_kernel), _P, _P, _I64, C.c_int32,
                          _P],
    "kmh_server_create": [_P, C.c_int32, C.c_uint64, C.POINTER(_P)],
    "kmh_server_set_lanes": [_P, C.c_int32],
    "kmh_server_reserve": [_P, C.c_int64, _P, C.c_int32, C.c_int32],
    "kmh_server_destroy": [_P],
    "kmh_server_submit": [_P, C.POINTER(kmh_serve_req), C.POINTER(C.c_int64)],
    "kmh_server_step": [_P, C.POINTER(C.c_int32)],
    "kmh_server_pending": [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)],
    "kmh_server_result": [_P, C.c_int64, C.POINTER(kmh_serve_result)],
    "kmh_server_factor": [_P, C.c_int64, C.POINTER(_P)],
    "kmh_server_release": [_P, C.c_int64],
    "kmh_hodlr_synchronize": [_P],
    "kmh_kd_order": [_P, _I64, C.c_int32, C.c_int32, _P],
    "kmh_rhodlr_build": [_P, _P, _I64, _P, _I64, C.c_int32, C.POINTER(kmh_kernel), C.POINTER(kmh_build_opts),
                         C.POINTER(C.c_void_p)],
    "kmh_rhodlr_levels": [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "kmh_rhodlr_ranks": [_P, _P],
    "kmh_rhodlr_matvec": [_P, _P, _P, _I64],
    "kmh_rhodlr_matvec_dev": [_P, _P, _P, _I64],
    "kmh_rhodlr_destroy": [_P],
}

# include/kmhodlr_testing.h
TEST_SIGNATURES = {
    "kmh_test_gemm": [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, C.c_int32, _P,
                      C.c_int32, _P, C.c_int32, C.c_double, C.c_double, C.c_int32],
    "kmh_test_potrf": [_P, C.c_int32, _P, C.c_int32],
    "kmh_test_trsm": [_P, C.c_int32, C.c_int32, _P, C.c_int32, C.c_int32, _P, C.c_int32],
    "kmh_test_qr": [_P, C.c_int32, C.c_int32, _P, C.c_int32, _P, C.c_int32],
    "kmh_test_capacitance": [_P, C.c_int32, _P, _P, C.c_int32, _P, _P, C.c_int32, _D],
    "kmh_test_arena": [_P, C.c_int32, C.c_int32, C.POINTER(C.c_int64)],
    "kmh_test_profile_spans": [_P, C.c_int32, C.POINTER(C.c_int32), _D, _D, _D, C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32)],
    "kmh_test_fail_scratch": [_P, C.c_int64],
    "kmh_test_scratch_live": [_P, C.POINTER(C.c_uint64)],
    "kmh_test_pool_allocs": [_P, C.POINTER(C.c_uint64)],
    "kmh_test_arena_shrink": [_P, C.c_uint64],
}


class LibraryMissing(RuntimeError):
    pass


class Lib:
    """Loaded libkmhodlr with typed entry points."""

    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise LibraryMissing(
                "libkmhodlr.so is not built (%s); run __graft_entry__.build() or "
                "python kernelmatrices.jl_amd/kmhodlr/build.py" % path)
        self.path = path
        self.dll = C.CDLL(path)
        for name, args in list(SIGNATURES.items()) + list(TEST_SIGNATURES.items()):
            f = getattr(self.dll, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "kmh_last_error" else C.c_int
            setattr(self, name, f)

    def last_error(self):
        m = self.kmh_last_error()
        return m.decode() if m else ""

    def device_count(self):
        n = C.c_int(0)
        self.kmh_device_count(C.byref(n))
        return n.value


_LIB = None


def load():
    global _LIB
    if _LIB is None:
        _LIB = Lib()
    return _LIB
