"""Host-side mirror of the reference's HODLR API over the C ABI.

Names, argument meaning and error behaviour follow KernelMatrices.jl 0.9.0:
  KernelMatrix(pts, pts2, parms, kernel)            src/structstypes.jl:17-23
  HODLR.FixedLevel / HODLR.LogLevel                 src/HODLR/src/structstypes.jl:30-37
  HODLR.KernelHODLR(K, ep, maxrank, lvl; nystrom)   src/HODLR/src/constructor.jl:6-58
  HODLR.symmetricfactorize!(HK)                     src/HODLR/src/symfact.jl:6-50
  HK * v, HK \\ v, logdet(HK), full(HK)              src/HODLR/src/baseoverloads.jl:152-326
  mul!/ldiv!(t, HK.W or adjoint(HK.W), v)           src/HODLR/src/baseoverloads.jl:95-150
  HODLR.stoch_gradient(K, HK, y, dfuns, vecs)       src/HODLR/src/stochasticgradient.jl:30-45
  negloglik(HK, y)                                  src/HODLR/src/maxlikfunctions.jl:2-9
Every computation runs in libkmhodlr on the GPU; this module only marshals
arguments.  Without the built library or a HIP device it raises.
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L


class ErrorException(RuntimeError):
    """Julia's ErrorException (the reference's `error(...)`)."""


class PosDefException(ErrorException):
    """LinearAlgebra.PosDefException from a failed Cholesky."""


class HBMBudgetExceeded(MemoryError):
    pass


def _check(lib, rc):
    if rc == L.KMH_OK:
        return
    msg = lib.last_error()
    if rc == L.KMH_E_NOT_PD:
        raise PosDefException(msg)
    if rc == L.KMH_E_OOM:
        raise MemoryError(msg)
    if rc == L.KMH_E_BUDGET:
        raise HBMBudgetExceeded(msg)
    raise ErrorException(msg)


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------

class Kernel:
    """A covariance function the C ABI can evaluate: family + derivative code
    + nugget added at equal points.  Parameters travel with KernelMatrix."""

    def __init__(self, family, nugget=0.0, deriv=L.KMH_D0):
        self.family = int(family)
        self.nugget = float(nugget)
        self.deriv = int(deriv)

    def derivative(self, code):
        return Kernel(self.family, 0.0, code)

    @property
    def d1(self):
        return self.derivative(L.KMH_D1)

    @property
    def d2(self):
        return self.derivative(L.KMH_D2)

    @property
    def d12(self):
        return self.derivative(L.KMH_D12)

    @property
    def d22(self):
        return self.derivative(L.KMH_D22)

    def desc(self, parms):
        k = L.kmh_kernel()
        k.family = self.family
        k.deriv = self.deriv
        p = list(parms) + [0.0] * (4 - len(parms))
        for i in range(4):
            k.params[i] = float(p[i])
        k.nugget = self.nugget
        return k


def rational_kernel(nugget=1.0):
    """p1^2/(1+(r/p2)^2)^2 + nugget*[x==y]  (README.md:42-48)."""
    return Kernel(L.KMH_RATIONAL2, nugget)


def cauchy_kernel(nugget=1e-12):
    """p1^2/(1+(p2 r)^2) + nugget*[x==y]  (example/basicexample.jl:62-68)."""
    return Kernel(L.KMH_CAUCHY, nugget)


def matern12(nugget=0.0):
    """t0*exp(-r/rho) + nugget*[x==y]  (mtn_kernfun, nu = 1/2)."""
    return Kernel(L.KMH_MATERN12, nugget)


def matern32(nugget=0.0):
    """t0*(1+sqrt3 r/rho)exp(-sqrt3 r/rho) + nugget*[x==y]  (nu = 3/2)."""
    return Kernel(L.KMH_MATERN32, nugget)


class KernelMatrix:
    """Lazy kernel matrix K[i,j] = kernel(x1[i], x2[j], parms)."""

    def __init__(self, x1, x2, parms, kernel):
        self.x1 = np.ascontiguousarray(x1, dtype=np.float64)
        self.x2 = self.x1 if x2 is x1 else np.ascontiguousarray(x2, dtype=np.float64)
        if self.x1.ndim != 2 or self.x2.ndim != 2 or self.x1.shape[1] != self.x2.shape[1]:
            raise ErrorException("Eltype of pts and pts2 must agree.")
        self.parms = tuple(float(p) for p in parms)
        self.kernel = kernel

    @property
    def shape(self):
        return (self.x1.shape[0], self.x2.shape[0])

    def size(self, j=None):
        return self.shape if j is None else self.shape[j - 1]


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gemm_tuning_gfx950.txt")


class Context:
    """A device, a stream and an HBM budget for resident factors.  The GEMM
    kernel-class table written by kmhodlr.autotune (TUNING_FILE) is loaded
    when present; tuning=None keeps the built-in shape rules."""

    def __init__(self, device=0, hbm_budget_bytes=0, tuning=TUNING_FILE):
        self.lib = L.load()
        p = C.c_void_p()
        _check(self.lib, self.lib.kmh_ctx_create(int(device), int(hbm_budget_bytes), C.byref(p)))
        self.ptr = p
        if tuning and os.path.exists(tuning):
            _check(self.lib, self.lib.kmh_gemm_tuning_load(self.ptr, tuning.encode()))

    def synchronize(self):
        _check(self.lib, self.lib.kmh_ctx_synchronize(self.ptr))

    def bytes_in_use(self):
        v = C.c_uint64(0)
        _check(self.lib, self.lib.kmh_ctx_bytes_in_use(self.ptr, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.kmh_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT = None


def default_context():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


# ---------------------------------------------------------------------------
# HODLR
# ---------------------------------------------------------------------------

class FixedLevel:
    def __init__(self, lv):
        self.lv = int(lv)
        self.kind = L.KMH_FIXED_LEVEL


class LogLevel:
    def __init__(self, lv):
        self.lv = int(lv)
        self.kind = L.KMH_LOG_LEVEL


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def kd_order(pts, leafsize=8):
    """Permutation that KD-orders the n x d points (host code of the
    library; see kmh_kd_order)."""
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(len(pts), -1))
    n, d = pts.shape
    perm = np.empty(n, dtype=np.int64)
    lib = L.load()
    _check(lib, lib.kmh_kd_order(pts.ctypes.data, n, d, int(leafsize), perm.ctypes.data))
    return perm


def maxlikdata(pts, data, sortmethod="KDTree", warn=True):
    """maxlikdata(; pts, data, sortmethod) (HODLR/src/constructor.jl:171-181):
    the points in KD order and the data reordered with them."""
    pts = np.asarray(pts, dtype=np.float64)
    data = np.asarray(data, dtype=np.float64)
    if sortmethod in ("Nothing", None):
        if warn:
            import warnings
            warnings.warn("Not sorting data can adversely affect approximation quality...")
        return pts, data
    if sortmethod != "KDTree":
        raise ErrorException("The options are :Nothing and :KDTree")
    perm = kd_order(pts)
    return np.ascontiguousarray(pts[perm]), np.ascontiguousarray(data[perm])


def _as_cols(x, n):
    """Vector or matrix -> column-major n x k buffer and k."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        if x.shape[0] != n:
            raise ErrorException("Incorrect vector size.")
        return np.ascontiguousarray(x), 1, True
    if x.shape[0] != n:
        raise ErrorException("Incorrect matrix size.")
    return np.asfortranarray(x), x.shape[1], False


class KernelHODLR:
    """HODLR.KernelHODLR(K, ep, maxrank, lvl; nystrom=false, plel=false,
    sorted=false): built on the GPU (constructor.jl:6-58).  `plel` is
    accepted and ignored (the GPU runs every block in parallel)."""

    def __init__(self, K, ep, maxrank, lvl, nystrom=False, plel=False, sorted=False, ctx=None, mixed=False):
        if not (K.x1 is K.x2 or np.array_equal(K.x1, K.x2)):
            raise ErrorException("This function builds symmetric matrices. For non-symmetric matrices, "
                                 "please use the recursive RKernelHODLR, which will only allow matvecs.")
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.K = K
        o = L.kmh_build_opts()
        o.method = L.KMH_NYSTROM if nystrom else L.KMH_ACA
        o.level_kind = lvl.kind
        o.level = lvl.lv
        o.maxrank = int(maxrank)
        o.tol = float(ep)
        # mixed: False (FP64), True or "panels" (KMH_MIXED: FP32 panels,
        # cascade panels and Q factors) or "q" (KMH_MIXED_Q: FP32 Q factors
        # and panel QR); FP64 leaves / capacitance / logdet, solves refined
        # to FP64 (no reference counterpart)
        o.precision = {False: L.KMH_FP64, True: L.KMH_MIXED, "panels": L.KMH_MIXED,
                       "q": L.KMH_MIXED_Q}[mixed]
        kd = K.kernel.desc(K.parms)
        p = C.c_void_p()
        pts = _f64(K.x1)
        _check(self.lib, self.lib.kmh_hodlr_build(self.ctx.ptr, pts.ctypes.data, pts.shape[0], pts.shape[1],
                                                  C.byref(kd), C.byref(o), C.byref(p)))
        self.ptr = p
        self.ep, self.mrnk, self.nys = float(ep), int(maxrank), bool(nystrom)
        info = self.info()
        self.n = info.n
        self.lvl = info.levels
        offs = (C.c_int64 * (info.nleaves + 1))()
        _check(self.lib, self.lib.kmh_hodlr_leaf_offsets(self.ptr, offs))
        self._leaf_off = list(offs)

    # -- structure -------------------------------------------------------
    def info(self):
        i = L.kmh_info()
        _check(self.lib, self.lib.kmh_hodlr_info(self.ptr, C.byref(i)))
        return i

    @property
    def leafindices(self):
        """1-based [a, b, c, d] per leaf (indices.jl:29-37)."""
        o = self._leaf_off
        return [(o[i] + 1, o[i + 1], o[i] + 1, o[i + 1]) for i in range(len(o) - 1)]

    @property
    def nonleafindices(self):
        """1-based [a, b, c, d] per level (indices.jl:40-51)."""
        o = self._leaf_off
        out = []
        for j in range(1, self.lvl + 1):
            skp, half = 2 ** j, 2 ** (j - 1)
            out.append([(o[i + half] + 1, o[i + skp], o[i] + 1, o[i + half])
                        for i in range(0, len(o) - 2, skp)])
        return out

    def ranks(self):
        nb = 2 ** self.lvl - 1
        r = (C.c_int32 * max(1, nb))()
        _check(self.lib, self.lib.kmh_hodlr_ranks(self.ptr, r))
        out, q = [], 0
        for j in range(1, self.lvl + 1):
            m = 2 ** (self.lvl - j)
            out.append(list(r[q:q + m]))
            q += m
        return out

    @property
    def is_factorized(self):
        return bool(self.info().factorized)

    @property
    def U(self):
        """Before factorization: per level, per block U matrices; None after
        (symfact.jl:45-46)."""
        return None if self.is_factorized else self._uv(0)

    @property
    def V(self):
        return None if self.is_factorized else self._uv(1)

    def _uv(self, which):
        out = []
        for j, lev in enumerate(self.nonleafindices, start=1):
            row = []
            for k, (a, b, c, d) in enumerate(lev):
                r = self.ranks()[j - 1][k]
                U = np.zeros((b - a + 1, r), order="F")
                V = np.zeros((d - c + 1, r), order="F")
                _check(self.lib, self.lib.kmh_hodlr_block_uv(self.ptr, j, k, U.ctypes.data, V.ctypes.data))
                row.append(U if which == 0 else V)
            out.append(row)
        return out

    @property
    def W(self):
        return FactorHODLR(self) if self.is_factorized else None

    # -- algebra ---------------------------------------------------------
    def full(self):
        out = np.zeros((self.n, self.n), order="F")
        _check(self.lib, self.lib.kmh_hodlr_full(self.ptr, out.ctypes.data))
        return out

    def matvec(self, x):
        xb, k, vec = _as_cols(x, self.n)
        y = np.zeros_like(xb)
        _check(self.lib, self.lib.kmh_hodlr_matvec(self.ptr, xb.ctypes.data, y.ctypes.data, k))
        return y

    __matmul__ = matvec

    def __mul__(self, x):
        return self.matvec(x)

    def solve(self, b):
        bb, k, vec = _as_cols(b, self.n)
        x = np.zeros_like(bb)
        _check(self.lib, self.lib.kmh_hodlr_solve(self.ptr, bb.ctypes.data, x.ctypes.data, k))
        return x

    def logdet(self):
        v = C.c_double(0.0)
        _check(self.lib, self.lib.kmh_hodlr_logdet(self.ptr, C.byref(v)))
        return v.value

    def factor_apply(self, op, x):
        xb, k, vec = _as_cols(x, self.n)
        y = np.zeros_like(xb)
        _check(self.lib, self.lib.kmh_hodlr_factor_apply(self.ptr, int(op), xb.ctypes.data, y.ctypes.data, k))
        return y

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.kmh_hodlr_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RKernelHODLR:
    """HODLR.RKernelHODLR(K, tol, maxrank=0, lvl=LogLevel(7))
    (constructor.jl:130-161): recursive HODLR of K(x1, x2) for matvecs; x1
    and x2 may differ, also in size.  Rows and columns split at floor(len/2);
    both off-diagonal blocks of every node are ACA-compressed, the last
    level's diagonal blocks are dense.  Level 0 keeps the whole matrix dense
    (the reference returns full(K) there, :141)."""

    def __init__(self, K, tol, maxrank=0, lvl=None, ctx=None):
        lvl = lvl if lvl is not None else LogLevel(7)
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.K = K
        o = L.kmh_build_opts()
        o.method = L.KMH_ACA
        o.level_kind = lvl.kind
        o.level = lvl.lv
        o.maxrank = int(maxrank)
        o.tol = float(tol)
        o.precision = L.KMH_FP64
        kd = K.kernel.desc(K.parms)
        p1, p2 = _f64(K.x1), _f64(K.x2)
        p = C.c_void_p()
        _check(self.lib, self.lib.kmh_rhodlr_build(self.ctx.ptr, p1.ctypes.data, p1.shape[0], p2.ctypes.data,
                                                   p2.shape[0], p1.shape[1], C.byref(kd), C.byref(o), C.byref(p)))
        self.ptr = p
        lv, n1, n2 = C.c_int32(), C.c_int64(), C.c_int64()
        _check(self.lib, self.lib.kmh_rhodlr_levels(self.ptr, C.byref(lv), C.byref(n1), C.byref(n2)))
        self.levels, self.shape = lv.value, (n1.value, n2.value)

    def size(self, j=None):
        return self.shape if j is None else self.shape[j - 1]

    def ranks(self):
        """[(rank K12, rank K21) per node] per depth, root first."""
        nb = 2 ** self.levels - 1
        r = (C.c_int32 * max(1, 2 * nb))()
        _check(self.lib, self.lib.kmh_rhodlr_ranks(self.ptr, r))
        out, q = [], 0
        for d in range(self.levels):
            out.append([(r[q + 2 * t], r[q + 2 * t + 1]) for t in range(2 ** d)])
            q += 2 ** (d + 1)
        return out

    def mul(self, x):
        """R * x (baseoverloads.jl:299-303) for a vector or an n2 x k matrix."""
        xb, k, vec = _as_cols(x, self.shape[1])
        y = np.zeros((self.shape[0],) if vec else (self.shape[0], k), order="F")
        _check(self.lib, self.lib.kmh_rhodlr_matvec(self.ptr, xb.ctypes.data, y.ctypes.data, k))
        return y

    __matmul__ = mul

    def __mul__(self, x):
        return self.mul(x)

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.kmh_rhodlr_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FactorHODLR:
    """HK.W with K = W W^T: W*x, W\\x and their adjoints."""

    def __init__(self, HK, adjoint=False):
        self.HK = HK
        self.adj = adjoint

    def adjoint(self):
        return FactorHODLR(self.HK, not self.adj)

    @property
    def T(self):
        return self.adjoint()

    def mul(self, x):
        return self.HK.factor_apply(L.KMH_WT if self.adj else L.KMH_W, x)

    def ldiv(self, x):
        return self.HK.factor_apply(L.KMH_WTINV if self.adj else L.KMH_WINV, x)

    __matmul__ = mul


def symmetricfactorize(HK, plel=False, verbose=False):
    """HODLR.symmetricfactorize!(HK) (symfact.jl:6-50), in place."""
    _check(HK.lib, HK.lib.kmh_hodlr_factorize(HK.ptr))


def full(HK):
    return HK.full()


def logdet(HK):
    return HK.logdet()


def ldiv(HK, b):
    return HK.solve(b)


def negloglik(HK, y):
    """maxlikfunctions.jl:2-9."""
    if not HK.is_factorized:
        raise ErrorException("The HODLR matrix needs to be factorized for this.")
    y = np.asarray(y, dtype=np.float64)
    return 0.5 * HK.logdet() + 0.5 * float(np.dot(y, HK.solve(y)))


def _is_device_tensor(x):
    """A float64 torch tensor in GPU memory (torch is imported lazily)."""
    mod = type(x).__module__
    if not mod.startswith("torch"):
        return False
    import torch
    return isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64


def stoch_gradient(K, HK, dat, dfuns, vecs, plel=False, verbose=False, shuffle=False):
    """HODLR.stoch_gradient (stochasticgradient.jl:30-45).  dfuns are
    derivative kernels (Kernel.derivative); vecs a list of n-vectors or an
    n x nv array of SAA vectors.  If dat and vecs are float64 torch tensors
    on the GPU (vecs n x nv, column-major i.e. vecs.t() contiguous), they are
    used in place (kmh_stoch_gradient_dev)."""
    if not HK.is_factorized:
        raise ErrorException("The matrix needs to be factorized for this to work.")
    if len(dfuns) != len(K.parms):
        raise ErrorException("You didn't supply the right number of gradient funs.")
    if shuffle:
        raise ErrorException("shuffle=true is not supported (the reference's path references an undefined opts)")
    n = HK.n
    arr = (L.kmh_kernel * len(dfuns))()
    for i, dk in enumerate(dfuns):
        arr[i] = dk.desc(K.parms)
    g = np.zeros(len(dfuns))
    if _is_device_tensor(dat) and _is_device_tensor(vecs):
        if vecs.dim() != 2 or vecs.shape[0] != n or vecs.stride(0) != 1 or vecs.stride(1) != n:
            raise ErrorException("Incorrect SAA Vector size.")
        if dat.dim() != 1 or dat.shape[0] != n or not dat.is_contiguous():
            raise ErrorException("data vector must be a contiguous n-vector")
        _check(HK.lib, HK.lib.kmh_stoch_gradient_dev(HK.ptr, arr, len(dfuns), dat.data_ptr(), vecs.data_ptr(),
                                                     vecs.shape[1], g.ctypes.data))
        return g
    if isinstance(vecs, np.ndarray) and vecs.ndim == 2:
        saa = np.asfortranarray(vecs, dtype=np.float64)
    else:
        saa = np.asfortranarray(np.stack([np.asarray(v, dtype=np.float64) for v in vecs], axis=1))
    if saa.shape[0] != n:
        raise ErrorException("Incorrect SAA Vector size.")
    y = _f64(dat)
    _check(HK.lib, HK.lib.kmh_stoch_gradient(HK.ptr, arr, len(dfuns), y.ctypes.data, saa.ctypes.data,
                                             saa.shape[1], g.ctypes.data))
    return g


def stoch_profile_gradient(K, HK, dat, dfuns, vecs, plel=False, verbose=False, shuffle=False):
    """HODLR.stoch_profile_gradient (stochasticgradient.jl:48-86): gradient of
    the profile negative log-likelihood 0.5 (logdet K + n log(y'K^-1 y))."""
    if not HK.is_factorized:
        raise ErrorException("The matrix needs to be factorized for this to work.")
    if len(dfuns) != len(K.parms):
        raise ErrorException("You didn't supply the right number of gradient funs.")
    if shuffle:
        raise ErrorException("shuffle=true is not supported (the reference's path references an undefined opts)")
    saa = _saa_matrix(vecs, HK.n)
    y = _f64(dat)
    arr = (L.kmh_kernel * len(dfuns))(*[dk.desc(K.parms) for dk in dfuns])
    g = np.zeros(len(dfuns))
    _check(HK.lib, HK.lib.kmh_stoch_profile_gradient(HK.ptr, arr, len(dfuns), y.ctypes.data, saa.ctypes.data,
                                                     saa.shape[1], g.ctypes.data))
    return g


def profile_negloglik(HK, y):
    """maxlikfunctions.jl:11-19: 0.5 (logdet K + n log(y' K^-1 y))."""
    if not HK.is_factorized:
        raise ErrorException("The HODLR matrix needs to be factorized for this.")
    y = np.asarray(y, dtype=np.float64)
    return 0.5 * (HK.logdet() + len(y) * float(np.log(np.dot(y, HK.solve(y)))))


def _saa_matrix(vecs, n):
    if isinstance(vecs, np.ndarray) and vecs.ndim == 2:
        saa = np.asfortranarray(vecs, dtype=np.float64)
    else:
        saa = np.asfortranarray(np.stack([np.asarray(v, dtype=np.float64) for v in vecs], axis=1))
    if saa.shape[0] != n:
        raise ErrorException("Incorrect SAA vector size.")
    return saa


def stoch_fisher(K, HK, dfuns, vecs, plel=False, shuffle=False):
    """HODLR.stoch_fisher (stochastichessian.jl:91-115) with the derivative
    HODLRs given by their kernels."""
    if not HK.is_factorized:
        raise ErrorException("The matrix needs to be factorized.")
    if len(dfuns) != len(K.parms):
        raise ErrorException("You didn't supply the right number of gradient funs.")
    saa = _saa_matrix(vecs, HK.n)
    p = len(dfuns)
    arr = (L.kmh_kernel * p)(*[dk.desc(K.parms) for dk in dfuns])
    F = np.zeros((p, p), order="F")
    _check(HK.lib, HK.lib.kmh_stoch_fisher(HK.ptr, arr, p, saa.ctypes.data, saa.shape[1], F.ctypes.data))
    return F


def stoch_hessian(K, HK, dat, d1funs, d2funs, vecs, plel=False, verbose=False, shuffle=False,
                  profile=False, exact_second=False):
    """HODLR.stoch_hessian (stochastichessian.jl:117-147).  d2funs[j][k-j] is
    the (j,k) second-derivative kernel or None (the reference's
    ZeroFunction).  exact_second=True uses the exact second derivative of
    the Nystrom blocks instead of the reference's block terms."""
    if profile:
        raise ErrorException("the profile-likelihood Hessian is not provided")
    if not HK.is_factorized:
        raise ErrorException("The matrix needs to be factorized.")
    p = len(d1funs)
    if p != len(K.parms):
        raise ErrorException("You didn't supply the right number of gradient funs.")
    flat = [f for row in d2funs for f in row]
    if len(flat) != p * (p + 1) // 2:
        raise ErrorException("You didn't supply the right number of Hessian funs.")
    saa = _saa_matrix(vecs, HK.n)
    y = _f64(dat)
    a1 = (L.kmh_kernel * p)(*[dk.desc(K.parms) for dk in d1funs])
    a2 = (L.kmh_kernel * len(flat))()
    for i, f in enumerate(flat):
        if f is None:
            a2[i].family = 0
        else:
            a2[i] = f.desc(K.parms)
    Hm = np.zeros((p, p), order="F")
    _check(HK.lib, HK.lib.kmh_stoch_hessian(HK.ptr, a1, p, a2, y.ctypes.data, saa.ctypes.data, saa.shape[1],
                                            1 if exact_second else 0, Hm.ctypes.data))
    return Hm
