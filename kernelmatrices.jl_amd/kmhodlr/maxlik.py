"""Gaussian-process likelihood drivers over the HODLR path
(HODLR/src/maxlikfunctions.jl, Maxlikopts/Maxlikdata in
HODLR/src/structstypes.jl:97-114, maxlikopts in constructor.jl:164-169).

Every objective builds KernelHODLR(K, 0.0, opts.mrnk, opts.lvl,
nystrom=True) from the (KD-ordered) data, factorizes it and evaluates the
requested quantity on the GPU through the C ABI, as the reference's
functions do on the CPU.
"""
import numpy as np

from . import api
from .api import ErrorException


class Maxlikopts:
    """Options of the likelihood drivers (structstypes.jl:97-108).

    kernfun: an api.Kernel (family + nugget); dfuns: its first-derivative
    kernels (Kernel.d1, Kernel.d2); d2funs: upper-triangle rows of
    second-derivative kernels (None entries are the reference's
    ZeroFunction); lvl: FixedLevel/LogLevel; mrnk: Nystrom rank; saav: SAA
    vectors (list of n-vectors or an n x nv array)."""

    def __init__(self, kernfun, dfuns, d2funs, lvl, mrnk, saav, apll=False, fpll=False, verb=False,
                 saa_fix=True):
        self.kernfun = kernfun
        self.dfuns = list(dfuns)
        self.d2funs = d2funs
        self.lvl = lvl
        self.mrnk = int(mrnk)
        self.saav = saav
        self.apll = bool(apll)
        self.fpll = bool(fpll)
        self.verb = bool(verb)
        self.saa_fix = bool(saa_fix)


def maxlikopts(kernfun, level, rank, saavecs=None, fix_saa=True, dfuns=(), d2funs=None, par_assem=False,
               par_factor=False, verbose=False):
    """maxlikopts(; kernfun, level, rank, ...) (constructor.jl:164-169)."""
    return Maxlikopts(kernfun, dfuns, d2funs, level, rank, saavecs if saavecs is not None else [], par_assem,
                      par_factor, verbose, fix_saa)


class Maxlikdata:
    """Points and data in the order the HODLR expects (structstypes.jl:111-114)."""

    def __init__(self, pts_s, dat_s):
        self.pts_s = np.ascontiguousarray(pts_s, dtype=np.float64)
        self.dat_s = np.ascontiguousarray(dat_s, dtype=np.float64)


def maxlikdata(pts, data, sortmethod="KDTree", warn=True):
    """maxlikdata(; pts, data, sortmethod) (constructor.jl:171-200)."""
    p, d = api.maxlikdata(pts, data, sortmethod=sortmethod, warn=warn)
    return Maxlikdata(p, d)


def _factorized(prms, data, opts):
    K = api.KernelMatrix(data.pts_s, data.pts_s, prms, opts.kernfun)
    HK = api.KernelHODLR(K, 0.0, opts.mrnk, opts.lvl, nystrom=True, plel=opts.apll, sorted=True)
    api.symmetricfactorize(HK, plel=opts.fpll)
    return K, HK


def _check_saa(opts):
    if not opts.saa_fix:
        raise ErrorException("saa_fix=false is not supported (the reference's shuffle path references an "
                             "undefined opts)")


def nll_objective(prms, grad, data, opts):
    """maxlikfunctions.jl:29-50: negative log-likelihood; fills grad (a
    writable array of length len(prms), or empty) with the stochastic
    gradient."""
    K, HK = _factorized(prms, data, opts)
    nll = api.negloglik(HK, data.dat_s)
    if grad is not None and len(grad) > 0:
        _check_saa(opts)
        grad[:] = api.stoch_gradient(K, HK, data.dat_s, opts.dfuns, opts.saav)
    return nll


def nlpl_objective(prms, grad, data, opts):
    """maxlikfunctions.jl:60-81: negative log PROFILE likelihood and its
    stochastic gradient."""
    K, HK = _factorized(prms, data, opts)
    nll = api.profile_negloglik(HK, data.dat_s)
    if grad is not None and len(grad) > 0:
        _check_saa(opts)
        grad[:] = api.stoch_profile_gradient(K, HK, data.dat_s, opts.dfuns, opts.saav)
    return nll


def scaleparm_mle(prms, data, opts):
    """maxlikfunctions.jl:21-27: y' K^-1 y / n (the profiled-out scale)."""
    _, HK = _factorized(prms, data, opts)
    y = data.dat_s
    return float(np.dot(y, HK.solve(y))) / len(y)


nlpl_scale = scaleparm_mle   # maxlikfunctions.jl:52-58 computes the same quantity


def nll_gradient(prms, data, opts):
    """maxlikfunctions.jl:83-98."""
    _check_saa(opts)
    K, HK = _factorized(prms, data, opts)
    return api.stoch_gradient(K, HK, data.dat_s, opts.dfuns, opts.saav)


def nll_hessian(prms, data, opts, exact_second=False):
    """maxlikfunctions.jl:100-115 (exact_second as api.stoch_hessian)."""
    if opts.d2funs is None:
        raise ErrorException("You need to supply second derivatives for this.")
    _check_saa(opts)
    K, HK = _factorized(prms, data, opts)
    return api.stoch_hessian(K, HK, data.dat_s, opts.dfuns, opts.d2funs, opts.saav, exact_second=exact_second)


def nlpl_hessian(prms, data, opts):
    """maxlikfunctions.jl:118-133 (stoch_hessian(...; profile=true))."""
    raise ErrorException("the profile-likelihood Hessian is not provided")


def fisher_matrix(prms, data, opts):
    """maxlikfunctions.jl:135-146."""
    _check_saa(opts)
    K, HK = _factorized(prms, data, opts)
    return api.stoch_fisher(K, HK, opts.dfuns, opts.saav)


def gpsimulate(kernel, locs, parms, opts=None, exact=False, kdtreesort=False, rng=None, z=None):
    """maxlikfunctions.jl:148-173: one draw of the GP at locs, W z with
    z ~ N(0, I) and W the HODLR symmetric factor (or the dense Cholesky
    factor when exact), returned as Maxlikdata in the (optionally KD-sorted)
    point order.  z may be given; otherwise it is drawn from rng (numpy)."""
    locs = np.asarray(locs, dtype=np.float64)
    if kdtreesort:
        locs = locs[api.kd_order(locs)]
    else:
        import warnings
        warnings.warn("Not sorting data can adversely affect approximation quality...")
    n = locs.shape[0]
    if z is None:
        z = (rng if rng is not None else np.random.default_rng()).standard_normal(n)
    z = np.ascontiguousarray(z, dtype=np.float64)
    K = api.KernelMatrix(locs, locs, parms, kernel)
    if exact:
        # one leaf: the factorization is the dense Cholesky of the whole matrix
        HK = api.KernelHODLR(K, 0.0, 0, api.FixedLevel(0))
    else:
        if opts is None:
            raise ErrorException("Supply HODLR options for approximated simulations.")
        HK = api.KernelHODLR(K, 0.0, opts.mrnk, opts.lvl, nystrom=True, plel=opts.apll)
    api.symmetricfactorize(HK)
    out = HK.W.mul(z)
    return Maxlikdata(locs, out)


def gpsimulate_box(kernel, parms, n, dim, boxsz, opts=None, exact=False, rng=None):
    """maxlikfunctions.jl:175-179: n uniform points in [0, boxsz]^dim, KD-sorted."""
    rng = rng if rng is not None else np.random.default_rng()
    pts = rng.random((n, dim)) * boxsz
    return gpsimulate(kernel, pts, parms, opts, exact=exact, kdtreesort=True, rng=rng)
