"""kmhodlr: MI355X-native HODLR Gaussian-process engine (Python binding).

The package mirrors the KernelMatrices.jl HODLR API; the work runs in
libkmhodlr.so (HIP, gfx950) through its C ABI.
"""
from . import _lib
from ._lib import load, LibraryMissing
from .api import (Context, ErrorException, FactorHODLR, FixedLevel, HBMBudgetExceeded, Kernel,
                  KernelHODLR, KernelMatrix, LogLevel, PosDefException, RKernelHODLR, cauchy_kernel,
                  default_context,
                  full, kd_order, maxlikdata, ldiv, logdet, matern12, matern32, negloglik, profile_negloglik,
                  rational_kernel, stoch_fisher, stoch_gradient, stoch_hessian, stoch_profile_gradient,
                  symmetricfactorize)


from . import maxlik


class HODLR:
    """Namespace mirroring `KernelMatrices.HODLR`."""
    KernelHODLR = KernelHODLR
    RKernelHODLR = RKernelHODLR
    FixedLevel = FixedLevel
    LogLevel = LogLevel
    symmetricfactorize = staticmethod(symmetricfactorize)
    stoch_gradient = staticmethod(stoch_gradient)
    stoch_fisher = staticmethod(stoch_fisher)
    stoch_hessian = staticmethod(stoch_hessian)
    stoch_profile_gradient = staticmethod(stoch_profile_gradient)
    profile_negloglik = staticmethod(profile_negloglik)
    Maxlikopts = maxlik.Maxlikopts
    Maxlikdata = maxlik.Maxlikdata
    maxlikopts = staticmethod(maxlik.maxlikopts)
    maxlikdata = staticmethod(maxlik.maxlikdata)
    nll_objective = staticmethod(maxlik.nll_objective)
    nlpl_objective = staticmethod(maxlik.nlpl_objective)
    nll_gradient = staticmethod(maxlik.nll_gradient)
    nll_hessian = staticmethod(maxlik.nll_hessian)
    fisher_matrix = staticmethod(maxlik.fisher_matrix)
    gpsimulate = staticmethod(maxlik.gpsimulate)
    full = staticmethod(full)
    logdet = staticmethod(logdet)
    negloglik = staticmethod(negloglik)
