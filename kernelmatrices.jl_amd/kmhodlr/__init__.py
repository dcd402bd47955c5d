"""kmhodlr: MI355X-native HODLR Gaussian-process engine (Python binding).

The package mirrors the KernelMatrices.jl HODLR API; the work runs in
libkmhodlr.so (HIP, gfx950) through its C ABI.
"""
from . import _lib
from ._lib import load, LibraryMissing
from .api import (Context, ErrorException, FactorHODLR, FixedLevel, HBMBudgetExceeded, Kernel,
                  KernelHODLR, KernelMatrix, LogLevel, PosDefException, cauchy_kernel, default_context,
                  full, kd_order, maxlikdata, ldiv, logdet, matern12, matern32, negloglik, rational_kernel, stoch_fisher, stoch_gradient, stoch_hessian,
                  symmetricfactorize)


class HODLR:
    """Namespace mirroring `KernelMatrices.HODLR`."""
    KernelHODLR = KernelHODLR
    FixedLevel = FixedLevel
    LogLevel = LogLevel
    symmetricfactorize = staticmethod(symmetricfactorize)
    stoch_gradient = staticmethod(stoch_gradient)
    stoch_fisher = staticmethod(stoch_fisher)
    stoch_hessian = staticmethod(stoch_hessian)
    full = staticmethod(full)
    logdet = staticmethod(logdet)
    negloglik = staticmethod(negloglik)
