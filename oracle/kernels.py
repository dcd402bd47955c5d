"""Covariance functions of the HODLR path, FP64 numpy (oracle; test use only).

A kernel is described the same way the C ABI describes it
(`kmh_kernel` in include/kmhodlr.h): a family, a derivative code, up to four
parameters and a nugget that is added where the two points are equal
(point equality `x == y`, README.md:44-46).

Families
  RATIONAL2  p1^2 / (1 + (r/p2)^2)^2           README.md:42-48 (config-1 kernel)
  CAUCHY     p1^2 / (1 + (p2 r)^2)             example/basicexample.jl:62-68
  MATERN12   t0 * mtn_cor(1/2, r/rho)          src/covariancefunctions.jl:258-266
  MATERN32   t0 * mtn_cor(3/2, r/rho)          src/covariancefunctions.jl:258-266

mtn_cor is the Stein-parameterised Matern correlation
(src/covariancefunctions.jl:141-212):
  mtn_cor(nu, x) = (sqrt(2 nu) x)^nu K_nu(sqrt(2 nu) x) / (Gamma(nu) 2^(nu-1)),
which returns exactly 1.0 for x <= 1e-8 (:209-212).  With the half-integer
Bessel forms (DLMF 10.39.2) this is e^{-x} for nu = 1/2 and (1 + z) e^{-z},
z = sqrt(3) x, for nu = 3/2.

Derivative codes follow the reference's fixed-nu two-parameter pattern
(mt1_kernfun* at src/covariancefunctions.jl:424-472, general-nu versions at
:268-330), parameters (t0, rho):
  D1   d/dt0      = mtn_cor(x)                   (1 at r == 0)      :268-276
  D2   d/drho     = -t0 * cor'(x) * r / rho^2    (0 at r == 0)      :278-286
  D11  d2/dt0^2   = 0                                               :298-300
  D12  d2/dt0drho = -cor'(x) * r / rho^2                            :302-310
  D22  d2/drho^2  = t0 * (cor''(x) (r/rho^2)^2 + cor'(x) 2 r / rho^3) :322-330
with cor'(x) = cor''(x) = 0 for x <= 1e-8 (mtn_cor_dx :213-216,
mtn_cor_dx_dx :227-233).  Derivative kernels carry no nugget.
"""
import numpy as np

RATIONAL2 = 1
CAUCHY = 2
MATERN12 = 3
MATERN32 = 4

D0, D1, D2, D11, D12, D22 = 0, 1, 2, 11, 12, 22

SQRT3 = 1.7320508075688772


class Kernel:
    """Kernel descriptor: family, derivative code, params, nugget."""

    def __init__(self, family, params, nugget=0.0, deriv=D0):
        self.family = int(family)
        self.params = tuple(float(p) for p in params)
        self.nugget = float(nugget)
        self.deriv = int(deriv)

    def derivative(self, code):
        """The derivative kernel (no nugget), as the reference's dfuns."""
        return Kernel(self.family, self.params, 0.0, code)

    def __call__(self, X1, X2):
        return kernel_matrix(self, X1, X2)

    def __repr__(self):
        return "Kernel(family=%d, params=%r, nugget=%g, deriv=%d)" % (
            self.family, self.params, self.nugget, self.deriv)


def distances(X1, X2):
    """r_ij = sqrt(dx^2 + dy^2), the StaticArrays norm(x - y) for 2-D points."""
    X1 = np.asarray(X1, dtype=np.float64)
    X2 = np.asarray(X2, dtype=np.float64)
    dx = X1[:, 0][:, None] - X2[:, 0][None, :]
    dy = X1[:, 1][:, None] - X2[:, 1][None, :]
    return np.sqrt(dx * dx + dy * dy)


def equal_points(X1, X2):
    X1 = np.asarray(X1, dtype=np.float64)
    X2 = np.asarray(X2, dtype=np.float64)
    return ((X1[:, 0][:, None] == X2[:, 0][None, :]) &
            (X1[:, 1][:, None] == X2[:, 1][None, :]))


def matern_cor(nu2, x):
    """mtn_cor for nu = nu2/2 in {1/2, 3/2}; exactly 1 for x <= 1e-8."""
    x = np.asarray(x, dtype=np.float64)
    if nu2 == 1:
        c = np.exp(-x)
    else:
        z = SQRT3 * x
        c = (1.0 + z) * np.exp(-z)
    return np.where(x <= 1.0e-8, 1.0, c)


def matern_cor_dx(nu2, x):
    x = np.asarray(x, dtype=np.float64)
    if nu2 == 1:
        c = -np.exp(-x)
    else:
        c = -3.0 * x * np.exp(-SQRT3 * x)
    return np.where(x <= 1.0e-8, 0.0, c)


def matern_cor_dxdx(nu2, x):
    x = np.asarray(x, dtype=np.float64)
    if nu2 == 1:
        c = np.exp(-x)
    else:
        z = SQRT3 * x
        c = 3.0 * (z - 1.0) * np.exp(-z)
    return np.where(x <= 1.0e-8, 0.0, c)


def kernel_values(k, r, eq):
    """Evaluate kernel k on a distance array r with equality mask eq."""
    p = k.params
    fam, d = k.family, k.deriv
    if fam == RATIONAL2:
        if d != D0:
            raise ValueError("derivatives of RATIONAL2 are not provided")
        t = r / p[1]
        out = (p[0] * p[0]) / np.square(1.0 + t * t)
    elif fam == CAUCHY:
        if d != D0:
            raise ValueError("derivatives of CAUCHY are not provided")
        t = p[1] * r
        out = (p[0] * p[0]) / (1.0 + t * t)
    elif fam in (MATERN12, MATERN32):
        nu2 = 1 if fam == MATERN12 else 3
        t0, rho = p[0], p[1]
        x = r / rho
        pos = r > 0.0
        if d == D0:
            out = np.where(pos, t0 * matern_cor(nu2, x), t0)
        elif d == D1:
            out = np.where(pos, matern_cor(nu2, x), 1.0)
        elif d == D2:
            out = np.where(pos, -t0 * matern_cor_dx(nu2, x) * r / (rho * rho), 0.0)
        elif d == D11:
            out = np.zeros_like(r)
        elif d == D12:
            out = np.where(pos, -matern_cor_dx(nu2, x) * r / (rho * rho), 0.0)
        elif d == D22:
            out = np.where(pos, t0 * (matern_cor_dxdx(nu2, x) * np.square(r / (rho * rho))
                                      + matern_cor_dx(nu2, x) * 2.0 * r / rho ** 3), 0.0)
        else:
            raise ValueError("unknown derivative code %d" % d)
    else:
        raise ValueError("unknown kernel family %d" % fam)
    out = np.array(out, dtype=np.float64)
    if d == D0 and k.nugget != 0.0:
        out = out + np.where(eq, k.nugget, 0.0)
    return out


def kernel_matrix(k, X1, X2):
    """Dense block [k(x, y) for x in X1, y in X2] (src/baseoverloads.jl:10-12)."""
    return kernel_values(k, distances(X1, X2), equal_points(X1, X2))
