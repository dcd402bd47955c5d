// Covariance functions on the device (FP64).  The formulas and their
// operation order follow the reference's kernel closures:
//   RATIONAL2  p1^2 / (1 + (r/p2)^2)^2              README.md:42-48
//   CAUCHY     p1^2 / (1 + (p2 r)^2)                example/basicexample.jl:62-68
//   MATERN12/32 t0 * mtn_cor(nu, r/rho) with the Stein parameterisation
//              (src/covariancefunctions.jl:141-212, 258-266); mtn_cor is
//              exactly 1 for x <= 1e-8 (:209-212) and its derivatives are 0
//              there (:213-233).  Closed forms for nu = 1/2, 3/2 (DLMF 10.39.2).
//   derivative codes D1, D2, D12, D22 follow mtn_kernfun_d1/_d2/_d1_d2/_d2_d2
//   (src/covariancefunctions.jl:268-330) in the fixed-nu two-parameter form
//   (:424-472).
// The distance is sqrt(dx*dx + dy*dy) with unfused multiplies and adds, as
// the norm of a 2-vector.  The nugget is added where both coordinates are
// equal (`x == y`, README.md:44-46) and only for the kernel itself.
#pragma once
#include <hip/hip_runtime.h>

#include "kmhodlr.h"

namespace kmh {

struct KDev {
    int family, deriv;
    double p0, p1, p2, p3;
    double nugget;
};

inline KDev to_kdev(const kmh_kernel& k) {
    return KDev{k.family, k.deriv, k.params[0], k.params[1], k.params[2], k.params[3], k.nugget};
}

constexpr double KMH_SQRT3 = 1.7320508075688772;

__device__ __forceinline__ double matern_cor(bool half, double x) {
    if (x <= 1.0e-8) return 1.0;
    if (half) return exp(-x);
    const double z = KMH_SQRT3 * x;
    return (1.0 + z) * exp(-z);
}

__device__ __forceinline__ double matern_cor_dx(bool half, double x) {
    if (x <= 1.0e-8) return 0.0;
    if (half) return -exp(-x);
    return (-3.0 * x) * exp(-(KMH_SQRT3 * x));
}

__device__ __forceinline__ double matern_cor_dxdx(bool half, double x) {
    if (x <= 1.0e-8) return 0.0;
    if (half) return exp(-x);
    const double z = KMH_SQRT3 * x;
    return (3.0 * (z - 1.0)) * exp(-z);
}

__device__ __forceinline__ double kernel_eval(const KDev& k, double x1, double y1, double x2, double y2) {
    const double dx = __dsub_rn(x1, x2), dy = __dsub_rn(y1, y2);
    const double r = sqrt(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)));
    double out;
    switch (k.family) {
        case KMH_RATIONAL2: {
            const double t = r / k.p1;
            const double s = __dadd_rn(1.0, __dmul_rn(t, t));
            out = __dmul_rn(k.p0, k.p0) / __dmul_rn(s, s);
            break;
        }
        case KMH_CAUCHY: {
            const double t = __dmul_rn(k.p1, r);
            out = __dmul_rn(k.p0, k.p0) / __dadd_rn(1.0, __dmul_rn(t, t));
            break;
        }
        default: {
            const bool half = (k.family == KMH_MATERN12);
            const double t0 = k.p0, rho = k.p1;
            const double x = r / rho;
            const bool pos = r > 0.0;
            const double rho2 = __dmul_rn(rho, rho);
            switch (k.deriv) {
                case KMH_D0: out = pos ? __dmul_rn(t0, matern_cor(half, x)) : t0; break;
                case KMH_D1: out = pos ? matern_cor(half, x) : 1.0; break;
                case KMH_D2:
                    out = pos ? __dmul_rn(__dmul_rn(-t0, matern_cor_dx(half, x)), r) / rho2 : 0.0;
                    break;
                case KMH_D12:
                    out = pos ? __dmul_rn(-matern_cor_dx(half, x), r) / rho2 : 0.0;
                    break;
                case KMH_D22: {
                    if (!pos) { out = 0.0; break; }
                    const double q = r / rho2;
                    const double a = __dmul_rn(matern_cor_dxdx(half, x), __dmul_rn(q, q));
                    const double b = __dmul_rn(__dmul_rn(matern_cor_dx(half, x), 2.0), r) / __dmul_rn(rho2, rho);
                    out = __dmul_rn(t0, __dadd_rn(a, b));
                    break;
                }
                default: out = 0.0; break;   // D11 and unknown codes
            }
        }
    }
    if (k.deriv == KMH_D0 && x1 == x2 && y1 == y2) out = __dadd_rn(out, k.nugget);
    return out;
}

}  // namespace kmh
