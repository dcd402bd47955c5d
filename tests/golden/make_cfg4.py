"""Writes tests/golden/cfg4_n131072.json: the oracle's results for the
stochastic-gradient / Fisher / Hessian workload at n = 131072 (Matern-3/2,
sigma 1 (t0 = 1), rho 0.08, nugget 1e-6, Nystrom rank 72, FixedLevel(9),
points uniform in [0,1]^2 from numpy default_rng(4), KD-ordered, y from the
same generator, 64 Rademacher vectors from default_rng(5)): logdet,
negloglik, stoch_gradient (d/dt0, d/drho), stoch_fisher and stoch_hessian
with the reference's second-derivative block terms (d2/dt0^2 is identically
zero: the reference's ZeroFunction); the exact second-derivative terms only
for n <= 32768 (the oracle forms them as dense blocks).  Takes about half an
hour on the CPU at full size (the oracle loops over the vectors in Python).

    python tests/golden/make_cfg4.py [n level nvec]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import hodlr as H          # noqa: E402
from oracle import kernels as kk       # noqa: E402
from oracle.kdsort import kd_sort      # noqa: E402

PARAMS = (1.0, 0.08)
NUGGET = 1e-6
RANK = 72


def inputs(n=131072, nvec=64, seed=4, saa_seed=5):
    rng = np.random.default_rng(seed)
    x = kd_sort(rng.random((n, 2)))
    y = rng.standard_normal(n)
    saa = np.asfortranarray(np.random.default_rng(saa_seed).choice([-1.0, 1.0], size=(n, nvec)))
    return x, y, saa


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 131072
    level = int(sys.argv[2]) if len(sys.argv) > 2 else 9
    nvec = int(sys.argv[3]) if len(sys.argv) > 3 else 64
    x, y, saa = inputs(n, nvec)
    vecs = [saa[:, j] for j in range(nvec)]
    kern = kk.Kernel(kk.MATERN32, PARAMS, nugget=NUGGET)
    K = H.KernelMatrix(x, x, kern)
    t = time.time()
    HK = H.KernelHODLR(K, 0.0, RANK, H.FixedLevel(level), nystrom=True)
    H.symmetricfactorize(HK)
    out = {"n": n, "level": level, "nvec": nvec, "seed": 4, "saa_seed": 5, "params": list(PARAMS),
           "nugget": NUGGET, "rank": RANK, "logdet": HK.logdet(), "negloglik": H.negloglik(HK, y)}
    print("factorized %.1f s" % (time.time() - t), flush=True)
    d1 = [kern.derivative(kk.D1), kern.derivative(kk.D2)]
    d2 = [[None, kern.derivative(kk.D12)], [kern.derivative(kk.D22)]]
    t = time.time()
    out["gradient"] = H.stoch_gradient(K, HK, y, d1, vecs).tolist()
    print("gradient %.1f s" % (time.time() - t), flush=True)
    t = time.time()
    DKs = [H.DerivativeHODLR(K, dk, HK) for dk in d1]
    out["fisher"] = H.stoch_fisher(K, HK, DKs, vecs).tolist()
    print("fisher %.1f s" % (time.time() - t), flush=True)
    for key, exact in (("hessian_reference_terms", False), ("hessian_exact_second", True)):
        if exact and n > 32768:
            continue
        t = time.time()
        out[key] = H.stoch_hessian(K, HK, y, d1, d2, vecs, exact=exact).tolist()
        print(key, "%.1f s" % (time.time() - t), flush=True)
    name = "cfg4_n%d.json" % n
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), name), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", name, flush=True)


if __name__ == "__main__":
    main()
