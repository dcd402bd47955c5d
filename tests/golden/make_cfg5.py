"""The largest serving request of the serving workload (n = 262144 =
2^18, FixedLevel(10): leaves of 256): the oracle's logdet, negloglik and
stochastic gradient (8 Rademacher vectors) for Matern-3/2 at sigma^2 = 1.69,
rho = 0.15, nugget 1e-6, Nystrom rank 72, points uniform in [0,1]^2 from
numpy default_rng(6), KD-ordered, y from the same generator, the vectors
from default_rng(7); written to tests/golden/cfg5_n262144.json.

    python tests/golden/make_cfg5.py
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

N, LEVEL, PARAMS, NUGGET, RANK, NVEC = 262144, 10, (1.69, 0.15), 1e-6, 72, 8


def inputs():
    rng = np.random.default_rng(6)
    x = kd_sort(rng.random((N, 2)))
    y = rng.standard_normal(N)
    saa = np.asfortranarray(np.random.default_rng(7).choice([-1.0, 1.0], size=(N, NVEC)))
    return x, y, saa


def main():
    x, y, saa = inputs()
    kern = kk.Kernel(kk.MATERN32, PARAMS, nugget=NUGGET)
    K = H.KernelMatrix(x, x, kern)
    t = time.time()
    HK = H.KernelHODLR(K, 0.0, RANK, H.FixedLevel(LEVEL), nystrom=True)
    H.symmetricfactorize(HK)
    out = {"n": N, "level": LEVEL, "params": list(PARAMS), "nugget": NUGGET, "rank": RANK, "nvec": NVEC,
           "logdet": HK.logdet(), "negloglik": H.negloglik(HK, y)}
    print("factorized %.1f s" % (time.time() - t), flush=True)
    t = time.time()
    out["gradient"] = H.stoch_gradient(K, HK, y, [kern.derivative(kk.D1), kern.derivative(kk.D2)],
                                       [saa[:, j] for j in range(NVEC)]).tolist()
    print("gradient %.1f s" % (time.time() - t), flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg5_n262144.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
