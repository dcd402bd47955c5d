"""Writes tests/golden/fullsize_n65536.json: the oracle's results for the
n = 65536 Matern-3/2 workload (sigma 1, rho 0.1, nugget 1e-6, FixedLevel(8),
points uniform in [0,1]^2 from numpy default_rng(0), KD-ordered, y from the
same generator) with Nystrom rank 72 and with ACA tol 1e-10: logdet,
per-block ranks, negloglik, and the solve K^-1 y summarised by its norm, its
dot with y and every 512th entry.  Takes a few minutes on 8 cores.

    python tests/golden/make_fullsize.py
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


def inputs(n=65536, seed=0):
    rng = np.random.default_rng(seed)
    x = kd_sort(rng.random((n, 2)))
    y = rng.standard_normal(n)
    return x, y


def main():
    x, y = inputs()
    K = H.KernelMatrix(x, x, kk.Kernel(kk.MATERN32, (1.0, 0.1), nugget=1e-6))
    out = {"n": len(y), "seed": 0, "kernel": "matern32", "params": [1.0, 0.1], "nugget": 1e-6, "level": 8}
    for name, nys, ep, mr in (("nystrom72", True, 0.0, 72), ("aca1e-10", False, 1e-10, 0)):
        t = time.time()
        HK = H.KernelHODLR(K, ep, mr, H.FixedLevel(8), nystrom=nys)
        ranks = [list(map(int, r)) for r in HK.ranks]
        H.symmetricfactorize(HK)
        s = HK.solve(y)
        out[name] = {"ep": ep, "maxrank": mr, "ranks": ranks, "logdet": HK.logdet(),
                     "negloglik": H.negloglik(HK, y), "solve_norm": float(np.linalg.norm(s)),
                     "solve_dot_y": float(np.dot(s, y)), "solve_every_512": s[::512].tolist()}
        print(name, "%.1f s" % (time.time() - t), flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fullsize_n65536.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
