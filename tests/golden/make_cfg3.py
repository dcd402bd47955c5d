"""Config-3 batch workload and the oracle's results for its three smallest
problems, written to tests/golden/cfg3_batch64.json.

The workload: 64 independent GP problems, n_i = round(exp(U[ln 4096,
ln 65536])), points uniform in [0,1]^2, KD-ordered, Matern-1/2 exp(-r/rho)
with rho_i ~ U[0.05, 0.3] and a 1e-6 nugget, FixedLevel(floor(log2(n_i /
256))) (leaves of about 256), ACA at tol 1e-6 (no rank limit), 4 N(0,1)
right-hand sides each; one numpy default_rng(3) stream draws everything in
problem order.  The oracle (FP64) gives per-block ranks, logdet and the
solve of the first right-hand side (its norm and dot with b) for the three
smallest problems.

    python tests/golden/make_cfg3.py
"""
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import hodlr as H          # noqa: E402
from oracle import kernels as kk       # noqa: E402
from oracle.kdsort import kd_sort      # noqa: E402

COUNT, NRHS, TOL, NUGGET = 64, 4, 1e-6, 1e-6


def problems(count=COUNT, seed=3):
    """(x, b (n x NRHS, column-major), rho, level) per problem."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(4096), math.log(65536)))))
        rho = float(rng.uniform(0.05, 0.3))
        x = kd_sort(rng.random((n, 2)))
        b = np.asfortranarray(rng.standard_normal((n, NRHS)))
        level = int(math.floor(math.log2(n / 256)))
        out.append((x, b, rho, level))
    return out


def main():
    probs = problems()
    order = sorted(range(len(probs)), key=lambda i: probs[i][0].shape[0])[:3]
    out = {"count": COUNT, "seed": 3, "nrhs": NRHS, "tol": TOL, "nugget": NUGGET, "oracle": {}}
    for i in order:
        x, b, rho, level = probs[i]
        t = time.time()
        K = H.KernelMatrix(x, x, kk.Kernel(kk.MATERN12, (1.0, rho), nugget=NUGGET))
        HK = H.KernelHODLR(K, TOL, 0, H.FixedLevel(level), nystrom=False)
        ranks = [list(map(int, r)) for r in HK.ranks]
        H.symmetricfactorize(HK)
        s = HK.solve(b[:, 0])
        out["oracle"][str(i)] = {"n": int(x.shape[0]), "rho": rho, "level": level, "ranks": ranks,
                                 "logdet": HK.logdet(), "solve_norm": float(np.linalg.norm(s)),
                                 "solve_dot_b": float(np.dot(s, b[:, 0]))}
        print("problem", i, "n", x.shape[0], "%.1f s" % (time.time() - t), flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg3_batch64.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
