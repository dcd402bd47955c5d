"""KD-tree point ordering for 2-D inputs (oracle side; test/bench input prep).

The reference orders points with NearestNeighbors.KDTree(pts).data
(constructor.jl:171-181, example/basicexample.jl:57): points in the same KD
cell become contiguous, which is what makes the HODLR off-diagonal blocks
low-rank.  NearestNeighbors' exact ordering is not reproduced (its median
selection is an implementation detail of an absent dependency); this is a
plain recursive median split on the widest coordinate, so the ordering is
parity-unpinned and both sides of every comparison receive the same sorted
points explicitly.
"""
import numpy as np


def kd_order(pts, leafsize=8):
    """Permutation that KD-sorts pts (n x d)."""
    pts = np.asarray(pts, dtype=np.float64)
    n = pts.shape[0]
    perm = np.arange(n)
    stack = [(0, n)]
    while stack:
        lo, hi = stack.pop()
        if hi - lo <= leafsize:
            continue
        sub = pts[perm[lo:hi]]
        spread = sub.max(axis=0) - sub.min(axis=0)
        dim = int(np.argmax(spread))
        mid = (hi - lo) // 2
        order = np.argsort(sub[:, dim], kind="stable")
        perm[lo:hi] = perm[lo:hi][order]
        stack.append((lo, lo + mid))
        stack.append((lo + mid, hi))
    return perm


def kd_sort(pts, leafsize=8):
    pts = np.asarray(pts, dtype=np.float64)
    return np.ascontiguousarray(pts[kd_order(pts, leafsize)])
