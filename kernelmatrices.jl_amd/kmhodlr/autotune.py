"""Write the GEMM kernel-class table (gemm_tuning_gfx950.txt next to this
file) by timing the kernel classes on the products of the HODLR path at the
served workload: 16 requests of n = 65536 points, leaves of 256, Nystrom
rank 72 (panel QR in 24-column blocks), one level at a time.

    python -m kmhodlr.autotune [--out PATH] [--reps 5]
"""
import argparse
import ctypes as C
import os
import sys


def shapes(requests=16, n=65536, leaf=256, r=72, qc=24):
    """(ta, tb, m, n, k, problems per launch) of the path's grouped products."""
    out = []
    L = 0
    while (leaf << L) < n:
        L += 1
    for lv in range(L):
        h = leaf << lv                          # child size of a level-(lv+1) block
        panels = requests * 2 * (n // (2 * h))  # U and V panels of the level's blocks
        out += [
            (0, 0, h, 2 * r, r, panels),        # Woodbury update  AB -= Q S2
            (1, 0, r, r, h, panels),            # Woodbury projection  Q^T AB, one half per block
            (1, 0, qc, qc, h, panels),          # QR projection, second block
            (1, 0, 2 * qc, qc, h, panels),      # QR projection, third block
            (0, 0, h, qc, qc, panels),          # QR update, second block
            (0, 0, h, qc, 2 * qc, panels),      # QR update, third block
            (1, 0, r, 1, h, panels),            # solve projection, one right-hand side
            (0, 0, h, 1, r, panels),            # solve update, one right-hand side
        ]
    leaves = requests * (n // leaf)
    out += [(0, 0, m, 2 * r, 64, leaves) for m in (192, 128, 64)]   # leaf TRSM updates
    out += [(0, 0, n, r, r, requests),          # Nystrom product K(x,P) K(P,P)^-1
            (0, 1, r, r, r, requests * 255),    # Bm = R1 R2^T
            (1, 0, r, r, r, requests * 255),    # Schur block Bm^T Bm
            (0, 0, 2 * r, 2 * r, 2 * r, requests * 255)]   # capacitance products
    return out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    import kmhodlr as km
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(here, "gemm_tuning_gfx950.txt"))
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    lib = km.load()
    ctx = km.Context(0, tuning=None)
    sh = shapes()
    arr = (C.c_int32 * (6 * len(sh)))(*[v for s in sh for v in s])
    rc = lib.kmh_gemm_autotune(ctx.ptr, len(sh), arr, args.reps, args.out.encode())
    if rc != 0:
        raise RuntimeError(lib.last_error())
    print(open(args.out).read(), flush=True)


if __name__ == "__main__":
    main()
