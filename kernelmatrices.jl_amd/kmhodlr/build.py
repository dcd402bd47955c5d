"""Build libkmhodlr.so (gfx950) from kernelmatrices.jl_amd/csrc with hipcc.

Every translation unit is compiled for gfx950 only, with packed FP32
instructions disabled, and linked into one shared library next to this file.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(HERE, "libkmhodlr.so")

SOURCES = ["ctx.hip", "la.hip", "qr.hip", "assemble.hip", "hodlr.hip", "gradient.hip", "hessian.hip", "capi.hip", "serve.hip", "ordering.hip", "rhodlr.hip", "testing.hip", "noalloc.hip", "graph.hip", "tune.hip", "leaf.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-I" + INCLUDE, "-I" + CSRC]


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(d, f) for d in (CSRC, INCLUDE)
                                        for f in os.listdir(d) if f.endswith(".h")]
    return any(os.path.getmtime(x) > t for x in deps)


def _compile(src, incremental=False):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, obj):
        return obj, ""
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, out))
    return obj, out


def build(verbose=False, incremental=False):
    """Compile every source (incremental=True: only those older than their
    source or a header) and link libkmhodlr.so."""
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(lambda s: _compile(s, incremental), SOURCES))
    for src, (obj, out) in zip(SOURCES, results):
        if verbose and out:
            print("[%s]\n%s" % (src, out), file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [r[0] for r in results]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
