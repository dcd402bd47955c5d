"""Benchmark: GP log-likelihood evaluations per second on one MI355X,
served (continuous batching) vs plain (one request at a time), n = 65536.

Workload: points uniform in [0,1]^2 (numpy PCG64 seed 0), KD-ordered by
the library, y ~ N(0,1); Matern nu=3/2 with a 1e-6 nugget; FixedLevel(8) (256 leaves of
256 points); Nystrom rank 72 (the served objective of the reference,
nll_objective, always uses Nystrom).  A request is one log-likelihood
evaluation, negloglik (maxlikfunctions.jl:2-9): HODLR build +
symmetricfactorize! + logdet + y' K^-1 y, FP64 on the GPU, at its own kernel
parameters (sigma ~ U[0.5, 2], rho ~ U[0.05, 0.2]), inputs resident in HBM.
Both legs take y' K^-1 y as |W^-1 y|^2 with the symmetric factor K = W W^T
(one W^-1 application instead of the full solve); the gradient legs apply
W^-T on top where the gradient needs K^-1 y.

Legs: `plain` runs the requests one after another through the per-request
C ABI; `served` keeps `--inflight` requests in the server, which advances
all of them stage by stage in shared grouped launches, spread over the
server's lanes (HIP streams stepped on parallel host threads).  Both legs
see the same requests; their logdets are compared.  `roofline` runs one
served batch on a single lane with every launch of the two timed kernels
(k_tile_qr<24>, k_gemm_nn_wide) timed by HIP event markers around them; the
headline legs time every PROF_STRIDE-th launch (a timed launch idles the
queue for a few microseconds).  The JSON line carries the roofline of the
one with more kernel time (the dominant kernel) with the other beside it,
and a CPU baseline (the numpy oracle, one BLAS thread).
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "kernelmatrices.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBPS = 8000.0
PROF_STRIDE = 4   # time every 4th k_gemm_nn_wide launch of a timed region
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "tile_qr_nn_wide_served_n65536_traffic.md")


KERNEL_SOURCES = ("la.hip", "qr.hip", "common.h", "la.h", "wavesum.h")


def kernel_source_hash():
    """sha256 of the sources of the timed kernels (the GEMM and tile-QR files)."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "kernelmatrices.jl_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (a line 'traffic_bytes_per_launch <kernel> <bytes>'), when the
    summary's 'source_sha256' line matches the current kernel sources;
    (None, reason) otherwise."""
    stamp, val = None, None
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2 and parts[0] == "source_sha256":
                    stamp = parts[1]
                if parts and parts[0] == "traffic_bytes_per_launch":
                    # the kernel name may hold spaces ("k_tile_qr<24, double, *>")
                    name, _, num = line.strip()[len("traffic_bytes_per_launch"):].strip().rpartition(" ")
                    if name.strip() == kernel:
                        val = int(float(num))
    except OSError:
        return None, "no committed rocprofv3 --pmc summary"
    if val is None:
        return None, "kernel not in the committed summary"
    if stamp != kernel_source_hash():
        return None, "stale: the kernel sources changed since the committed rocprofv3 --pmc run"
    return val, "rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel (%s)" % os.path.basename(TRAFFIC_FILE)


def make_inputs(n, seed=0):
    """Uniform points in [0,1]^2 and y ~ N(0,1), put in KD order by the
    library (maxlikdata, HODLR/src/constructor.jl:171-181)."""
    import numpy as np
    import kmhodlr as km
    rng = np.random.default_rng(seed)
    x = rng.random((n, 2))
    y = rng.standard_normal(n)
    return km.maxlikdata(x, y)


def request_params(count, seed=1):
    import numpy as np
    rng = np.random.default_rng(seed)
    sig = rng.uniform(0.5, 2.0, count)
    rho = rng.uniform(0.05, 0.2, count)
    return [(float(s * s), float(r)) for s, r in zip(sig, rho)]


def factor_flops(n, m, ranks, nblocks):
    """Algorithmic FP64 flops of symmetricfactorize! with structured M
    (leaf potrf + leaf trsm + panel QR + capacitance + Woodbury cascade)."""
    L = len(ranks)
    f = n * m * m / 3.0 + n * m * sum(ranks)
    for l, r in enumerate(ranks):
        f += 4.0 * n * r * r - nblocks[l] * 2 * 4.0 * r ** 3 / 3.0
        f += nblocks[l] * (2.0 * r ** 3 / 3.0 + 8.0 * r ** 3 / 3.0)
    for lv in range(L):
        for lev in range(lv + 1, L):
            f += 4.0 * n * ranks[lv] * ranks[lev] + 12.0 * ranks[lv] ** 2 * ranks[lev] * nblocks[lv] + n * ranks[lev]
    return f


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_baseline(n, lvl, params, full=False, timed=3):
    """The oracle (numpy/scipy restatement of the reference's op sequence)
    timed on the host at the bench's own workload: one negloglik eval
    (build + symmetricfactorize! + logdet + y'K^-1 y) at n = 65536 takes
    about 90 s whatever the thread count (the restatement is bound by its
    per-block Python loops), so the default run times one eval on all the
    host threads the job may use; full=True times, for each of 1 BLAS thread
    (the reference's BLAS.set_num_threads(1), test/runtests.jl:9) and all
    threads, the median of `timed` evals after one warm-up."""
    import statistics
    from threadpoolctl import threadpool_limits
    from oracle import hodlr as H
    from oracle import kernels as kk
    x, y = make_inputs(n)

    def one():
        t0 = time.perf_counter()
        K = H.KernelMatrix(x, x, kk.Kernel(kk.MATERN32, params, nugget=1e-6))
        HK = H.KernelHODLR(K, 0.0, 72, H.FixedLevel(lvl), nystrom=True)
        H.symmetricfactorize(HK)
        nll = H.negloglik(HK, y)
        return time.perf_counter() - t0, nll

    nthreads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    res = {}
    for th in ((1, nthreads) if full else (nthreads,)):
        with threadpool_limits(th):
            if full:
                progress("cpu baseline warm-up, %d threads" % th)
                one()
            ts = []
            for q in range(timed if full else 1):
                ts.append(one()[0])
                progress("cpu baseline eval on %d threads: %.1f s" % (th, ts[-1]))
        res[th] = statistics.median(ts)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return res, nthreads, model


def poisson_leg(args, lib, L, ctx, kernel_of, check):
    """Open-loop serving of likelihood + gradient requests arriving as a
    Poisson stream: n log-uniform in [4096, 262144] (FixedLevel so leaves
    hold about 256 points), sigma ~ U[0.5, 2], rho ~ U[0.02, 0.3],
    Matern-3/2, Nystrom rank 72, y ~ N(0,1), 16 SAA vectors, a 200 GB HBM
    budget for the in-flight requests.  Inputs are generated and put in HBM
    before the stream starts.  Reports served evals/s over the stream,
    p50/p99 latency and queue time, beside the plain per-request FP64 path
    over the same requests."""
    import numpy as np
    import torch
    import kmhodlr as km
    rng = np.random.default_rng(9)
    count, nv, rate = args.poisson_requests, 16, args.poisson_rate
    dev = torch.device("cuda:0")
    reqs = []
    for i in range(count):
        nn = int(round(math.exp(rng.uniform(math.log(4096), math.log(262144)))))
        x = rng.random((nn, 2))
        x = np.ascontiguousarray(x[km.kd_order(x)])
        y = rng.standard_normal(nn)
        saa = rng.choice([-1.0, 1.0], size=(nv, nn))       # row j = vector j: column-major n x nv
        sig, rho = rng.uniform(0.5, 2.0), rng.uniform(0.02, 0.3)
        lvl = max(1, int(math.floor(math.log2(nn / 256))))
        reqs.append({"n": nn, "x": torch.from_numpy(x).to(dev), "y": torch.from_numpy(y).to(dev),
                     "saa": torch.from_numpy(saa).to(dev), "p": (float(sig * sig), float(rho)), "lvl": lvl})
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    torch.cuda.synchronize()

    def make(q):
        r = L.kmh_serve_req()
        r.pts = q["x"].data_ptr()
        r.n = q["n"]
        r.d = 2
        r.on_device = 1
        r.kernel = kernel_of(q["p"])
        r.opts.method = L.KMH_NYSTROM
        r.opts.level_kind = L.KMH_FIXED_LEVEL
        r.opts.level = q["lvl"]
        r.opts.maxrank = 72
        r.opts.precision = L.KMH_FP64
        r.y = q["y"].data_ptr()
        r.ndk = 2
        r.saa = q["saa"].data_ptr()
        r.nv = nv
        r.dks[0] = kernel_of(q["p"], L.KMH_D1)
        r.dks[1] = kernel_of(q["p"], L.KMH_D2)
        return r

    srv = C.c_void_p()
    check(lib.kmh_server_create(ctx.ptr, args.inflight, 200 * (1 << 30), C.byref(srv)))
    keep, live, res = [], {}, {}
    t0 = time.perf_counter()
    nxt = 0
    while nxt < count or live:
        now = time.perf_counter() - t0
        while nxt < count and arrivals[nxt] <= now:
            r = make(reqs[nxt])
            keep.append(r)
            rid = C.c_int64()
            check(lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)))
            live[rid.value] = nxt
            nxt += 1
        if not live:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
            continue
        done = C.c_int32()
        check(lib.kmh_server_step(srv, C.byref(done)))
        if done.value:
            for rid in list(live):
                o = L.kmh_serve_result()
                rc = lib.kmh_server_result(srv, rid, C.byref(o))
                if rc == L.KMH_E_STATE:
                    continue
                check(rc)
                if o.status != 0:
                    raise RuntimeError("request failed: %s" % lib.last_error())
                res[live.pop(rid)] = (o.nll, o.latency_us * 1e-3, o.queue_us * 1e-3)
    el = time.perf_counter() - t0
    lib.kmh_server_destroy(srv)
    lat = [v[1] for v in res.values()]
    qt = [v[2] for v in res.values()]
    # the plain per-request FP64 path over the same requests
    t1 = time.perf_counter()
    plain_err = 0.0
    for i, q in enumerate(reqs):
        h = C.c_void_p()
        kern = kernel_of(q["p"])
        o = L.kmh_build_opts()
        o.method, o.level_kind, o.level, o.maxrank, o.precision = L.KMH_NYSTROM, L.KMH_FIXED_LEVEL, q["lvl"], 72, 0
        check(lib.kmh_hodlr_build_dev(ctx.ptr, q["x"].data_ptr(), q["n"], 2, C.byref(kern), C.byref(o), C.byref(h)))
        check(lib.kmh_hodlr_factorize(h))
        ld = C.c_double()
        check(lib.kmh_hodlr_logdet(h, C.byref(ld)))
        z = torch.empty_like(q["y"])
        check(lib.kmh_hodlr_factor_apply_dev(h, L.KMH_WINV, q["y"].data_ptr(), z.data_ptr(), 1))
        check(lib.kmh_ctx_synchronize(ctx.ptr))
        nll = 0.5 * ld.value + 0.5 * float(torch.dot(z, z))
        dks = (L.kmh_kernel * 2)(kernel_of(q["p"], L.KMH_D1), kernel_of(q["p"], L.KMH_D2))
        g = np.zeros(2)
        check(lib.kmh_stoch_gradient_dev(h, dks, 2, q["y"].data_ptr(), q["saa"].data_ptr(), nv, g.ctypes.data))
        check(lib.kmh_hodlr_destroy(h))
        plain_err = max(plain_err, abs(nll - res[i][0]) / abs(nll))
    pel = time.perf_counter() - t1
    ns = [q["n"] for q in reqs]
    return {"requests": count, "arrival_rate_per_s": rate, "evals_per_s": count / el, "p50_ms": pct(lat, 0.5),
            "p99_ms": pct(lat, 0.99), "queue_p50_ms": pct(qt, 0.5), "queue_p99_ms": pct(qt, 0.99),
            "n_min": min(ns), "n_max": max(ns), "n_mean": float(np.mean(ns)), "saa_vectors": nv,
            "hbm_budget_bytes": 200 * (1 << 30), "plain_evals_per_s": count / pel,
            "max_rel_nll_err_served_vs_plain": plain_err,
            "note": "open loop: a request is submitted at its arrival time; evals/s over the whole stream "
                    "(arrival span included)"}


def pct(v, q):
    s = sorted(v)
    return s[min(len(s) - 1, max(0, int(math.ceil(q * len(s))) - 1))]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48, help="timed requests per leg")
    ap.add_argument("--warmup", type=int, default=8, help="untimed requests per leg")
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--inflight", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline on 1 thread and all threads, median of 3 after a warm-up (about 15 minutes)")
    ap.add_argument("--legs", default="plain,served,plain_grad,served_grad,served_mixed,roofline,aca,followup,poisson")
    ap.add_argument("--lanes", type=int, default=0, help="server lanes (0: the library's default)")
    ap.add_argument("--grad-vecs", type=int, default=32, help="SAA vectors of the gradient legs")
    ap.add_argument("--poisson-requests", type=int, default=48)
    ap.add_argument("--poisson-rate", type=float, default=60.0, help="arrivals per second of the Poisson leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(independent GP requests shard across GPUs without collectives)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import kmhodlr as km
    from kmhodlr import _lib as L

    lib = km.load()
    if lib.device_count() < 1:
        raise RuntimeError("no HIP device")
    n = args.n
    x, y = make_inputs(n)
    dev = torch.device("cuda:0")
    pts_d = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    y_d = torch.from_numpy(y).to(dev)
    sol_d = torch.empty_like(y_d)
    torch.cuda.synchronize()
    ctx = km.Context(0)
    total = args.warmup + args.steps
    params = request_params(total)

    def check(rc):
        if rc != 0:
            raise RuntimeError(lib.last_error())

    def kernel_of(p, deriv=0):
        k = L.kmh_kernel()
        k.family = L.KMH_MATERN32
        k.deriv = deriv
        k.params[0], k.params[1] = p
        k.nugget = 1e-6 if deriv == 0 else 0.0
        return k

    # gradient legs: nll_objective (maxlikfunctions.jl:29-50) = negloglik +
    # stoch_gradient with Rademacher SAA vectors
    saa = np.asfortranarray(np.random.default_rng(2).choice([-1.0, 1.0], size=(n, args.grad_vecs)))
    # the SAA vectors are fixed across evaluations: resident in HBM like y
    saa_d = torch.from_numpy(np.ascontiguousarray(saa.T)).to(dev)   # column-major n x nv
    y_host = np.ascontiguousarray(y)

    opts = L.kmh_build_opts()
    opts.method = L.KMH_NYSTROM
    opts.level_kind = L.KMH_FIXED_LEVEL
    opts.level = args.level
    opts.maxrank = 72
    opts.tol = 0.0
    opts.precision = L.KMH_FP64
    info = {}
    out = {}

    # ---- plain: one request at a time ----------------------------------
    def quad_form(h):
        """y' K^-1 y = |W^-1 y|^2 with the symmetric factor K = W W^T, the
        same arithmetic as the server's"""
        check(lib.kmh_hodlr_factor_apply_dev(h, L.KMH_WINV, y_d.data_ptr(), sol_d.data_ptr(), 1))
        check(lib.kmh_ctx_synchronize(ctx.ptr))
        return float(torch.dot(sol_d, sol_d))

    def plain_eval(p):
        h = C.c_void_p()
        kern = kernel_of(p)
        check(lib.kmh_hodlr_build_dev(ctx.ptr, pts_d.data_ptr(), n, 2, C.byref(kern), C.byref(opts), C.byref(h)))
        check(lib.kmh_hodlr_factorize(h))
        ld = C.c_double()
        check(lib.kmh_hodlr_logdet(h, C.byref(ld)))
        quad = quad_form(h)
        if not info:
            ii = L.kmh_info()
            check(lib.kmh_hodlr_info(h, C.byref(ii)))
            info["levels"] = ii.levels
            info["device_bytes"] = ii.device_bytes
        check(lib.kmh_hodlr_destroy(h))
        return 0.5 * ld.value + 0.5 * quad, ld.value

    def plain_grad_eval(p):
        h = C.c_void_p()
        kern = kernel_of(p)
        check(lib.kmh_hodlr_build_dev(ctx.ptr, pts_d.data_ptr(), n, 2, C.byref(kern), C.byref(opts), C.byref(h)))
        check(lib.kmh_hodlr_factorize(h))
        ld = C.c_double()
        check(lib.kmh_hodlr_logdet(h, C.byref(ld)))
        quad = quad_form(h)
        dks = (L.kmh_kernel * 2)(kernel_of(p, L.KMH_D1), kernel_of(p, L.KMH_D2))
        g = np.zeros(2)
        check(lib.kmh_stoch_gradient_dev(h, dks, 2, y_d.data_ptr(), saa_d.data_ptr(), saa.shape[1], g.ctypes.data))
        check(lib.kmh_hodlr_destroy(h))
        return 0.5 * ld.value + 0.5 * quad, g

    legs = args.legs.split(",")
    PROF_MASK = (1 << L.KMH_PROF_NN_WIDE) | (1 << L.KMH_PROF_TILE_QR)
    check(lib.kmh_ctx_profile_config(ctx.ptr, PROF_MASK, PROF_STRIDE))

    def prof_begin():
        check(lib.kmh_ctx_profile_begin(ctx.ptr))

    def prof_end():
        """k_gemm_nn_wide and k_tile_qr<24> totals of the timed launches
        since prof_begin"""
        fl, ms, cnt = C.c_double(), C.c_double(), C.c_int64()
        check(lib.kmh_ctx_profile_end(ctx.ptr, C.byref(fl), C.byref(ms), C.byref(cnt)))
        out_ = {}
        for tag, key in ((L.KMH_PROF_NN_WIDE, "nn_wide"), (L.KMH_PROF_TILE_QR, "tile_qr")):
            b, t, nl = C.c_double(), C.c_double(), C.c_int64()
            check(lib.kmh_ctx_profile_query(ctx.ptr, tag, C.byref(b), C.byref(t), C.byref(nl)))
            k = max(1, nl.value)
            out_.update({key + "_bytes_per_launch": b.value / k, key + "_us_per_launch": 1e3 * t.value / k,
                         key + "_ms": t.value, key + "_timed_launches": nl.value,
                         key + "_gbps": b.value / (t.value * 1e-3) / 1e9 if t.value else 0.0})
        out_["nn_wide_tflops"] = fl.value / (ms.value * 1e-3) / 1e12 if ms.value else 0.0
        return out_

    plain_ld, plain_nll = {}, {}
    progress("legs: " + ",".join(legs))
    if "plain" in legs:
        for i in range(args.warmup):
            plain_eval(params[i])
        torch.cuda.synchronize()
        prof_begin()
        lat = []
        t0 = time.perf_counter()
        for i in range(args.warmup, total):
            s = time.perf_counter()
            nll, ld = plain_eval(params[i])
            lat.append(1e3 * (time.perf_counter() - s))
            plain_ld[i] = ld
            plain_nll[i] = nll
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out["plain"] = {"evals_per_s": args.steps / el, "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99)}
        progress("plain: %.1f evals/s" % out["plain"]["evals_per_s"])
        out["plain"].update(prof_end())

    plain_grad = {}
    if "plain_grad" in legs:
        for i in range(args.warmup):
            plain_grad_eval(params[i])
        torch.cuda.synchronize()
        lat = []
        t0 = time.perf_counter()
        for i in range(args.warmup, total):
            s_ = time.perf_counter()
            plain_grad[i] = plain_grad_eval(params[i])
            lat.append(1e3 * (time.perf_counter() - s_))
        el = time.perf_counter() - t0
        out["plain_grad"] = {"evals_per_s": args.steps / el, "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
                             "saa_vectors": args.grad_vecs}
        progress("plain_grad: %.1f evals/s" % out["plain_grad"]["evals_per_s"])

    # ---- served: continuous batching -----------------------------------
    if "served" in legs or "served_grad" in legs or "roofline" in legs:
        srv = C.c_void_p()
        check(lib.kmh_server_create(ctx.ptr, args.inflight, 0, C.byref(srv)))
        if args.lanes:
            check(lib.kmh_server_set_lanes(srv, args.lanes))
        # the server's device pools sized for the requests in flight (server
        # set-up, before any leg: serving then makes no device allocation)
        check(lib.kmh_server_reserve(srv, n, C.byref(opts), args.grad_vecs, args.inflight))
        reqs = []

        grad_mode = [False]
        prec_mode = [L.KMH_FP64]

        def submit(i):
            r = L.kmh_serve_req()
            r.pts = pts_d.data_ptr()
            r.n = n
            r.d = 2
            r.ndk = 0
            r.on_device = 1
            r.kernel = kernel_of(params[i])
            r.opts = opts
            r.opts.precision = prec_mode[0]
            r.y = y_d.data_ptr()
            if grad_mode[0]:
                r.ndk = 2
                r.saa = saa_d.data_ptr()
                r.nv = saa.shape[1]
                r.dks[0] = kernel_of(params[i], L.KMH_D1)
                r.dks[1] = kernel_of(params[i], L.KMH_D2)
            reqs.append(r)
            rid = C.c_int64()
            check(lib.kmh_server_submit(srv, C.byref(r), C.byref(rid)))
            return rid.value

        def run(idx, profile):
            """closed loop: keep `inflight` requests outstanding"""
            pending = list(idx)
            live = {}
            res = {}
            if profile:
                prof_begin()
            t0 = time.perf_counter()
            while pending or live:
                while pending and len(live) < args.inflight:
                    i = pending.pop(0)
                    live[submit(i)] = i
                done = C.c_int32()
                check(lib.kmh_server_step(srv, C.byref(done)))
                if done.value:
                    for rid in list(live):
                        o = L.kmh_serve_result()
                        rc = lib.kmh_server_result(srv, rid, C.byref(o))
                        if rc == L.KMH_E_STATE:
                            continue
                        check(rc)
                        if o.status != 0:
                            raise RuntimeError("request failed: %s" % lib.last_error())
                        res[live.pop(rid)] = (o.nll, o.logdet, o.latency_us * 1e-3, [o.grad[0], o.grad[1]])
            el = time.perf_counter() - t0
            return res, el, (prof_end() if profile else None)

        if "served" in legs:
            run(range(args.warmup), False)
            torch.cuda.synchronize()
            res, el, prof = run(range(args.warmup, total), True)
            lat = [v[2] for v in res.values()]
            out["served"] = {"evals_per_s": args.steps / el, "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
                             "inflight": args.inflight}
            progress("served: %.1f evals/s" % out["served"]["evals_per_s"])
            out["served"].update(prof)
            if plain_ld:
                errs = [abs(res[i][1] - plain_ld[i]) / abs(plain_ld[i]) for i in plain_ld]
                out["served"]["max_rel_logdet_err_vs_plain"] = max(errs)
        if "served_mixed" in legs:
            # the same requests in the mixed form KMH_MIXED_Q (FP32 panel QR and
            # Q factors, FP64 cascade, leaves, capacitance and logdet; y'K^-1 y
            # from the solve refined against the FP64 operator)
            prec_mode[0] = L.KMH_MIXED_Q
            run(range(args.warmup), False)
            torch.cuda.synchronize()
            res, el, _ = run(range(args.warmup, total), False)
            prec_mode[0] = L.KMH_FP64
            lat = [v[2] for v in res.values()]
            out["served_mixed"] = {"evals_per_s": args.steps / el, "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
                                   "inflight": args.inflight, "precision": "KMH_MIXED_Q"}
            if plain_ld:
                errs = [abs(res[i][1] - plain_ld[i]) / abs(plain_ld[i]) for i in plain_ld]
                nerr = [abs(res[i][0] - plain_nll[i]) / abs(plain_nll[i]) for i in plain_nll]
                out["served_mixed"]["max_rel_logdet_err_vs_plain_fp64"] = max(errs)
                out["served_mixed"]["max_rel_nll_err_vs_plain_fp64"] = max(nerr)
        if "served_grad" in legs:
            grad_mode[0] = True
            run(range(args.warmup), False)
            torch.cuda.synchronize()
            res, el, _ = run(range(args.warmup, total), False)
            lat = [v[2] for v in res.values()]
            out["served_grad"] = {"evals_per_s": args.steps / el, "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
                                  "saa_vectors": args.grad_vecs, "inflight": args.inflight}
            if plain_grad:
                gerr = max(abs(res[i][3][j] - plain_grad[i][1][j]) / abs(plain_grad[i][1][j])
                           for i in plain_grad for j in range(2))
                out["served_grad"]["max_rel_grad_err_vs_plain"] = gerr
        if "roofline" in legs:
            # the timed kernels without another stream's kernels beside them:
            # one served batch on one lane, every launch timed
            grad_mode[0] = False
            check(lib.kmh_server_set_lanes(srv, 1))
            check(lib.kmh_server_reserve(srv, n, C.byref(opts), 0, args.inflight))
            check(lib.kmh_ctx_profile_config(ctx.ptr, PROF_MASK, 1))
            k = min(args.inflight, total)
            run(range(k), False)
            torch.cuda.synchronize()
            res, el, prof = run(range(k), True)
            check(lib.kmh_ctx_profile_config(ctx.ptr, PROF_MASK, PROF_STRIDE))
            lat = [v[2] for v in res.values()]
            out["roofline_pass"] = {"requests": k, "lanes": 1, "evals_per_s": k / el, "p50_ms": pct(lat, 0.5),
                                    "p99_ms": pct(lat, 0.99)}
            out["roofline_pass"].update(prof)
        lib.kmh_server_destroy(srv)

    # ---- the same workload compressed by ACA at tol 1e-10 ---------------
    if "aca" in legs:
        progress("aca leg")
        aopts = L.kmh_build_opts()
        aopts.method = L.KMH_ACA
        aopts.level_kind = L.KMH_FIXED_LEVEL
        aopts.level = args.level
        aopts.maxrank = 0
        aopts.tol = 1e-10
        aopts.precision = L.KMH_FP64
        kern = kernel_of((1.0, 0.1))          # sigma 1, rho 0.1
        times = []
        for rep_ in range(3):                 # one warm-up, two timed
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            h = C.c_void_p()
            check(lib.kmh_hodlr_build_dev(ctx.ptr, pts_d.data_ptr(), n, 2, C.byref(kern), C.byref(aopts), C.byref(h)))
            check(lib.kmh_ctx_synchronize(ctx.ptr))
            t1 = time.perf_counter()
            check(lib.kmh_hodlr_factorize(h))
            ld = C.c_double()
            check(lib.kmh_hodlr_logdet(h, C.byref(ld)))
            check(lib.kmh_hodlr_solve_dev(h, y_d.data_ptr(), sol_d.data_ptr(), 1))
            check(lib.kmh_ctx_synchronize(ctx.ptr))
            t2 = time.perf_counter()
            ii = L.kmh_info()
            check(lib.kmh_hodlr_info(h, C.byref(ii)))
            nb = 2 ** ii.levels - 1
            rk = (C.c_int32 * nb)()
            check(lib.kmh_hodlr_ranks(h, rk))
            check(lib.kmh_hodlr_destroy(h))
            if rep_:
                times.append((t1 - t0, t2 - t1))
        bt = min(t[0] for t in times)
        ft = min(t[1] for t in times)
        out["aca"] = {"evals_per_s": 1.0 / (bt + ft), "build_s": bt, "factorize_logdet_solve_s": ft,
                      "tol": 1e-10, "max_rank": max(rk), "top_level_rank": rk[nb - 1], "logdet": ld.value,
                      "note": "sigma 1, rho 0.1, best of two timed runs after one warm-up"}

    # ---- follow-up solves on a resident factor: eager vs captured graphs --
    if "followup" in legs:
        progress("followup solves")
        h = C.c_void_p()
        kern = kernel_of(params[0])
        check(lib.kmh_hodlr_build_dev(ctx.ptr, pts_d.data_ptr(), n, 2, C.byref(kern), C.byref(opts), C.byref(h)))
        check(lib.kmh_hodlr_factorize(h))
        fo = {}
        for mode, en in (("eager", 0), ("graph", 1)):
            check(lib.kmh_hodlr_solve_graphs(h, en, None, None))
            for _ in range(3):
                check(lib.kmh_hodlr_solve_dev(h, y_d.data_ptr(), sol_d.data_ptr(), 1))
            check(lib.kmh_ctx_synchronize(ctx.ptr))
            lat = []
            t0 = time.perf_counter()
            for _ in range(50):
                s_ = time.perf_counter()
                check(lib.kmh_hodlr_solve_dev(h, y_d.data_ptr(), sol_d.data_ptr(), 1))
                check(lib.kmh_ctx_synchronize(ctx.ptr))
                lat.append(1e3 * (time.perf_counter() - s_))
            fo[mode] = {"solves_per_s": 50 / (time.perf_counter() - t0), "p50_ms": pct(lat, 0.5),
                        "p99_ms": pct(lat, 0.99)}
        cap, rep_ = C.c_int64(), C.c_int64()
        check(lib.kmh_hodlr_solve_graphs(h, -1, C.byref(cap), C.byref(rep_)))
        check(lib.kmh_hodlr_destroy(h))
        out["followup_solve"] = {"eager_solves_per_s": fo["eager"]["solves_per_s"],
                                 "eager_p50_ms": fo["eager"]["p50_ms"],
                                 "graph_solves_per_s": fo["graph"]["solves_per_s"],
                                 "graph_p50_ms": fo["graph"]["p50_ms"], "graph_p99_ms": fo["graph"]["p99_ms"],
                                 "graph_captures": cap.value, "graph_replays": rep_.value,
                                 "note": "K^-1 y on a resident FP64 factor, synchronized per solve"}

    # ---- open-loop Poisson serving (n log-uniform in [4096, 262144]) -----
    if "poisson" in legs:
        progress("poisson serving leg")
        out["poisson"] = poisson_leg(args, lib, L, ctx, kernel_of, check)

    head = (out.get("served") or out.get("plain") or out.get("served_grad") or out.get("plain_grad")
            or out.get("roofline_pass") or out.get("served_mixed") or out.get("aca") or out.get("poisson")
            or {"evals_per_s": None})

    KERNELS = {
        "tile_qr": ("k_tile_qr<24, double, *>",
                    "Householder QR of <= 256 x 24 tiles (four rows per lane; tiles of <= 128 rows two) with Q "
                    "left in compact-WY form (TSQR leaves and inner nodes), one wave per tile; the leaf tiles of "
                    "a panel's second and third column blocks remove the reorthogonalisation projection Q T2 as "
                    "they load; bytes = tiles and prior columns read, Householder vectors and R written; bound "
                    "by LDS traffic and dependent latency, so below its HBM roofline"),
        "nn_wide": ("k_gemm_nn_wide<0>", "skinny FP64 GEMM on v_mfma_f64_16x16x4f64 (Woodbury updates, leaf-solve "
                                      "updates, Nystrom product); bytes = A, B, C written, C read when beta != 0"),
    }

    def roofline(h):
        """Roofline of the dominant kernel: of the two kernels timed live
        (k_tile_qr<24>, k_gemm_nn_wide), the one with more kernel time in the
        roofline pass (one served batch on one lane, every launch timed with
        HIP event markers around it), or else in the headline leg (every
        PROF_STRIDE-th launch, the other lane's kernels beside it).  Both are
        HBM-bound in the ideal (a 24-column panel QR does about 6 flops per
        byte, under the 78.6 TFLOP/s / 8 TB/s balance of MI355X), so the
        bound is the 8 TB/s HBM peak.  achieved = algorithmic bytes per
        launch / average launch time; traffic = measured HBM bytes per launch
        (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel) from
        the committed summary.  The other kernel's figures ride along."""
        where = "headline leg, lane 0, every %d-th launch" % PROF_STRIDE
        if "roofline_pass" in out:
            h = out["roofline_pass"]
            where = "roofline pass: one served batch of %d requests on one lane, every launch" % h["requests"]
        keys = [k for k in ("tile_qr", "nn_wide") if h.get(k + "_ms")]
        if not keys:
            return {"bound": "hbm", "kernel": None, "achieved": None, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": None, "traffic": None}

        def one(k):
            name, what = KERNELS[k]
            return {"bound": "hbm", "kernel": name, "what": what,
                    "achieved": round(h[k + "_gbps"], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(h[k + "_gbps"] / HBM_PEAK_GBPS, 4),
                    "algorithmic_bytes_per_launch": round(h[k + "_bytes_per_launch"]),
                    "us_per_launch": round(h[k + "_us_per_launch"], 2), "kernel_ms": round(h[k + "_ms"], 3),
                    "timed_launches": h[k + "_timed_launches"], "measured_in": where,
                    "traffic": measured_traffic(name)[0], "traffic_source": measured_traffic(name)[1]}
        keys.sort(key=lambda k: -h[k + "_ms"])
        r = one(keys[0])
        if len(keys) > 1:
            r["other_timed_kernel"] = one(keys[1])
        r["gemm_tflops_nn_wide"] = round(h.get("nn_wide_tflops", 0.0), 3)
        r["fp64_mfma_peak_tflops"] = FP64_MFMA_PEAK_TFLOPS
        r["factor_flops_structured_per_eval"] = ff
        return r

    L_ = info.get("levels", args.level)
    nblocks = [2 ** (L_ - l - 1) for l in range(L_)]
    m = n // (2 ** L_)
    ff = factor_flops(n, m, [72] * L_, nblocks)
    line = {
        "metric": "GP loglik evals/s + p99 latency, n=2^16, served vs plain GPU at 1e-8 rel err",
        "value": round(head["evals_per_s"], 4) if head["evals_per_s"] else None,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / head["evals_per_s"], 3) if head["evals_per_s"] else None,
        "p50_ms": round(head["p50_ms"], 3) if "p50_ms" in head else None,
        "p99_ms": round(head["p99_ms"], 3) if "p99_ms" in head else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "n=65536 Matern-3/2 (sigma~U[0.5,2], rho~U[0.05,0.2], nugget 1e-6) "
                               "Nystrom r=72 FixedLevel(8) negloglik requests, %s form"
                               % ("served+batched FP64" if "served" in out else "plain FP64"),
                   "n": n, "leaf": m, "nystrom_rank": 72, "levels": L_, "inflight": args.inflight,
                   "server_lanes": args.lanes or (2 if args.inflight >= 2 else 1)},
        "legs": {k: {kk: (float("%.5g" % v) if isinstance(v, float) else v) for kk, v in d.items()}
                 for k, d in out.items()},
        "roofline": roofline(head),
    }
    if "served" in out and "plain" in out:
        line["speedup_served_vs_plain"] = round(out["served"]["evals_per_s"] / out["plain"]["evals_per_s"], 3)
    if "served_grad" in out and "plain_grad" in out:
        line["speedup_served_vs_plain_with_gradient"] = round(
            out["served_grad"]["evals_per_s"] / out["plain_grad"]["evals_per_s"], 3)
    if not args.no_cpu:
        # the oracle at the same n = 65536 workload
        progress("cpu baseline (the oracle at n=%d)" % n)
        res, nth, model = cpu_baseline(n, args.level, params[args.warmup], full=args.cpu_full)
        cb = {"value": round(1.0 / res[nth], 6), "unit": "evals/s", "cores": nth, "kind": "port"}
        if 1 in res:
            cb["value_1_core"] = round(1.0 / res[1], 6)
        how = ("median of 3 after 1 warm-up: %.2f s on %d threads, %.2f s on 1 thread" % (res[nth], nth, res[1])
               if args.cpu_full else "one eval on %d threads: %.2f s" % (nth, res[nth]))
        cb["sample"] = ("oracle (numpy/scipy restatement of the reference's op sequence), negloglik eval (build + "
                        "symmetricfactorize! + logdet + y'K^-1 y) of the bench workload itself, n=%d FixedLevel(%d) "
                        "Nystrom 72, %s; host %s, %d threads available"
                        % (n, args.level, how, model, len(os.sched_getaffinity(0))))
        line["cpu_baseline"] = cb
    print(json.dumps(line))


if __name__ == "__main__":
    main()
