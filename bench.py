#!/usr/bin/env python3
"""Benchmark of the per-partition graph operator path on one MI355X.

Default workload (the configuration the headline GTEPS is quoted on):
PageRank, pull form, on RMAT scale 24 (edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 1, duplicates and self-loops removed), one compute
partition and four memory partitions resident in HBM, with the on-device
bitvector master/mirror exchange; FP32 ranks with Kahan-compensated
contributions, FP64 dangling mass and residual, damping 0.85, converged when
the L1 change drops below 1e-6.

A step is one complete PageRank solve (init + iterations to convergence) on
the resident graph; GTEPS = arcs x iterations / time.  The dominant kernel
(the pull over the memory partitions' CSC) is bracketed by HIP events on its
own stream inside the timed steps for the roofline figure.

    python bench.py [--steps K] [--warmup W] [--scale S] [--parts M]
                    [--workload pr_pull|cc|sssp|pr] [--no-cpu] [--no-check]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(1, os.path.join(ROOT, "tests"))  # oracle legs only (FP64 check, CPU baseline)

METRIC = "GTEPS on 1 MI355X, RMAT PR/CC/SSSP; learned-vs-plain schedule speedup"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def traffic_from_profiles(kernel_prefix):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary (profiles/pmc_<kernel>.md: 'traffic_bytes_per_launch: N')."""
    path = os.path.join(ROOT, "profiles", "pmc_%s.md" % kernel_prefix)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def calibrate_copy(mib):
    """Device-to-device copy of a cold buffer of known size (the runtime's
    copy kernel); run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it gives
    the counters' scale factor: bytes counted / bytes copied."""
    import ctypes as ct

    hip = ct.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ct.c_void_p(), ct.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ct.byref(p), ct.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ct.c_size_t(n))
    # flush the caches' view of `a` with an unrelated buffer larger than the
    # 256 MB last-level cache before the measured copy
    c = ct.c_void_p()
    hip.hipMalloc(ct.byref(c), ct.c_size_t(1 << 30))
    hip.hipMemset(c, 2, ct.c_size_t(1 << 30))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ct.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    for p in (a, b, c):
        hip.hipFree(p)
    log("calibration copy: %d MiB device-to-device" % mib)


def schedule_speedup(api, ctx, scale, parts):
    """Learned vs plain (and heuristic) schedule on RMAT-`scale`: CC on the
    symmetrized graph, SSSP hop count from gid 0 and the reference delta
    PageRank on the directed graph.  Speedups are time ratios of the same
    work; GTEPS follow the definitions of the PR (arcs x plain iterations /
    time), CC and SSSP (arcs / time) workloads."""
    import numpy as np

    wpath = os.path.join(ROOT, "grudon_amd", "selector_bf16.bin")
    ctx.selector_load(np.fromfile(wpath, dtype="<u2"), 0.0)
    out = {"graph": "RMAT-%d EF16 seed 1, 1 compute + %d memory partitions" % (scale, parts)}
    scheds = [("plain", api.SCHED_PLAIN), ("heuristic", api.SCHED_HEURISTIC), ("learned", api.SCHED_LEARNED)]
    for sym, algo, name, kw in [(True, api.CC, "cc", {}), (False, api.SSSP, "sssp", {"source": 0}),
                                (False, api.PR, "pr", {"max_iter": 1000})]:
        G = api.Graph.rmat(ctx, scale, 16, seed=1, symmetrize=sym, C=1, M=parts)
        S = api.State(G, algo)
        ms = {}
        iters_plain = None
        for label, sc in scheds:
            S.init(sched=sc, **kw)
            S.run()
            best = None
            for _ in range(3):
                S.init(sched=sc, **kw)
                t = time.perf_counter()
                st = S.run()
                dt = (time.perf_counter() - t) * 1e3
                best = dt if best is None else min(best, dt)
            ms[label] = best
            if sc == api.SCHED_PLAIN:
                iters_plain = st["iterations"]
        work = G.m * (iters_plain if algo == api.PR else 1)
        out[name] = {"ms": {k: round(v, 3) for k, v in ms.items()},
                     "gteps": {k: round(work / (v * 1e-3) / 1e9, 2) for k, v in ms.items()},
                     "learned_vs_plain": round(ms["plain"] / ms["learned"], 3),
                     "learned_vs_heuristic": round(ms["heuristic"] / ms["learned"], 3)}
        S.free()
        G.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--parts", type=int, default=4)
    ap.add_argument("--compute", type=int, default=1)
    ap.add_argument("--workload", default="pr_pull", choices=["pr_pull", "cc", "sssp", "pr"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 / bit-exact check")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--speedup-scale", type=int, default=22,
                    help="RMAT scale of the learned-vs-plain schedule leg (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a cold device copy of MIB MiB (PMC counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU (partitions are resident in one HBM); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from grudon_amd import api

    ctx = api.Context(0)
    t0 = time.perf_counter()
    sym = args.workload == "cc"
    rp, col = ctx.generate_rmat(args.scale, args.ef, seed=1, symmetrize=sym)
    t_gen = time.perf_counter() - t0
    wts = None
    if args.workload == "sssp":
        sym = True
        rp, col = ctx.generate_rmat(args.scale, args.ef, seed=1, symmetrize=True)
    t1 = time.perf_counter()
    G = api.Graph.from_csr(ctx, rp, col, wts, C=args.compute, M=args.parts)
    t_part = time.perf_counter() - t1
    m = G.m
    log("graph: n=%d m=%d proxies=%d cross=%d gen %.1fs partition+upload %.1fs device %.2f GB"
        % (G.n, m, G.proxies, G.cross, t_gen, t_part, G.device_bytes / 1e9))

    algo = {"pr_pull": api.PR_PULL, "cc": api.CC, "sssp": api.SSSP, "pr": api.PR}[args.workload]
    S = api.State(G, algo)
    sched = api.SCHED_HEURISTIC

    def one_step(time_kernel):
        S.init(max_iter=1000, tolerance=1e-6, sched=sched, time_kernel=time_kernel)
        return S.run()

    for _ in range(args.warmup):
        one_step(False)
    t_start = time.perf_counter()
    iters = 0
    kms = 0.0
    klaunch = 0
    kbytes = 0
    for _ in range(args.steps):
        st = one_step(True)
        iters += st["iterations"]
        kms += st["kernel_ms"]
        klaunch += st["kernel_launches"]
        kbytes += st["kernel_bytes"]
    t_total = time.perf_counter() - t_start
    last = st
    gteps = m * iters / t_total / 1e9
    ms_per_step = t_total * 1e3 / args.steps
    log("timed %d steps: %.3f s, %d iterations total, %.2f ms/step, %.1f GTEPS" %
        (args.steps, t_total, iters, ms_per_step, gteps))

    achieved = (kbytes / klaunch) / (kms / klaunch * 1e-3) / 1e9 if klaunch else 0.0
    kname = {"pr_pull": "k_prp_pull", "cc": "k_pull", "sssp": "k_pull", "pr": "k_pull"}[args.workload]
    traffic = traffic_from_profiles(kname + "_rmat%d" % args.scale)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname, "launches": klaunch,
                "avg_launch_us": round(kms / klaunch * 1e3, 2) if klaunch else None,
                "algorithmic_bytes_per_launch": int(kbytes / klaunch) if klaunch else None}

    extra = {}
    if not args.no_check and args.workload == "pr_pull":
        import oracle_check
        res = oracle_check.pr_pull_l1(rp, col, S.result(), last["iterations"], args.cpu_threads)
        extra["fp64_l1"] = res
        log("FP64 oracle L1 = %.3e" % res)

    # free the headline workload before the schedule-comparison leg
    S.free()
    G.free()
    if args.speedup_scale:
        extra["learned_vs_plain"] = schedule_speedup(api, ctx, args.speedup_scale, args.parts)
        log("learned vs plain: %s" % json.dumps(extra["learned_vs_plain"]))

    cpu = None
    if not args.no_cpu:
        import oracle_check
        model, cores = cpu_info()
        nthreads = args.cpu_threads or min(cores, 64)
        cpu = oracle_check.cpu_baseline(rp, col, args.compute, args.parts, nthreads, algo)
        cpu["cores"] = nthreads
        cpu["cpu_model"] = model
        log("cpu baseline: %s" % json.dumps(cpu))

    out = {
        "metric": METRIC, "value": round(gteps, 2), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32" if algo in (0, 3) else "u32",
        "data": "synthetic",
        "config": {"workload": {"pr_pull": "PageRank pull (FP32 Kahan, FP64 dangling+residual, d=0.85, tol 1e-6)",
                                "cc": "CC label propagation", "sssp": "SSSP hop count from gid 0",
                                "pr": "reference delta-push PageRank"}[args.workload]
                   + ", RMAT-%d EF%d seed 1%s, %d compute + %d memory partitions resident in HBM, "
                     "on-device bitvector master/mirror sync" % (args.scale, args.ef, " symmetrized" if sym else "",
                                                                   args.compute, args.parts),
                   "n": G.n, "arcs": m, "iterations_per_step": last["iterations"],
                   "schedule": "heuristic", "exchange_bytes_per_step": last["bytes_exchanged"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
