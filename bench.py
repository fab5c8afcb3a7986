#!/usr/bin/env python3
"""Benchmark of the per-partition graph operator path on one MI355X.

Default workload (`headline`, the configuration the metric is quoted on):
RMAT scale 26 (67,108,864 vertices, edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 1, duplicates and self-loops removed, 1.06 B arcs),
one compute partition and four memory partitions resident in HBM, with the
on-device bitvector master/mirror exchange.

  value    PageRank, pull form: FP32 ranks with Kahan-compensated
           contributions, FP64 dangling mass and residual, damping 0.85,
           converged when the L1 change drops below 1e-6.  A step is one
           complete solve (init + iterations) on the resident graph; GTEPS =
           arcs x iterations / time.  Checked: L1 against the oracle's FP64
           power iteration run for the same number of iterations.
  legs     on the same scale, same partitioning, one JSON line:
           cc        CC on RMAT-26 symmetrized (2.1 B arcs); GTEPS = arcs / time
                     of one solve; labels checked against a parallel union-find
           sssp      weighted SSSP (weights [1,255] seed 2, delta 256) from 16
                     sources (seed 3, degree >= 1) on the same graph; GTEPS =
                     harmonic mean of arcs / time; every distance vector
                     checked against the Bellman equations
           learned_vs_plain  plain vs heuristic vs learned schedule on the
                     directed graph with weights: reference PageRank, CC,
                     weighted SSSP from gid 0; best of 3 solves each
  roofline the pull kernel of the PageRank solves (HIP events on the
           context's stream around every launch); each leg carries its own
  cpu_baseline  the oracle's threaded FP64 pull PageRank on the same graph
           (2 iterations per run, median of 5 after a warm-up), and beside it
           the restatement of the reference's delta-push superstep

Other workloads (one JSON line each, same fields), e.g. for profiling:
  pr_pull  the headline's PageRank alone (default scale 24 unless --scale)
  cc       CC on RMAT-S symmetrized; labels checked
  sssp     weighted SSSP from 16 sources on RMAT-S symmetrized; checked
  pr       the reference's delta-push PageRank (FP32) on RMAT-S
  batch    64 graphs of uneven size (32 RMAT scale 14-20 EF 8-32, 32
           Erdos-Renyi n 2^14-2^20 mean degree 4-64, seed 4) in one batch,
           autotuned; step = PageRank + CC on all of them
  learned  learned vs heuristic vs plain schedule (RMAT-26 by default)
  schedules  the reference's offloaded push, its NDPEngine vote, edge
           shipping every superstep and the heuristic: CC, SSSP hop and
           weighted; value = geometric mean speedup of edge shipping
  selector  every option timed at every superstep of held-out CC / SSSP
           runs: the hindsight-best schedule against the heuristic and the
           learned selector; value = heuristic / best (the ceiling of any
           per-superstep selector's gain over the heuristic)

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--scale S]
                    [--parts M] [--no-cpu] [--no-check]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(1, os.path.join(ROOT, "tests"))  # the cpu_baseline leg only

METRIC = "GTEPS on 1 MI355X, RMAT PR/CC/SSSP; learned-vs-plain schedule speedup"
HBM_PEAK_GBS = 8000.0
# the kernels of one pull step per pull mode, timed together between HIP events
PULL_KERNELS = {"plain": "k_prp_pull_rb (+ k_prp_pull_fixup)",
                "hotcold": "k_pull_hot<HotPR> + k_prp_pull_rb (+ fix-ups)",
                "tiled": "k_prp_pull_tiled + k_prp_tier_merge (+ fix-ups)",
                "segmented": "k_prp_pull_seg + k_prp_tier_merge (+ fix-ups)"}
INF = 0xFFFFFFFF


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


def cpu_threads(args):
    model, cores = cpu_info()
    return model, (args.cpu_threads or min(cores, 64))


def host_desc():
    """Affinity core count and NUMA nodes of the host the CPU baseline ran on."""
    _, cores = cpu_info()
    try:
        nodes = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and d[4:].isdigit()])
    except OSError:
        nodes = 0
    return {"affinity_cores": cores, "numa_nodes": nodes}


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


# ---------------------------------------------------------------------------
# full-size checks (numpy / scipy, independent of the library)
# ---------------------------------------------------------------------------

def fp64_pagerank(rp, col, iters, d=0.85):
    """FP64 power iteration r' = (1-d)/n + d (A^T (r/deg) + D/n), r0 = 1/n,
    run for exactly `iters` iterations."""
    import numpy as np
    import scipy.sparse as sp

    n = len(rp) - 1
    rp64 = rp.astype(np.int64)
    deg = np.diff(rp64).astype(np.float64)
    sink = deg == 0
    A = sp.csr_matrix((np.ones(len(col), dtype=np.float64), col, rp64), shape=(n, n))
    AT = A.T  # CSC view: A^T x without an explicit transpose
    r = np.full(n, 1.0 / n)
    inv = np.where(sink, 0.0, 1.0 / np.where(sink, 1.0, deg))
    for _ in range(int(iters)):
        D = r[sink].sum()
        r = (1.0 - d) / n + d * (AT @ (r * inv) + D / n)
    return r


def cc_labels_scipy(rp, col):
    """Min gid of every vertex's connected component (symmetric input)."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    n = len(rp) - 1
    A = sp.csr_matrix((np.ones(len(col), dtype=np.int8), col, rp.astype(np.int64)), shape=(n, n))
    _, comp = cg.connected_components(A, directed=False)
    mins = np.full(comp.max() + 1, np.iinfo(np.int64).max, dtype=np.int64)
    np.minimum.at(mins, comp, np.arange(n, dtype=np.int64))
    return mins[comp].astype(np.uint32)


# ---------------------------------------------------------------------------
# common pieces
# ---------------------------------------------------------------------------

def roofline_of(kms, klaunch, kbytes, kname, traffic):
    achieved = (kbytes / klaunch) / (kms / klaunch * 1e-3) / 1e9 if klaunch else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kname, "launches": klaunch,
            "avg_launch_us": round(kms / klaunch * 1e3, 2) if klaunch else None,
            "algorithmic_bytes_per_launch": int(kbytes / klaunch) if klaunch else None}


def line(args, value, unit, ms_per_step, dtype, config, roofline, cpu, extra=None, higher=True):
    out = {"metric": METRIC, "value": round(value, 3), "unit": unit, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}
    if extra:
        out.update(extra)
    return out


def layout_desc(args, scale, sym):
    return ("RMAT-%d EF%d seed 1%s, %d compute + %d memory partitions resident in HBM, "
            "on-device bitvector master/mirror sync" % (scale, args.ef, " symmetrized" if sym else "",
                                                         args.compute, args.parts))


def timed_solves(S, args, init_kw, sources=None, exact_bytes=False):
    """Warm-up, then exactly args.steps timed steps; a step is one solve (or
    one solve per source).  Returns (seconds, iterations, kernel stats, last
    run stats); the last run's stats also carry "timed_edges_examined", the
    edges the traversals examined over all timed steps (for work GTEPS).
    With exact_bytes (CC / SSSP, whose pulls leave out gathers), the timed
    steps time the kernels with events only (time_kernel 2) and one more,
    untimed step counts their algorithmic bytes exactly (time_kernel 1); the
    solves are deterministic, so its bytes per launch are the timed ones'."""
    srcs = sources if sources is not None else [None]

    def step(time_kernel):
        acc = [0, 0.0, 0, 0, 0]
        st = None
        for s in srcs:
            kw = dict(init_kw)
            if s is not None:
                kw["source"] = int(s)
            S.init(time_kernel=time_kernel, **kw)
            st = S.run()
            acc[0] += st["iterations"]
            acc[1] += st["kernel_ms"]
            acc[2] += st["kernel_launches"]
            acc[3] += st["kernel_bytes"]
            acc[4] += st["edges_examined"]
        return acc, st

    for _ in range(args.warmup):
        step(False)
    iters, kms, kl, kb, ee = 0, 0.0, 0, 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        acc, last = step(2 if exact_bytes else 1)
        iters += acc[0]
        kms += acc[1]
        kl += acc[2]
        kb += acc[3]
        ee += acc[4]
    secs = time.perf_counter() - t0
    if exact_bytes and kl:
        acc, _ = step(1)
        if acc[2]:
            kb = int(round(acc[3] / acc[2] * kl))
    last["timed_edges_examined"] = ee
    return secs, iters, (kms, kl, kb), last


def schedule_speedup(api, ctx, scale, parts, ef=16, weighted=False, reps=3, sym_cc=True, graph=None):
    """Learned vs plain (and heuristic) schedule on RMAT-`scale`: CC (on the
    symmetrized graph when sym_cc), SSSP from gid 0 (weighted or hop count)
    and the reference delta PageRank on the directed graph.  Speedups are
    time ratios of the same work; GTEPS follow the definitions of the PR
    (arcs x plain iterations / time), CC and SSSP (arcs / time) workloads.
    Results are compared across the schedules (bit for bit for CC / SSSP)."""
    import numpy as np

    wpath = os.path.join(ROOT, "grudon_amd", "selector_bf16.bin")
    ctx.selector_load(np.fromfile(wpath, dtype="<u2"), api.SELECTOR_MARGIN)
    out = {"graph": "RMAT-%d EF%d seed 1, 1 compute + %d memory partitions%s"
                    % (scale, ef, parts, ", weights [1,255] seed 2" if weighted else "")}
    scheds = [("plain", api.SCHED_PLAIN), ("heuristic", api.SCHED_HEURISTIC), ("learned", api.SCHED_LEARNED)]
    jobs = [(sym_cc, api.CC, "cc", {}), (False, api.SSSP, "sssp", {"source": 0, "weighted": weighted}),
            (False, api.PR, "pr", {"max_iter": 1000})]
    G = graph
    G_sym = None
    for sym, algo, name, kw in jobs:
        if sym:
            if G_sym is None:
                G_sym = api.Graph.rmat(ctx, scale, ef, seed=1, symmetrize=True, wseed=2 if weighted else 0,
                                       C=1, M=parts)
            Gx = G_sym
        else:
            if G is None:
                if G_sym is not None:
                    G_sym.free()
                    G_sym = None
                G = api.Graph.rmat(ctx, scale, ef, seed=1, symmetrize=False, wseed=2 if weighted else 0,
                                   C=1, M=parts)
            Gx = G
        S = api.State(Gx, algo)
        ms = {}
        res = {}
        iters_plain = None
        for label, sc in scheds:
            S.init(sched=sc, **kw)
            S.run()
            best = None
            for _ in range(reps):
                S.init(sched=sc, **kw)
                t = time.perf_counter()
                st = S.run()
                dt = (time.perf_counter() - t) * 1e3
                best = dt if best is None else min(best, dt)
            ms[label] = best
            res[label] = S.result()
            if sc == api.SCHED_PLAIN:
                iters_plain = st["iterations"]
        if algo == api.PR:
            base = res["plain"].astype(np.float64)
            same = all(np.abs(r.astype(np.float64) / r.sum() - base / base.sum()).sum() < 1e-5
                       for r in res.values())
        else:
            same = all(np.array_equal(r, res["plain"]) for r in res.values())
        work = Gx.m * (iters_plain if algo == api.PR else 1)
        out[name] = {"arcs": Gx.m, "ms": {k: round(v, 3) for k, v in ms.items()},
                     "gteps": {k: round(work / (v * 1e-3) / 1e9, 2) for k, v in ms.items()},
                     "learned_vs_plain": round(ms["plain"] / ms["learned"], 3),
                     "learned_vs_heuristic": round(ms["heuristic"] / ms["learned"], 3),
                     "results_identical": bool(same)}
        log("schedule %s: %s" % (name, json.dumps(out[name])))
        S.free()
    for g in (G, G_sym):
        if g is not None and g is not graph:
            g.free()
    return out


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def gather_ceiling(api, ctx, n, m, parts, kms, kl):
    """The pull's arc rate (arcs per second of one pull step: hot arcs read
    from LDS, cold ones gathered through the L2) beside the device's
    uniform-random gather ceiling for a table of one memory partition's own
    vertices and for an L2-resident table (4 MB)."""
    ceil = {"pull_gathers_per_s": round(m / (kms / kl * 1e-3), 1)}
    for name, entries in (("partition_table", -(-n // parts)), ("l2_table", 1 << 20)):
        _, rate = api.probe_gather(ctx.device, entries)
        ceil[name] = {"entries": entries, "uniform_random_gathers_per_s": round(rate, 1)}
    log("gather ceiling: %s" % json.dumps(ceil))
    return ceil


def run_pagerank(api, ctx, args):
    import numpy as np

    pull = args.workload == "pr_pull"
    t0 = time.perf_counter()
    rp, col = ctx.generate_rmat(args.scale, args.ef, seed=1)
    t_gen = time.perf_counter() - t0
    t1 = time.perf_counter()
    G = api.Graph.from_csr(ctx, rp, col, C=args.compute, M=args.parts)
    t_part = time.perf_counter() - t1
    log("graph: n=%d m=%d proxies=%d cross=%d gen %.1fs partition+upload %.1fs device %.2f GB"
        % (G.n, G.m, G.proxies, G.cross, t_gen, t_part, G.device_bytes / 1e9))
    algo = api.PR_PULL if pull else api.PR
    S = api.State(G, algo)
    secs, iters, (kms, kl, kb), last = timed_solves(
        S, args, {"max_iter": 1000, "tolerance": 1e-6, "sched": api.SCHED_HEURISTIC,
                  "offload_only": args.offload_only})
    gteps = G.m * iters / secs / 1e9
    log("timed %d steps: %.3f s, %d iterations, %.2f ms/step, %.1f GTEPS"
        % (args.steps, secs, iters, secs * 1e3 / args.steps, gteps))
    kname = PULL_KERNELS[args.pull] if pull else "k_bsp_traverse"
    traffic = traffic_from_profiles("k_prp_pull_%s_rmat%d" % (args.pull, args.scale)) if pull else None
    roof = roofline_of(kms, kl, kb, kname, traffic)
    # edges the traversals examined (a pull examines every arc) per second
    extra = {"work_gteps": round(last["timed_edges_examined"] / secs / 1e9, 3)}
    plan = None
    if pull and (not args.no_check or not args.no_cpu):
        import oracle_lib as O
        plan = O.PullPlan(rp, col, cpu_threads(args)[1])
    if pull and not args.no_check:
        ref, _, _, _ = plan.run(iters=last["iterations"])
        extra["fp64_l1"] = float(np.abs(S.result().astype(np.float64) - ref).sum())
        log("FP64 L1 = %.3e (bar 1e-6)" % extra["fp64_l1"])
    n, m, bytes_x = G.n, G.m, last["bytes_exchanged"]
    S.free()
    G.free()
    if pull and kl:
        extra["gather_ceiling"] = gather_ceiling(api, ctx, n, m, args.parts, kms, kl)
    if pull and args.speedup_scale:
        extra["learned_vs_plain"] = schedule_speedup(api, ctx, args.speedup_scale, args.parts)
    cpu = None
    if not args.no_cpu:
        import oracle_check
        model, nt = cpu_threads(args)
        if pull:
            cpu = oracle_check.cpu_baseline_pull(plan, nt)
        else:
            cpu = oracle_check.cpu_baseline(rp, col, args.compute, args.parts, nt, api.PR)
        cpu.update(cores=nt, cpu_model=model, **host_desc())
        log("cpu baseline: %s" % json.dumps(cpu))
    if plan is not None:
        plan.free()
    wl = ("PageRank pull (FP32 Kahan, FP64 dangling+residual, d=0.85, tol 1e-6)" if pull
          else "reference delta-push PageRank (FP32, d=0.85, per-vertex threshold 1e-6)")
    config = {"workload": wl + ", " + layout_desc(args, args.scale, False), "n": n, "arcs": m,
              "iterations_per_step": last["iterations"], "schedule": "heuristic",
              "exchange_bytes_per_step": bytes_x}
    return line(args, gteps, "GTEPS", secs * 1e3 / args.steps, "f32", config, roof, cpu, extra)


def run_cc(api, ctx, args):
    import numpy as np

    rp, col = ctx.generate_rmat(args.scale, args.ef, seed=1, symmetrize=True)
    t1 = time.perf_counter()
    G = api.Graph.from_csr(ctx, rp, col, C=args.compute, M=args.parts)
    log("graph: n=%d m=%d partition+upload %.1fs" % (G.n, G.m, time.perf_counter() - t1))
    S = api.State(G, api.CC)
    secs, iters, (kms, kl, kb), last = timed_solves(S, args, {"sched": api.SCHED_HEURISTIC,
                                                              "offload_only": args.offload_only,
                                                              "pull_every_row": args.pull_every_row},
                                                    exact_bytes=True)
    gteps = G.m * args.steps / secs / 1e9
    log("timed %d steps: %.3f s, %.1f GTEPS" % (args.steps, secs, gteps))
    # edges the traversals examined per second (a CC / SSSP pull leaves out
    # the row blocks whose rows cannot improve)
    extra = {"work_gteps": round(last["timed_edges_examined"] / secs / 1e9, 3)}
    if not args.no_check:
        import oracle_check
        extra["labels_bit_exact"] = oracle_check.cc_exact(rp, col, S.result(), cpu_threads(args)[1])
        log("labels bit-exact vs the parallel union-find: %s" % extra["labels_bit_exact"])
    roof = roofline_of(kms, kl, kb, "k_bsp_traverse", None)
    n, m = G.n, G.m
    S.free()
    G.free()
    cpu = None
    if not args.no_cpu:
        import oracle_check
        model, nt = cpu_threads(args)
        cpu = oracle_check.cpu_baseline(rp, col, args.compute, args.parts, nt, api.CC)
        cpu.update(cores=nt, cpu_model=model, **host_desc())
    config = {"workload": "CC label propagation (u32 labels), " + layout_desc(args, args.scale, True),
              "n": n, "arcs": m, "iterations_per_step": last["iterations"], "schedule": "heuristic",
              "gteps_definition": "arcs / time of one solve"}
    return line(args, gteps, "GTEPS", secs * 1e3 / args.steps, "u32", config, roof, cpu, extra)


def run_sssp(api, ctx, args):
    import numpy as np

    rp, col = ctx.generate_rmat(args.scale, args.ef, seed=1, symmetrize=True)
    w = ctx.edge_weights(rp, col, 2)
    deg = np.diff(rp.astype(np.int64))
    rng = np.random.default_rng(3)
    sources = rng.choice(np.flatnonzero(deg > 0), size=16, replace=False)
    t1 = time.perf_counter()
    G = api.Graph.from_csr(ctx, rp, col, w, C=args.compute, M=args.parts)
    log("graph: n=%d m=%d partition+upload %.1fs" % (G.n, G.m, time.perf_counter() - t1))
    S = api.State(G, api.SSSP)
    kw = {"weighted": True, "delta": args.delta, "sched": api.SCHED_HEURISTIC, "offload_only": args.offload_only,
          "pull_every_row": args.pull_every_row}
    secs, iters, (kms, kl, kb), last = timed_solves(S, args, kw, sources=sources, exact_bytes=True)
    # harmonic mean over sources of arcs / t_s = 16 arcs / sum t_s
    gteps = G.m * len(sources) * args.steps / secs / 1e9
    log("timed %d steps x %d sources: %.3f s, %.1f GTEPS" % (args.steps, len(sources), secs, gteps))
    # edges the traversals examined per second (a CC / SSSP pull leaves out
    # the row blocks whose rows cannot improve)
    extra = {"work_gteps": round(last["timed_edges_examined"] / secs / 1e9, 3)}
    if not args.no_check:
        import oracle_check
        nt = cpu_threads(args)[1]
        ok = []
        for s in sources:
            S.init(source=int(s), **kw)
            S.run()
            ok.append(oracle_check.sssp_exact(rp, col, w, S.result(), int(s), nt))
            log("Bellman check, source %d: %s" % (int(s), ok[-1]))
        extra["distances_exact"] = bool(all(ok))
    roof = roofline_of(kms, kl, kb, "k_bsp_traverse", None)
    n, m = G.n, G.m
    S.free()
    G.free()
    cpu = None
    if not args.no_cpu:
        import oracle_check
        model, nt = cpu_threads(args)
        cpu = oracle_check.cpu_baseline(rp, col, args.compute, args.parts, nt, api.SSSP, supersteps=12, w=w,
                                        source=int(sources[0]))
        cpu.update(cores=nt, cpu_model=model, **host_desc())
    config = {"workload": "SSSP delta-stepping (u32 distances, delta %d), weights [1,255] seed 2, 16 sources "
                          "seed 3 among degree >= 1, %s" % (args.delta, layout_desc(args, args.scale, True)),
              "n": n, "arcs": m, "iterations_per_step": iters // max(1, args.steps), "schedule": "heuristic",
              "sources": [int(s) for s in sources],
              "gteps_definition": "harmonic mean over the 16 sources of arcs / time"}
    return line(args, gteps, "GTEPS", secs * 1e3 / args.steps, "u32", config, roof, cpu, extra)


def batch_graphs(ctx, seed=4):
    """32 RMAT (scale 14-20, EF 8-32) and 32 Erdos-Renyi (n 2^14-2^20, mean
    degree 4-64) graphs, drawn from `seed`."""
    import numpy as np

    rng = np.random.default_rng(seed)
    graphs, desc = [], []
    for i in range(32):
        s, ef = int(rng.integers(14, 21)), int(rng.integers(8, 33))
        graphs.append(ctx.generate_rmat(s, ef, seed=1000 + i))
        desc.append("rmat%d/ef%d" % (s, ef))
    for i in range(32):
        n = int(round(2.0 ** rng.uniform(14, 20)))
        d = float(rng.uniform(4, 64))
        graphs.append(ctx.generate_er(n, d, seed=2000 + i))
        desc.append("er%d/%.1f" % (n, d))
    return graphs, desc


def run_batch(api, ctx, args):
    import numpy as np

    graphs, desc = batch_graphs(ctx)
    B = api.Batch(ctx, graphs)
    t0 = time.perf_counter()
    chosen = B.autotune()
    log("autotune %.1fs: %s" % (time.perf_counter() - t0, chosen))
    arcs = [len(c) for _, c in graphs]

    def step():
        it_pr, _ = B.pagerank(tolerance=1e-6, max_iter=1000)
        it_cc, _ = B.cc(max_iter=1000)
        return it_pr, it_cc

    for _ in range(args.warmup):
        step()
    work = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it_pr, it_cc = step()
        work += sum(a * int(i) for a, i in zip(arcs, it_pr)) + sum(a * int(i) for a, i in zip(arcs, it_cc))
    secs = time.perf_counter() - t0
    gteps = work / secs / 1e9
    log("timed %d steps: %.3f s, %.1f GTEPS" % (args.steps, secs, gteps))
    extra = {"autotuned": chosen}
    if not args.no_check:
        # PageRank: every graph within 1e-6 L1 of the FP64 power iteration;
        # CC: the Erdos-Renyi half (symmetric) against scipy components, the
        # RMAT half against the partitioned BSP path of this library
        worst = 0.0
        cc_ok = True
        it_pr, _ = B.pagerank(tolerance=1e-6, max_iter=1000)
        for i, (rp, col) in enumerate(graphs):
            ref = fp64_pagerank(rp, col, int(it_pr[i]))
            worst = max(worst, float(np.abs(B.result(i, api.PR_PULL).astype(np.float64) - ref).sum()))
        B.cc(max_iter=1000)
        for i, (rp, col) in enumerate(graphs):
            lab = B.result(i, api.CC)
            if desc[i].startswith("er"):
                cc_ok &= bool(np.array_equal(lab, cc_labels_scipy(rp, col)))
            else:
                G = api.Graph.from_csr(ctx, rp, col, C=1, M=2)
                S = api.State(G, api.CC)
                S.init(sched=api.SCHED_HEURISTIC)
                S.run()
                cc_ok &= bool(np.array_equal(lab, S.result()))
                S.free()
                G.free()
        extra["pr_fp64_l1_max"] = worst
        extra["cc_labels_exact"] = bool(cc_ok)
        log("batch checks: PR worst L1 %.3e, CC exact %s" % (worst, cc_ok))
    B.free()
    cpu = None
    if not args.no_cpu:
        import oracle_check
        model, nt = cpu_threads(args)
        cpu = oracle_check.cpu_baseline_batch(graphs, nt)
        cpu.update(cores=nt, cpu_model=model, **host_desc())
    config = {"workload": "batched PageRank pull (FP32 Kahan, FP64 residual, tol 1e-6) + CC over 64 graphs "
                          "(32 RMAT scale 14-20 EF 8-32, 32 Erdos-Renyi n 2^14-2^20 degree 4-64, seed 4) in "
                          "persistent launches with hipGraph-captured iterations",
              "graphs": len(graphs), "arcs": int(sum(arcs)), "n": int(sum(len(r) - 1 for r, _ in graphs)),
              "gteps_definition": "sum over graphs of arcs x iterations (PR + CC) / time"}
    return line(args, gteps, "GTEPS", secs * 1e3 / args.steps, "f32", config, None, cpu, extra)


def _timed_leg(api, S, args, init_kw, arcs, sources=None, per_iteration=False, exact_bytes=False):
    """Timed solves of one leg: (GTEPS, ms per step, iterations per step,
    kernel stats of its dominant traversal kernel, last run stats, seconds).
    GTEPS = arcs x iterations / time with per_iteration (PageRank), else
    arcs x solves / time (one solve per source)."""
    secs, iters, (kms, kl, kb), last = timed_solves(S, args, init_kw, sources=sources, exact_bytes=exact_bytes)
    nsrc = 1 if sources is None else len(sources)
    work = arcs * iters if per_iteration else arcs * nsrc * args.steps
    return work / secs / 1e9, secs * 1e3 / args.steps, iters // max(1, args.steps), (kms, kl, kb), last, secs


def run_headline(api, ctx, args):
    """PageRank pull on RMAT-S (value), CC and weighted SSSP on RMAT-S
    symmetrized and the learned-vs-plain schedule speedup on RMAT-S (legs),
    each checked at full size; the CPU baseline of the same pull algorithm."""
    import numpy as np

    import oracle_check
    import oracle_lib as O

    scale = args.scale
    model, nt = cpu_threads(args)
    # ---- directed graph: the PageRank value and the learned leg ----
    t0 = time.perf_counter()
    rp, col = ctx.generate_rmat(scale, args.ef, seed=1)
    w = ctx.edge_weights(rp, col, 2)
    t_gen = time.perf_counter() - t0
    t1 = time.perf_counter()
    G = api.Graph.from_csr(ctx, rp, col, w, C=args.compute, M=args.parts)
    t_part = time.perf_counter() - t1
    log("graph: n=%d m=%d proxies=%d cross=%d gen %.1fs partition+upload %.1fs device %.2f GB"
        % (G.n, G.m, G.proxies, G.cross, t_gen, t_part, G.device_bytes / 1e9))
    S = api.State(G, api.PR_PULL)
    gteps, ms_step, iters, (kms, kl, kb), last, secs = _timed_leg(
        api, S, args, {"max_iter": 1000, "tolerance": 1e-6}, G.m, per_iteration=True)
    pr_iters = last["iterations"]
    log("PR pull: %d steps, %.2f ms/step, %d iterations, %.1f GTEPS" % (args.steps, ms_step, pr_iters, gteps))
    roof = roofline_of(kms, kl, kb, PULL_KERNELS[args.pull],
                       traffic_from_profiles("k_prp_pull_%s_rmat%d" % (args.pull, scale)))
    extra = {"work_gteps": round(last["timed_edges_examined"] / secs / 1e9, 3)}
    if kl:
        extra["gather_ceiling"] = gather_ceiling(api, ctx, G.n, G.m, args.parts, kms, kl)
    plan = None
    if not args.no_check or not args.no_cpu:
        t2 = time.perf_counter()
        plan = O.PullPlan(rp, col, nt)
        log("CPU pull plan (transpose) %.1fs" % (time.perf_counter() - t2))
    if not args.no_check:
        ref, it, _, _ = plan.run(iters=pr_iters)
        extra["fp64_l1"] = float(np.abs(S.result().astype(np.float64) - ref).sum())
        log("FP64 L1 = %.3e (bar 1e-6)" % extra["fp64_l1"])
    S.free()
    cpu = None
    if not args.no_cpu:
        cpu = oracle_check.cpu_baseline_pull(plan, nt)
        cpu.update(cores=nt, cpu_model=model, **host_desc())
        log("cpu baseline (pull): %s" % json.dumps(cpu))
    if plan is not None:
        plan.free()
        plan = None
    legs = {}
    if args.speedup_scale:
        legs["learned_vs_plain"] = schedule_speedup(api, ctx, scale, args.parts, ef=args.ef, weighted=True,
                                                    reps=3, sym_cc=False, graph=G)
    n, m, bytes_x, proxies = G.n, G.m, last["bytes_exchanged"], G.proxies
    G.free()
    del rp, col, w
    # ---- symmetrized graph: CC and weighted SSSP ----
    t0 = time.perf_counter()
    rps, cols = ctx.generate_rmat(scale, args.ef, seed=1, symmetrize=True)
    ws = ctx.edge_weights(rps, cols, 2)
    Gs = api.Graph.from_csr(ctx, rps, cols, ws, C=args.compute, M=args.parts)
    log("symmetrized graph: m=%d, gen + partition %.1fs" % (Gs.m, time.perf_counter() - t0))
    Sc = api.State(Gs, api.CC)
    g_cc, ms_cc, it_cc, kcc, last_cc, secs_cc = _timed_leg(api, Sc, args, {"sched": api.SCHED_HEURISTIC}, Gs.m,
                                                           exact_bytes=True)
    leg = {"gteps": round(g_cc, 3), "ms_per_step": round(ms_cc, 3), "arcs": Gs.m, "iterations": it_cc,
           "work_gteps": round(last_cc["timed_edges_examined"] / secs_cc / 1e9, 3),
           "roofline": roofline_of(*kcc, "k_bsp_traverse", None),
           "workload": "CC label propagation (u32), RMAT-%d EF%d seed 1 symmetrized, heuristic schedule, "
                       "1 step = 1 solve" % (scale, args.ef)}
    if not args.no_check:
        leg["labels_exact"] = oracle_check.cc_exact(rps, cols, Sc.result(), nt)
    log("CC leg: %s" % json.dumps(leg))
    legs["cc"] = leg
    Sc.free()
    deg = np.diff(rps.astype(np.int64))
    rng = np.random.default_rng(3)
    sources = rng.choice(np.flatnonzero(deg > 0), size=16, replace=False)
    Ss = api.State(Gs, api.SSSP)
    kw = {"weighted": True, "delta": args.delta, "sched": api.SCHED_HEURISTIC, "max_iter": 100000}
    g_ss, ms_ss, it_ss, kss, last_ss, secs_ss = _timed_leg(api, Ss, args, kw, Gs.m, sources=sources,
                                                           exact_bytes=True)
    leg = {"gteps": round(g_ss, 3), "ms_per_step": round(ms_ss, 3), "arcs": Gs.m, "iterations_per_step": it_ss,
           "work_gteps": round(last_ss["timed_edges_examined"] / secs_ss / 1e9, 3),
           "roofline": roofline_of(*kss, "k_bsp_traverse", None),
           "sources": [int(x) for x in sources],
           "workload": "SSSP delta-stepping (u32, delta %d), weights [1,255] seed 2, 16 sources seed 3 among "
                       "degree >= 1, RMAT-%d EF%d seed 1 symmetrized, heuristic schedule, 1 step = the 16 "
                       "solves; GTEPS = harmonic mean over the sources of arcs / time" % (args.delta, scale, args.ef)}
    if not args.no_check:
        ok = []
        for x in sources:
            Ss.init(source=int(x), **kw)
            Ss.run()
            ok.append(oracle_check.sssp_exact(rps, cols, ws, Ss.result(), int(x), nt))
        leg["distances_exact"] = bool(all(ok))
    log("SSSP leg: %s" % json.dumps(leg))
    legs["sssp"] = leg
    Ss.free()
    Gs.free()
    del rps, cols, ws
    if not args.no_cpu:
        # the reference's own superstep (delta-push PageRank) beside the pull
        # baseline, on a bounded sample: RMAT-22, C=1 M=4, 2 supersteps
        rp22, col22 = ctx.generate_rmat(min(scale, 22), args.ef, seed=1)
        dp = oracle_check.cpu_baseline(rp22, col22, args.compute, args.parts, nt, api.PR)
        dp.update(cores=nt)
        extra["cpu_baseline_delta_push"] = dp
        log("cpu baseline (delta push): %s" % json.dumps(dp))
    extra["legs"] = legs
    wl = ("PageRank pull (FP32 Kahan, FP64 dangling+residual, d=0.85, tol 1e-6), " + layout_desc(args, scale, False)
          + "; legs: CC and weighted SSSP on RMAT-%d symmetrized, learned vs plain schedule" % scale)
    config = {"workload": wl, "n": n, "arcs": m, "proxies": proxies, "iterations_per_step": pr_iters,
              "exchange_bytes_per_step": bytes_x,
              "gteps_definition": "PR: arcs x iterations / time of one solve (init included)"}
    return line(args, gteps, "GTEPS", ms_step, "f32", config, roof, cpu, extra)


def run_learned(api, ctx, args):
    scale = args.scale
    sp = schedule_speedup(api, ctx, scale, args.parts, ef=args.ef, weighted=True, reps=max(1, args.steps),
                          sym_cc=False)
    ratios = [sp[k]["learned_vs_plain"] for k in ("pr", "cc", "sssp")]
    geo = math.exp(sum(math.log(r) for r in ratios) / len(ratios))
    ms = sum(sp[k]["ms"]["learned"] for k in ("pr", "cc", "sssp"))
    config = {"workload": "learned vs heuristic vs plain schedule: reference PageRank, CC, weighted SSSP from "
                          "gid 0 on RMAT-%d EF%d seed 1 (weights [1,255] seed 2), 1 compute + %d memory "
                          "partitions; best of %d solves per schedule" % (scale, args.ef, args.parts,
                                                                         max(1, args.steps)),
              "arcs": sp["pr"]["arcs"]}
    extra = {"learned_vs_plain": sp}
    cpu = {"value": None, "unit": "x", "cores": 0, "kind": "port",
           "sample": "none: the value is a GPU-to-GPU schedule ratio"}
    return line(args, geo, "x (learned / plain, geometric mean of PR, CC, SSSP)", ms, "f32+u32", config, None,
                cpu, extra)


def run_selector(api, ctx, args):
    """What a per-superstep schedule selector can gain over the heuristic.
    Held-out seeded graphs (RMAT-22/24 EF16 seed 77, Erdos-Renyi n 2^22
    degree 16 seed 78), C=1 M=4: CC (symmetrized), SSSP hop count and
    weighted delta-stepping from the highest-degree vertex.  Their supersteps
    are deterministic, so every fixed option visits the same states: with
    each option timed at every superstep, the sum over supersteps of the
    fastest option's time is the best any per-superstep selector could reach
    (hindsight best).  heuristic / best bounds what the learned selector can
    gain over the heuristic; learned / heuristic is what it does gain.  All
    times are per-superstep device times of the same runs (time_iters)."""
    import numpy as np

    w = np.fromfile(os.path.join(ROOT, "grudon_amd", "selector_bf16.bin"), dtype="<u2")
    ctx.selector_load(w, api.SELECTOR_MARGIN)
    options = [api.OPT_PUSH_SPARSE, api.OPT_PULL_SPARSE, api.OPT_PUSH_DENSE, api.OPT_PULL_DENSE, api.OPT_SHIP_EDGES]
    graphs = [("rmat22", 22), ("rmat24", 24), ("er22", 0)]
    rows = []
    for gname, scale in graphs:
        for sym, algo, aname, kw in [(True, api.CC, "cc", {}), (False, api.SSSP, "sssp_hop", {}),
                                     (False, api.SSSP, "sssp_weighted", {"weighted": True, "delta": args.delta})]:
            if scale:
                G = api.Graph.rmat(ctx, scale, 16, seed=77, symmetrize=sym, wseed=2 if kw.get("weighted") else 0,
                                   C=1, M=4)
            else:
                G = api.Graph.er(ctx, 1 << 22, 16.0, seed=78, wseed=2 if kw.get("weighted") else 0, C=1, M=4)
            if algo == api.SSSP:
                kw = dict(kw, source=int(np.argmax(G.out_degree())))
            S = api.State(G, algo)

            def traced(**ikw):
                S.init(max_iter=100000, time_iters=True, **kw, **ikw)
                S.run(trace_cap=1 << 16)  # warm-up
                S.init(max_iter=100000, time_iters=True, **kw, **ikw)
                st, tr = S.run(trace_cap=1 << 16)
                return st, np.array([r.ms for r in tr]), S.result()

            per_opt, ref = [], None
            for o in options:
                st, t, res = traced(sched=api.SCHED_FIXED, fixed_option=o)
                ref = res if ref is None else ref
                assert np.array_equal(res, ref), "a fixed option changed the result"
                per_opt.append(t)
            n = min(len(t) for t in per_opt)
            best = float(np.min(np.stack([t[:n] for t in per_opt]), axis=0).sum())
            st_h, t_h, res_h = traced(sched=api.SCHED_HEURISTIC)
            st_l, t_l, res_l = traced(sched=api.SCHED_LEARNED)
            assert np.array_equal(res_h, ref) and np.array_equal(res_l, ref)
            row = {"graph": gname, "algo": aname, "supersteps": int(n), "arcs": G.m,
                   "ms": {"best": round(best, 3), "heuristic": round(float(t_h.sum()), 3),
                          "learned": round(float(t_l.sum()), 3),
                          "fixed": [round(float(t.sum()), 3) for t in per_opt]},
                   "heuristic_over_best": round(float(t_h.sum()) / best, 3),
                   "learned_vs_heuristic": round(float(t_h.sum()) / float(t_l.sum()), 3)}
            log("selector %s" % json.dumps(row))
            rows.append(row)
            S.free()
            G.free()
    geo = lambda xs: math.exp(sum(math.log(x) for x in xs) / len(xs))
    ceiling = geo([r["heuristic_over_best"] for r in rows])
    gain = geo([r["learned_vs_heuristic"] for r in rows])
    config = {"workload": "per-superstep schedule selection: CC, SSSP hop and weighted delta-stepping (delta %d) on "
                          "held-out RMAT-22/24 EF16 seed 77 and Erdos-Renyi 2^22 degree 16 seed 78, 1 compute + 4 "
                          "memory partitions" % args.delta,
              "value_definition": "geometric mean of heuristic time / hindsight-best time: the largest gain any "
                                  "per-superstep selector could have over the heuristic"}
    cpu = {"value": None, "unit": "x", "cores": 0, "kind": "port",
           "sample": "none: the value is a GPU-to-GPU schedule ratio"}
    ms = sum(r["ms"]["learned"] for r in rows)
    return line(args, ceiling, "x (heuristic / hindsight best, geometric mean)", ms, "u32", config, None, cpu,
                {"selector": rows, "learned_vs_heuristic_geomean": round(gain, 4)})


def run_schedules(api, ctx, args):
    """The reference's two supersteps and its offload vote on RMAT-`scale`
    (C compute + M memory partitions): plain (the offloaded push every
    superstep, as the reference runs), the NDPEngine vote, edge shipping
    every superstep, and the heuristic; CC on the symmetrized graph, hop and
    weighted delta-stepping SSSP from the highest-degree vertex.  Best of
    `steps` solves each; results must be identical across schedules."""
    import numpy as np

    scheds = [("plain", api.SCHED_PLAIN, 0), ("ndp_engine", api.SCHED_NDP_ENGINE, 0),
              ("shipping", api.SCHED_FIXED, api.OPT_SHIP_EDGES), ("heuristic", api.SCHED_HEURISTIC, 0)]
    out = {}
    for name, sym, algo, kw in [("cc", True, api.CC, {}), ("sssp_hop", False, api.SSSP, {}),
                                ("sssp_weighted", False, api.SSSP, {"weighted": True, "delta": args.delta})]:
        G = api.Graph.rmat(ctx, args.scale, args.ef, seed=1, symmetrize=sym, wseed=2 if kw.get("weighted") else 0,
                           C=args.compute, M=args.parts)
        if algo == api.SSSP:
            kw = dict(kw, source=int(np.argmax(G.out_degree())))
        row = {"arcs": G.m, "ms": {}, "options": {}, "bytes_exchanged": {}}
        ref = None
        same = True
        for label, sc, fo in scheds:
            best = None
            for _ in range(max(1, args.warmup) + max(1, args.steps)):
                S = api.State(G, algo)
                S.init(sched=sc, fixed_option=fo, max_iter=100000, **kw)
                st = S.run()
                r = S.result()
                S.free()
                best = st["ms"] if best is None else min(best, st["ms"])
            ref = r if ref is None else ref
            same = same and bool(np.array_equal(r, ref))
            row["ms"][label] = round(best, 3)
            row["options"][label] = st["option_counts"]
            row["bytes_exchanged"][label] = st["bytes_exchanged"]
        row["shipping_vs_plain"] = round(row["ms"]["plain"] / row["ms"]["shipping"], 3)
        row["ndp_engine_vs_plain"] = round(row["ms"]["plain"] / row["ms"]["ndp_engine"], 3)
        row["results_identical"] = same
        log("schedules %s: %s" % (name, json.dumps(row)))
        out[name] = row
        G.free()
    ratios = [out[k]["shipping_vs_plain"] for k in out]
    geo = math.exp(sum(math.log(r) for r in ratios) / len(ratios))
    config = {"workload": "offloaded push vs NDPEngine vote vs edge shipping vs heuristic: CC (symmetrized), SSSP "
                          "hop and weighted (delta %d) from the highest-degree vertex, %s"
                          % (args.delta, layout_desc(args, args.scale, False)),
              "schedules": [s[0] for s in scheds]}
    cpu = {"value": None, "unit": "x", "cores": 0, "kind": "port",
           "sample": "none: the value is a GPU-to-GPU schedule ratio"}
    ms = sum(out[k]["ms"]["shipping"] for k in out)
    return line(args, geo, "x (edge shipping / plain, geometric mean of CC, SSSP hop, SSSP weighted)", ms, "u32",
                config, None, cpu, {"schedules": out})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=0, help="RMAT scale (default 26 for headline/learned, else 24)")
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--parts", type=int, default=4)
    ap.add_argument("--compute", type=int, default=1)
    ap.add_argument("--delta", type=int, default=256, help="SSSP delta-stepping bucket width")
    ap.add_argument("--offload-only", action="store_true",
                    help="cc/sssp/pr: the heuristic never runs a push superstep as edge shipping")
    ap.add_argument("--workload", default="headline", choices=["headline", "pr_pull", "cc", "sssp", "pr", "batch",
                                                                   "learned", "schedules", "selector"])
    ap.add_argument("--pull", default="hotcold", choices=["plain", "tiled", "segmented", "hotcold"],
                    help="pull PageRank: the row-block pull, LDS tiles, L2 segments or the hot / cold split "
                         "(default)")
    ap.add_argument("--tier-bits", type=int, default=0,
                    help="tiled / segmented / hot-cold pull: 2^bits sources per tier or window (0 = 15 / 19 / 15)")
    ap.add_argument("--tile-arcs", type=int, default=0, help="tiled pull: arcs per tile (0 = 65536)")
    ap.add_argument("--pull-every-row", action="store_true",
                    help="cc / sssp workloads: pull every row (by default a pull leaves out the rows that "
                         "cannot improve)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 / exactness check")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--speedup-scale", type=int, default=-1,
                    help="pr_pull: RMAT scale of a learned-vs-plain leg (0 = skip, default); headline: 0 skips "
                         "its learned-vs-plain leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a cold device copy of MIB MiB (PMC counter calibration)")
    ap.add_argument("--calibrate-stream", type=int, default=0, metavar="MIB",
                    help="only run the probe library's streaming read and streaming write of MIB MiB, "
                         "one launch each (PMC counter calibration with kernels of known bytes)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.calibrate_stream:
        from grudon_amd import api as _api
        nb = args.calibrate_stream << 20
        log("calibration stream: read %d MiB: %.3f ms" % (args.calibrate_stream, _api.probe_stream(0, nb, 1)))
        log("calibration stream: write %d MiB: %.3f ms" % (args.calibrate_stream, _api.probe_stream(0, nb, 1, True)))
        return
    if not args.scale:
        args.scale = 26 if args.workload in ("headline", "learned") else 24
    if args.speedup_scale < 0:
        args.speedup_scale = args.scale if args.workload == "headline" else 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU (partitions are resident in one HBM); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from grudon_amd import api

    ctx = api.Context(0)
    ctx.set_pull_tiling({"plain": 0, "tiled": 1, "segmented": 2, "hotcold": 3}[args.pull], args.tier_bits,
                        args.tile_arcs)
    run = {"headline": run_headline, "pr_pull": run_pagerank, "pr": run_pagerank, "cc": run_cc, "sssp": run_sssp, "batch": run_batch,
           "learned": run_learned, "schedules": run_schedules, "selector": run_selector}[args.workload]
    out = run(api, ctx, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
