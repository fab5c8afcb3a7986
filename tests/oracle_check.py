"""The cpu_baseline legs of bench.py: the oracle timed on the host cores.

The reference itself cannot be built (its Galois fork, Boost, cxxopts and
spdlog are absent), so the CPU baseline times the oracle: the threaded FP64
pull PageRank (the headline's algorithm) and the restatement of the
reference's BSP superstep (delta-push PageRank, CC, SSSP).  Every figure is
the median of `reps` timed runs after one untimed warm-up run, on a stated,
bounded sample of the benchmark's own graph.
"""
import statistics
import time

import numpy as np

import oracle_lib as O


def _median_runs(fn, reps):
    fn()  # warm-up
    return statistics.median(fn() for _ in range(reps))


def cpu_baseline_pull(plan, nthreads, iters=2, reps=5):
    """`iters` iterations of the threaded FP64 pull PageRank over the same
    graph as the GPU run (the transposed graph built once, excluded); GTEPS =
    arcs x iterations / median time of the iteration loop."""
    def one():
        _, it, _, sec = plan.run(iters=iters)
        assert it == iters
        return sec

    sec = _median_runs(one, reps)
    return {"value": round(plan.arcs * iters / sec / 1e9, 4), "unit": "GTEPS", "kind": "port",
            "sample": "the oracle's threaded FP64 pull PageRank (the headline's recurrence) over the same "
                      "%d-arc graph, %d iterations per run, median of %d runs after 1 warm-up: %.3f s; the "
                      "transposed graph is built once outside the timed region" % (plan.arcs, iters, reps, sec)}


def cpu_baseline(rp, col, C, M, nthreads, algo, supersteps=2, w=None, source=0, reps=5):
    """`supersteps` BSP supersteps of the oracle's restatement of the
    reference superstep (partitions in-process); GTEPS = edges examined by
    gen_updates / median time (1 warm-up, `reps` timed runs)."""
    g = O.Graph(rp, col, w)
    t0 = time.perf_counter()
    P = O.Partition(g, C, M)
    t_build = time.perf_counter() - t0
    edges = [0]

    def one():
        t1 = time.perf_counter()
        if algo == 1:
            _, st = P.cc(max_iter=supersteps, nthreads=nthreads)
        elif algo == 2:
            _, st = P.sssp(source, weighted=w is not None, max_iter=supersteps, nthreads=nthreads)
        else:
            _, st = P.pr(fp64=False, max_iter=supersteps, nthreads=nthreads)
        dt = time.perf_counter() - t1
        edges[0] = int(st[2])
        return dt

    dt = _median_runs(one, reps)
    name = {1: "CC", 2: "SSSP (weighted)" if w is not None else "SSSP hop"}.get(algo, "delta-push PageRank (FP32)")
    return {"value": round(edges[0] / dt / 1e9, 4), "unit": "GTEPS", "kind": "port",
            "sample": "first %d BSP supersteps of the oracle's restatement of the reference %s "
                      "(C=%d, M=%d, %d arcs, partitions in-process), %d edges examined, median of %d runs after "
                      "1 warm-up: %.3f s; layout build %.1f s excluded"
                      % (supersteps, name, C, M, len(col), edges[0], reps, dt, t_build)}


def cpu_baseline_batch(graphs, nthreads, budget_s=15.0):
    """The oracle's threaded FP64 pull PageRank (the recurrence the batched
    form computes) over the batch's graphs, in order, until about budget_s
    seconds of wall time; GTEPS = arcs x iterations / time."""
    done, arcs_iters, t = 0, 0, 0.0
    for rp, col in graphs:
        g = O.Graph(rp, col)
        t0 = time.perf_counter()
        _, it, _ = O.pr_pull(g, iters=0, tol=1e-6, nthreads=nthreads)
        t += time.perf_counter() - t0
        arcs_iters += g.m * it
        done += 1
        if t >= budget_s:
            break
    return {"value": round(arcs_iters / t / 1e9, 4), "unit": "GTEPS", "kind": "port",
            "sample": "the oracle's FP64 pull PageRank (tol 1e-6) on the first %d of %d batch graphs, "
                      "%d arc-iterations in %.2f s" % (done, len(graphs), arcs_iters, t)}


def cc_exact(rp, col, labels, nthreads):
    """Labels equal the minimum gid of each component (parallel union-find)."""
    return bool(np.array_equal(labels, O.cc_unionfind_mt(rp, col, nthreads)))


def sssp_exact(rp, col, w, dist, src, nthreads):
    """dist satisfies the Bellman equations from src (exact distances)."""
    return O.sssp_violations(rp, col, w, dist, src, nthreads) == 0
