"""The cpu_baseline leg of bench.py: the oracle timed on the host cores.

The CPU baseline times the oracle's restatement of the reference's BSP
superstep (the reference itself cannot be built: its Galois fork, Boost,
cxxopts and spdlog are absent) on a bounded sample of the benchmark graph.
"""
import time

import numpy as np

import oracle_lib as O


def cpu_baseline(rp, col, C, M, nthreads, algo, supersteps=2, w=None, source=0):
    """Time `supersteps` BSP supersteps of the reference-semantics operator on
    the host; GTEPS = edges examined by gen_updates / time."""
    g = O.Graph(rp, col, w)
    t0 = time.perf_counter()
    P = O.Partition(g, C, M)
    t_build = time.perf_counter() - t0
    t1 = time.perf_counter()
    if algo == 1:
        _, st = P.cc(max_iter=supersteps, nthreads=nthreads)
        name = "CC"
    elif algo == 2:
        _, st = P.sssp(source, weighted=w is not None, max_iter=supersteps, nthreads=nthreads)
        name = "SSSP (weighted)" if w is not None else "SSSP hop"
    else:
        _, st = P.pr(fp64=False, max_iter=supersteps, nthreads=nthreads)
        name = "delta-push PageRank (FP32)"
    dt = time.perf_counter() - t1
    edges = int(st[2])
    return {"value": round(edges / dt / 1e9, 4), "unit": "GTEPS", "kind": "port",
            "sample": "first %d BSP supersteps of the oracle's restatement of the reference %s "
                      "(C=%d, M=%d, same graph, partitions in-process), %d edges examined in %.2f s; "
                      "layout build %.1f s excluded" % (supersteps, name, C, M, edges, dt, t_build)}


def cpu_baseline_pull(graphs, nthreads, budget_s=15.0):
    """Time the oracle's threaded FP64 pull PageRank (the recurrence the
    batched form computes) over the batch's graphs, in order, until about
    budget_s seconds of wall time; GTEPS = arcs x iterations / time."""
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
