"""Oracle legs of bench.py: the FP64 check and the timed CPU baseline.

The CPU baseline times the oracle's restatement of the reference's BSP
superstep (the reference itself cannot be built: its Galois fork, Boost,
cxxopts and spdlog are absent) on a bounded sample of the benchmark graph.
"""
import time

import numpy as np

import oracle_lib as O


def pr_pull_l1(rp, col, gpu_ranks, iterations, nthreads=0):
    """L1 distance of the GPU pull PageRank to the FP64 power iteration run
    for the same number of iterations."""
    import os

    g = O.Graph(rp, col)
    nt = nthreads or min(len(os.sched_getaffinity(0)), 64)
    ref, _, _ = O.pr_pull(g, iters=int(iterations), nthreads=nt)
    return float(np.abs(np.asarray(gpu_ranks, dtype=np.float64) - ref).sum())


def cpu_baseline(rp, col, C, M, nthreads, algo, supersteps=2):
    """Time `supersteps` BSP supersteps of the reference-semantics operator on
    the host; GTEPS = edges examined by gen_updates / time."""
    g = O.Graph(rp, col)
    t0 = time.perf_counter()
    P = O.Partition(g, C, M)
    t_build = time.perf_counter() - t0
    t1 = time.perf_counter()
    if algo == 1:
        _, st = P.cc(max_iter=supersteps, nthreads=nthreads)
        name = "CC"
    elif algo == 2:
        _, st = P.sssp(0, max_iter=supersteps, nthreads=nthreads)
        name = "SSSP hop"
    else:
        _, st = P.pr(fp64=False, max_iter=supersteps, nthreads=nthreads)
        name = "reference delta-push PageRank (FP32)"
    dt = time.perf_counter() - t1
    edges = int(st[2])
    return {"value": round(edges / dt / 1e9, 4), "unit": "GTEPS", "kind": "port",
            "sample": "first %d BSP supersteps of the oracle's restatement of the reference %s "
                      "(C=%d, M=%d, same graph, partitions in-process), %d edges examined in %.2f s; "
                      "layout build %.1f s excluded" % (supersteps, name, C, M, edges, dt, t_build)}
