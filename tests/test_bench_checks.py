"""The full-size checks of bench.py (numpy / scipy / the threaded oracle)
against the oracle, and planted mistakes that they must catch (no GPU)."""
import numpy as np

import bench
import oracle_lib as O


def test_fp64_pagerank_matches_oracle():
    g = O.rmat(11, 8, seed=21)
    for iters in (1, 7, 30):
        ref, _, _ = O.pr_pull(g, iters=iters)
        assert np.abs(bench.fp64_pagerank(g.rp, g.col, iters) - ref).sum() < 1e-12


def test_cc_labels_scipy_matches_oracle():
    g = O.rmat(11, 4, seed=22, symmetrize=True)
    ref, _ = O.Partition(g, 1, 2).cc()
    lab = bench.cc_labels_scipy(g.rp, g.col)
    assert np.array_equal(lab, ref)
    assert np.array_equal(lab, O.cc_unionfind(g))


def test_bellman_check_accepts_dijkstra_and_rejects_mistakes():
    """The check bench.py runs on SSSP distances (oracle_check.sssp_exact)."""
    import oracle_check

    g = O.rmat(11, 8, seed=23, symmetrize=True).with_weights(2)
    src = int(np.argmax(g.out_degree()))
    d = O.dijkstra(g, src)
    assert oracle_check.sssp_exact(g.rp, g.col, g.w, d, src, 4)
    # a distance one too large, one too small, and a reached vertex marked unreachable
    reached = np.flatnonzero((d != bench.INF) & (np.arange(g.n) != src))
    for delta in (+1, -1):
        bad = d.astype(np.int64)
        bad[reached[5]] += delta
        assert not oracle_check.sssp_exact(g.rp, g.col, g.w, bad.astype(np.uint32), src, 4)
    bad = d.copy()
    bad[reached[-1]] = bench.INF
    assert not oracle_check.sssp_exact(g.rp, g.col, g.w, bad, src, 4)
    # hop distances (unit weights) of the oracle's BSP pass the same check
    h, _ = O.Partition(g, 1, 3).sssp(src)
    assert oracle_check.sssp_exact(g.rp, g.col, np.ones_like(g.w), h, src, 4)


def test_bench_refuses_several_gpus():
    import subprocess
    import sys

    p = subprocess.run([sys.executable, bench.__file__, "--gpus", "2"], capture_output=True, text=True)
    assert p.returncode == 2 and "exactly one GPU" in p.stderr
