"""GPU parity at the benchmark configurations' own sizes (where the oracle
finishes in seconds) and for the learned schedule.

- RMAT-18 EF16 seed 1, C = 1, M = 1: the reference's delta-push PageRank
  (its CPU-runnable configuration, mpirun -n 2);
- RMAT-18 EF16 seed 1 symmetrized, C = 1, M = 4: CC and hop SSSP bit-exact
  with equal iteration counts; weighted delta-stepping (weights [1,255] seed
  2, delta 256) from the 16 sources drawn with seed 3 among the vertices of
  degree >= 1, distances bit-exact against Dijkstra;
- RMAT-18 C = 1, M = 4 pull PageRank: within 1e-6 L1 of the FP64 power
  iteration, stopping within one iteration of the oracle's own convergence;
- the 64-graph batch generator at reduced scales through grd_batch_*;
- the learned schedule (packaged weights) on reference PageRank against the
  oracle.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from grudon_amd import api

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NT = 8  # oracle threads


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def rmat18():
    return O.rmat(18, 16, seed=1)


@pytest.fixture(scope="module")
def rmat18_sym():
    return O.rmat(18, 16, seed=1, symmetrize=True)


def _pr_bound(out, ref64, ref32):
    out = out.astype(np.float64)
    l1 = np.abs(out / out.sum() - ref64 / ref64.sum()).sum()
    l1_32 = np.abs(ref32 / ref32.sum() - ref64 / ref64.sum()).sum()
    return l1, max(1e-6, 4 * l1_32)


@pytest.mark.parametrize("sched", [api.SCHED_PLAIN, api.SCHED_HEURISTIC])
def test_rmat18_reference_pagerank(ctx, rmat18, sched):
    _say("C1 reference PR rmat18", sched)
    P = O.Partition(rmat18, 1, 1)
    ref64, st = P.pr(fp64=True, nthreads=NT)
    ref32, _ = P.pr(fp64=False, nthreads=NT)
    G = api.Graph.from_csr(ctx, rmat18.rp, rmat18.col, C=1, M=1)
    S = api.State(G, api.PR)
    S.init(sched=sched)
    stats = S.run()
    l1, bound = _pr_bound(S.result(), ref64, ref32)
    _say("   L1", l1, "bound", bound, "iters", stats["iterations"], int(st[0]))
    assert l1 <= bound
    assert abs(stats["iterations"] - int(st[0])) <= 3
    S.free()
    G.free()


@pytest.mark.parametrize("sched", [api.SCHED_PLAIN, api.SCHED_HEURISTIC, api.SCHED_NDP_ENGINE])
def test_rmat18_sym_cc(ctx, rmat18_sym, sched):
    _say("C3 cc rmat18 sym", sched)
    ref, st = O.Partition(rmat18_sym, 1, 4).cc(nthreads=NT)
    assert np.array_equal(ref, O.cc_unionfind(rmat18_sym))
    G = api.Graph.from_csr(ctx, rmat18_sym.rp, rmat18_sym.col, C=1, M=4)
    S = api.State(G, api.CC)
    S.init(sched=sched)
    stats = S.run()
    assert np.array_equal(S.result(), ref)
    assert stats["iterations"] == int(st[0])
    S.free()
    G.free()


def test_rmat18_sym_sssp_hop(ctx, rmat18_sym):
    _say("C3 sssp hop rmat18 sym")
    ref, st = O.Partition(rmat18_sym, 1, 4).sssp(0, nthreads=NT)
    assert np.array_equal(ref, O.bfs(rmat18_sym, 0))
    G = api.Graph.from_csr(ctx, rmat18_sym.rp, rmat18_sym.col, C=1, M=4)
    S = api.State(G, api.SSSP)
    for sched in (api.SCHED_PLAIN, api.SCHED_HEURISTIC):
        S.init(source=0, sched=sched)
        stats = S.run()
        assert np.array_equal(S.result(), ref), sched
        assert stats["iterations"] == int(st[0]), sched
    S.free()
    G.free()


def test_rmat18_sym_sssp_weighted_16_sources(ctx, rmat18_sym):
    """The benchmark's 16-source draw (seed 3 among degree >= 1), weights seed
    2, delta 256: every distance vector equals Dijkstra's; for two sources
    the iteration count equals the oracle's delta-stepping BSP."""
    _say("C3 sssp weighted 16 sources")
    gw = rmat18_sym.with_weights(2)
    deg = gw.out_degree()
    rng = np.random.default_rng(3)
    sources = rng.choice(np.flatnonzero(deg > 0), size=16, replace=False)
    G = api.Graph.from_csr(ctx, gw.rp, gw.col, gw.w, C=1, M=4)
    S = api.State(G, api.SSSP)
    P = O.Partition(gw, 1, 4)
    for i, s in enumerate(sources.tolist()):
        S.init(source=s, weighted=True, delta=256, max_iter=100000, sched=api.SCHED_HEURISTIC)
        stats = S.run()
        assert np.array_equal(S.result(), O.dijkstra(gw, s)), s
        if i < 2:
            _, st = P.sssp(s, weighted=True, delta=256, nthreads=NT)
            assert stats["iterations"] == int(st[0]), s
    S.free()
    G.free()


def test_rmat18_pr_pull(ctx, rmat18):
    _say("C2-shaped pull PR rmat18 M=4")
    G = api.Graph.from_csr(ctx, rmat18.rp, rmat18.col, C=1, M=4)
    S = api.State(G, api.PR_PULL)
    S.init(tolerance=1e-6)
    stats = S.run()
    ref, _, _ = O.pr_pull(rmat18, iters=stats["iterations"], nthreads=NT)
    l1 = np.abs(S.result().astype(np.float64) - ref).sum()
    _, it_conv, _ = O.pr_pull(rmat18, iters=0, tol=1e-6, nthreads=NT)
    _say("   L1", l1, "iters", stats["iterations"], "oracle converges at", it_conv)
    assert l1 <= 1e-6
    assert abs(stats["iterations"] - it_conv) <= 1
    S.free()
    G.free()


def _batch_graphs(seed=4, rmat_scales=(10, 14), er_log2=(10, 14)):
    """The benchmark's 64-graph mix (32 RMAT with EF 8-32, 32 Erdos-Renyi
    with mean degree 4-64, drawn from the seed) at reduced scales."""
    rng = np.random.default_rng(seed)
    gs = []
    for i in range(32):
        s, ef = int(rng.integers(rmat_scales[0], rmat_scales[1] + 1)), int(rng.integers(8, 33))
        gs.append(O.rmat(s, ef, seed=1000 + i))
    for i in range(32):
        n = int(round(2.0 ** rng.uniform(*er_log2)))
        d = float(rng.uniform(4, 64))
        gs.append(O.er(n, d, seed=2000 + i))
    return gs


def test_batch_64_graphs_reduced(ctx):
    _say("C4 batch of 64 at reduced scales")
    gs = _batch_graphs()
    B = api.Batch(ctx, [(g.rp, g.col) for g in gs])
    chosen = B.autotune()
    _say("   autotuned", chosen)
    try:
        it_pr, _ = B.pagerank(tolerance=1e-6, max_iter=1000)
        # the batch keeps the results of the algorithm it ran last
        pr = [B.result(i, api.PR_PULL) for i in range(len(gs))]
        it_cc, _ = B.cc(max_iter=1000)
        for i, g in enumerate(gs):
            ref, _, _ = O.pr_pull(g, iters=int(it_pr[i]))
            l1 = np.abs(pr[i].astype(np.float64) - ref).sum()
            assert l1 <= 1e-6, (i, l1)
            _, it_conv, _ = O.pr_pull(g, iters=0, tol=1e-6)
            assert abs(int(it_pr[i]) - it_conv) <= 1, i
            lab, st = O.Partition(g, 1, 1).cc()
            assert np.array_equal(B.result(i, api.CC), lab), i
            assert int(it_cc[i]) == int(st[0]), i
    finally:
        B.free()


@pytest.mark.parametrize("C,M", [(1, 4), (2, 3)])
def test_learned_schedule_reference_pagerank(ctx, C, M):
    """Reference PageRank under the learned schedule (packaged weights, the
    heuristic as fallback) against the oracle, as for the fixed schedules."""
    _say("learned PR", C, M)
    w = np.fromfile(os.path.join(ROOT, "grudon_amd", "selector_bf16.bin"), dtype="<u2")
    ctx.selector_load(w, min_margin=0.0)
    g = O.rmat(15, 16, seed=9)
    P = O.Partition(g, C, M)
    ref64, st = P.pr(fp64=True, nthreads=NT)
    ref32, _ = P.pr(fp64=False, nthreads=NT)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=C, M=M)
    S = api.State(G, api.PR)
    S.init(sched=api.SCHED_LEARNED)
    stats = S.run()
    l1, bound = _pr_bound(S.result(), ref64, ref32)
    diag = "L1 %.3e bound %.3e iterations %d oracle %d options %s fallbacks %d" % (
        l1, bound, stats["iterations"], int(st[0]), stats["option_counts"], stats["fallbacks"])
    _say("   " + diag)
    assert l1 <= bound, diag
    assert abs(stats["iterations"] - int(st[0])) <= 3, diag
    S.free()
    G.free()
