"""GPU tests of the source-blocked pull-PageRank modes (tiers.hip: LDS tiles,
L2 segments, the hot / cold split) against the FP64 oracle.

Small tiers and tiles make small graphs exercise what large graphs do: many
tiers per partition, rows whose arcs span several tiers (partials merged in
tier order), pairs longer than one row block (segments merged by the
fix-up), windows cut short at a partition's end, C > 1 layouts with cross
regions inside the partition range.  Bar: L1 <= 1e-6 against the FP64 power
iteration for the same iteration count, stopping within one iteration of the
oracle's convergence, bit-identical repeat runs.
"""
import numpy as np
import pytest

import oracle_lib as O
from grudon_amd import api

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _solve(ctx, g, C, M, mode, tier_bits=0, tile_arcs=0, tol=1e-6):
    ctx.set_pull_tiling(mode, tier_bits, tile_arcs)
    try:
        G = api.Graph.from_csr(ctx, g.rp, g.col, C=C, M=M)
        S = api.State(G, api.PR_PULL)
        S.init(tolerance=tol)
        st = S.run()
        r = S.result()
        S.free()
        G.free()
    finally:
        ctx.set_pull_tiling(3, 0, 0)  # the default
    return r, st


@pytest.mark.parametrize("mode,scale,C,M,tier_bits,tile_arcs", [
    (1, 12, 1, 1, 4, 0), (1, 12, 1, 4, 6, 4096), (1, 13, 2, 3, 5, 2048), (1, 14, 1, 2, 8, 0),
    (1, 14, 3, 2, 10, 8192), (1, 16, 1, 1, 15, 0), (1, 16, 1, 4, 12, 0),
    (2, 12, 1, 1, 4, 0), (2, 13, 2, 3, 6, 0), (2, 14, 1, 4, 8, 0), (2, 16, 1, 1, 10, 0), (2, 16, 1, 4, 0, 0),
    (3, 12, 1, 1, 4, 0), (3, 13, 2, 3, 6, 0), (3, 14, 1, 4, 8, 0), (3, 16, 1, 1, 10, 0), (3, 16, 1, 4, 0, 0),
    (3, 17, 3, 2, 12, 0),
])
def test_tiled_pull_fp64_bound(ctx, mode, scale, C, M, tier_bits, tile_arcs):
    _say("tiled pull", mode, scale, C, M, tier_bits, tile_arcs)
    g = O.rmat(scale, 16, seed=scale + 7)
    r, st = _solve(ctx, g, C, M, mode, tier_bits, tile_arcs)
    ref, _, _ = O.pr_pull(g, iters=st["iterations"], nthreads=4)
    l1 = float(np.abs(r.astype(np.float64) - ref).sum())
    _, it_conv, _ = O.pr_pull(g, iters=0, tol=1e-6, nthreads=4)
    _say("   L1", l1, "iters", st["iterations"], it_conv)
    assert l1 <= 1e-6
    assert abs(st["iterations"] - it_conv) <= 1


def test_tiled_equals_plain_pull_and_repeats_bit_exact(ctx):
    _say("tiled vs plain")
    g = O.rmat(15, 16, seed=3)
    rp, st_p = _solve(ctx, g, 1, 4, 0)
    for mode, tb, ta in ((1, 9, 4096), (2, 11, 4096), (3, 8, 0)):
        rt, st_t = _solve(ctx, g, 1, 4, mode, tb, ta)
        rt2, _ = _solve(ctx, g, 1, 4, mode, tb, ta)
        assert np.array_equal(rt.view(np.uint32), rt2.view(np.uint32)), mode  # deterministic
        assert st_t["iterations"] == st_p["iterations"], mode
        assert float(np.abs(rt.astype(np.float64) - rp.astype(np.float64)).sum()) <= 1e-6, mode


def test_tiled_pull_er_and_sink_heavy(ctx):
    """An Erdos-Renyi graph (flat degrees: every tier has arcs) and a
    sink-heavy directed graph with isolated vertices."""
    _say("tiled er / sinks")
    for g in (O.er(20000, 12.0, seed=5), O.from_edges(9, [(0, 1), (1, 2), (2, 0), (3, 4), (5, 4), (6, 4)])):
        for M in (1, 3):
            for mode in (1, 2, 3):
                r, st = _solve(ctx, g, 1, M, mode, 4, 0 if mode == 3 else 1024)
                ref, _, _ = O.pr_pull(g, iters=st["iterations"])
                assert float(np.abs(r.astype(np.float64) - ref).sum()) <= 1e-6, (g.n, M, mode)


def test_hot_cold_long_hot_and_cold_rows(ctx):
    """Rows longer than a row block on both sides of the split: every vertex
    points at 0 and 1, so with a 2^12-source window rows 0 and 1 each have
    about 4096 hot arcs (two hot segments) and 4900 cold ones (three cold
    segments); plus a ring so that the other rows are not empty."""
    _say("hot/cold long rows")
    n = 9000
    edges = [(i, 0) for i in range(1, n)] + [(i, 1) for i in range(2, n)] + [(i, (i + 1) % n) for i in range(n)]
    g = O.from_edges(n, edges)
    for M in (1, 2):
        r, st = _solve(ctx, g, 1, M, 3, 12, 0)
        ref, _, _ = O.pr_pull(g, iters=st["iterations"])
        assert float(np.abs(r.astype(np.float64) - ref).sum()) <= 1e-6, M
        rp, _ = _solve(ctx, g, 1, M, 0)
        assert float(np.abs(r.astype(np.float64) - rp.astype(np.float64)).sum()) <= 1e-6, M


@pytest.fixture(scope="module")
def bsp_graphs():
    gs = O.rmat(12, 16, seed=5, symmetrize=True)
    return gs, O.rmat(12, 16, seed=7), gs.with_weights(11)


@pytest.mark.parametrize("C,M,tb", [(1, 1, 4), (1, 4, 6), (2, 3, 5), (1, 2, 8)])
@pytest.mark.parametrize("option", [api.OPT_PULL_SPARSE, api.OPT_PULL_DENSE])
def test_hot_cold_bsp_pull(ctx, bsp_graphs, C, M, tb, option):
    """The BSP pull through the hot / cold split with small windows (so that
    rows have hot and cold arcs, and long hot and cold rows exist), every
    superstep a pull: CC labels, hop and weighted SSSP distances bit-exact
    with equal iteration counts, reference PageRank within its FP32 bound."""
    _say("hot/cold bsp pull", C, M, tb, option)
    gs, gd, gw = bsp_graphs
    ctx.set_pull_tiling(3, tb, 0)
    try:
        ref, st = O.Partition(gs, C, M).cc()
        G = api.Graph.from_csr(ctx, gs.rp, gs.col, C=C, M=M)
        S = api.State(G, api.CC)
        S.init(sched=api.SCHED_FIXED, fixed_option=option)
        stats = S.run()
        assert np.array_equal(S.result(), ref)
        assert stats["iterations"] == int(st[0])
        S.free()
        G.free()

        ref, st = O.Partition(gd, C, M).sssp(0)
        G = api.Graph.from_csr(ctx, gd.rp, gd.col, C=C, M=M)
        S = api.State(G, api.SSSP)
        S.init(source=0, sched=api.SCHED_FIXED, fixed_option=option)
        stats = S.run()
        assert np.array_equal(S.result(), ref)
        assert stats["iterations"] == int(st[0])
        S.free()

        Sp = api.State(G, api.PR)
        Sp.init(sched=api.SCHED_FIXED, fixed_option=option)
        stats = Sp.run()
        out = Sp.result().astype(np.float64)
        P = O.Partition(gd, C, M)
        ref64, st = P.pr(fp64=True)
        ref32, _ = P.pr(fp64=False)
        l1 = np.abs(out / out.sum() - ref64 / ref64.sum()).sum()
        l1_32 = np.abs(ref32 / ref32.sum() - ref64 / ref64.sum()).sum()
        assert l1 <= max(1e-6, 4 * l1_32)
        assert abs(stats["iterations"] - int(st[0])) <= 3
        Sp.free()
        G.free()

        src = int(np.argmax(gw.out_degree() > 0))
        _, st = O.Partition(gw, C, M).sssp(src, weighted=True, delta=0)
        G = api.Graph.from_csr(ctx, gw.rp, gw.col, gw.w, C=C, M=M)
        S = api.State(G, api.SSSP)
        S.init(source=src, weighted=True, delta=0, max_iter=100000, sched=api.SCHED_FIXED, fixed_option=option)
        stats = S.run()
        assert np.array_equal(S.result(), O.dijkstra(gw, src))
        assert stats["iterations"] == int(st[0])
        S.free()
        G.free()
    finally:
        ctx.set_pull_tiling(3, 0, 0)


@pytest.mark.parametrize("scale_w,add_w", [(1, 0), (1000, 7), (1, 1)])
@pytest.mark.parametrize("option", [api.OPT_PULL_SPARSE, api.OPT_PULL_DENSE, api.OPT_SHIP_EDGES])
def test_hot_cold_weight_storage(ctx, bsp_graphs, scale_w, add_w, option):
    """The hot / cold split, and the edge-shipping wire, keep arc weights as
    bytes when every weight is below 256 and as 32-bit words otherwise:
    weights in [1, 255] (bytes), in [1007, 255007] and in [2, 256] (one
    weight of 256 is enough for words); weighted SSSP through
    every-superstep pulls or shipped supersteps is Dijkstra-exact with the
    oracle's iteration count either way."""
    _say("hot/cold weight storage", scale_w, add_w, option)
    gw0 = bsp_graphs[2]
    gw = O.Graph(gw0.rp, gw0.col, gw0.w.astype(np.uint64) * scale_w + add_w)
    assert (int(gw.w.max()) < 256) == (scale_w == 1 and add_w == 0)
    src = int(np.argmax(gw.out_degree() > 0))
    _, st = O.Partition(gw, 1, 4).sssp(src, weighted=True, delta=0)
    ctx.set_pull_tiling(3, 6, 0)
    try:
        G = api.Graph.from_csr(ctx, gw.rp, gw.col, gw.w, C=1, M=4)
        S = api.State(G, api.SSSP)
        S.init(source=src, weighted=True, delta=0, max_iter=100000, sched=api.SCHED_FIXED, fixed_option=option)
        stats = S.run()
        assert np.array_equal(S.result(), O.dijkstra(gw, src))
        assert stats["iterations"] == int(st[0])
        S.free()
        G.free()
    finally:
        ctx.set_pull_tiling(3, 0, 0)


@pytest.mark.parametrize("mode,tb", [(0, 0), (3, 6), (3, 9)])
@pytest.mark.parametrize("C,M", [(1, 4), (2, 3)])
@pytest.mark.parametrize("algo", ["cc", "hop", "weighted", "delta"])
def test_pull_leaves_out_done_rows(ctx, bsp_graphs, mode, tb, C, M, algo):
    """A CC / SSSP pull leaves out the row blocks whose rows with arcs cannot
    improve (the vertex's value is at or below the smallest value broadcast):
    every superstep a pull, the labels / distances and iteration counts equal
    the oracle's and a run that pulls every row, with no more arcs examined and
    kernel bytes counted (fewer for CC); the heuristic schedule (pushes, pulls
    and shipping mixed) gives the same results either way."""
    _say("rows done", mode, tb, C, M, algo)
    gs, gd, gw = bsp_graphs
    g = {"cc": gs, "hop": gd, "weighted": gw, "delta": gw}[algo]
    kw = {"cc": {}, "hop": {"source": 0}, "weighted": {"source": 0, "weighted": True},
          "delta": {"source": 0, "weighted": True, "delta": 32}}[algo]
    P = O.Partition(g, C, M)
    if algo == "cc":
        ref, st = P.cc()
    else:
        ref, st = P.sssp(0, weighted=kw.get("weighted", False), delta=kw.get("delta", 0))
    ctx.set_pull_tiling(mode, tb, 0)
    try:
        G = api.Graph.from_csr(ctx, g.rp, g.col, g.w if algo in ("weighted", "delta") else None, C=C, M=M)
        algo_id = api.CC if algo == "cc" else api.SSSP
        runs = {}
        for every in (True, False):
            for sched in (api.SCHED_FIXED, api.SCHED_HEURISTIC):
                S = api.State(G, algo_id)
                S.init(max_iter=100000, sched=sched, fixed_option=api.OPT_PULL_DENSE, time_kernel=True,
                       pull_every_row=every, **kw)
                stats = S.run()
                assert np.array_equal(S.result(), ref), (every, sched)
                assert stats["iterations"] == int(st[0]), (every, sched)
                runs[every, sched] = stats
                S.free()
        full, part = runs[True, api.SCHED_FIXED], runs[False, api.SCHED_FIXED]
        assert full["edges_examined"] == full["iterations"] * G.m
        # CC labels reach the frontier's smallest label within a few supersteps,
        # so whole blocks are left out; how many SSSP rows can be left out
        # depends on the distances (none, on some of these small graphs)
        if algo == "cc":
            assert part["edges_examined"] < full["edges_examined"]
            assert part["kernel_bytes"] < full["kernel_bytes"]
        assert part["edges_examined"] <= full["edges_examined"]
        # the pull that may leave rows out also reads the rows-done bitmap
        assert part["kernel_bytes"] <= full["kernel_bytes"] + part["iterations"] * G.proxies // 8
        _say("   arcs examined", part["edges_examined"], "of", full["edges_examined"])
        G.free()
    finally:
        ctx.set_pull_tiling(3, 0, 0)


@pytest.mark.parametrize("algo", ["cc", "weighted"])
def test_events_only_timing(ctx, bsp_graphs, algo):
    """time_kernel 2 times the same launches as time_kernel 1 without the
    exact count of the gathers that rows left out: the same results,
    iterations and launches, and kernel bytes that count those gathers as
    made (never fewer than the exact count; more for CC, whose pulls always
    leave some out)."""
    _say("events-only timing", algo)
    gs, _, gw = bsp_graphs
    g = gs if algo == "cc" else gw
    kw = {} if algo == "cc" else {"source": 0, "weighted": True}
    ctx.set_pull_tiling(3, 6, 0)
    try:
        G = api.Graph.from_csr(ctx, g.rp, g.col, g.w if algo == "weighted" else None, C=1, M=4)
        runs = {}
        for tk in (1, 2):
            S = api.State(G, api.CC if algo == "cc" else api.SSSP)
            S.init(max_iter=100000, sched=api.SCHED_FIXED, fixed_option=api.OPT_PULL_DENSE, time_kernel=tk, **kw)
            runs[tk] = (S.run(), S.result())
            S.free()
        (s1, r1), (s2, r2) = runs[1], runs[2]
        assert np.array_equal(r1, r2)
        assert s1["iterations"] == s2["iterations"] and s1["kernel_launches"] == s2["kernel_launches"]
        assert s2["kernel_bytes"] >= s1["kernel_bytes"]
        if algo == "cc":
            assert s2["kernel_bytes"] > s1["kernel_bytes"]
        G.free()
    finally:
        ctx.set_pull_tiling(3, 0, 0)


def test_bsp_pull_on_graph_tiled_by_another_mode(ctx, bsp_graphs):
    """A graph regrouped by the LDS-tiled pull first keeps that regrouping;
    a CC state created on it later pulls over the plain CSC, and is exact."""
    _say("bsp pull after mode 1")
    gs = bsp_graphs[0]
    ref, st = O.Partition(gs, 1, 4).cc()
    G = api.Graph.from_csr(ctx, gs.rp, gs.col, C=1, M=4)
    try:
        ctx.set_pull_tiling(1, 6, 1024)
        Sp = api.State(G, api.PR_PULL)
        ctx.set_pull_tiling(3, 6, 0)
        S = api.State(G, api.CC)
        S.init(sched=api.SCHED_FIXED, fixed_option=api.OPT_PULL_DENSE)
        stats = S.run()
        assert np.array_equal(S.result(), ref)
        assert stats["iterations"] == int(st[0])
        S.free()
        Sp.free()
    finally:
        ctx.set_pull_tiling(3, 0, 0)
        G.free()


def test_pull_tiling_argument_checks(ctx):
    with pytest.raises(api.GrdError):
        ctx.set_pull_tiling(1, 3, 0)
    with pytest.raises(api.GrdError):
        ctx.set_pull_tiling(1, 16, 0)
    with pytest.raises(api.GrdError):
        ctx.set_pull_tiling(2, 25, 0)
    with pytest.raises(api.GrdError):
        ctx.set_pull_tiling(4, 0, 0)
    with pytest.raises(api.GrdError):
        ctx.set_pull_tiling(3, 0, 4096)
    with pytest.raises(api.GrdError):
        ctx.set_pull_tiling(3, 16, 0)
    ctx.set_pull_tiling(3, 15, 0)
    ctx.set_pull_tiling(2, 24, 0)
    ctx.set_pull_tiling(3, 0, 0)
