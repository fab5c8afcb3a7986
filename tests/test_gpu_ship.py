"""GPU parity of the edge-shipping (NO_OFFLOAD) schedule, the coverage
vector and the NDPEngine vote, against the oracle through the C-ABI.

Bar: coverage equal gid for gid; CC labels and SSSP distances bit-exact with
equal iteration counts; the option of every iteration equal to the oracle's
vote; exchange bytes equal to the oracle's count; reference PageRank within
the FP32-vs-FP64 bound of the offloaded schedule's test.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from grudon_amd import api

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(*a):
    print(*a, flush=True)


GRAPHS = {
    "rmat11_sym": lambda: O.rmat(11, 8, seed=3, symmetrize=True),
    "doubled_sym": lambda: O.parity_doubled(O.rmat(11, 8, seed=6, symmetrize=True)),
    "doubled_dir_w": lambda: O.parity_doubled(O.rmat(11, 8, seed=4).with_weights(5)),
    "rmat14_sym": lambda: O.rmat(14, 16, seed=8, symmetrize=True),
}
_cache = {}


def graph(name):
    if name not in _cache:
        _cache[name] = GRAPHS[name]()
    return _cache[name]


def _options(trace):
    return np.array([r.option for r in trace])


@pytest.mark.parametrize("name", ["rmat11_sym", "doubled_sym", "doubled_dir_w"])
@pytest.mark.parametrize("C,M", [(1, 1), (1, 3), (2, 2), (3, 2)])
def test_coverage_equals_oracle(ctx, name, C, M):
    _say("coverage", name, C, M)
    g = graph(name)
    G = api.Graph.from_csr(ctx, g.rp, g.col, g.w, C=C, M=M)
    cov, unc = G.coverage()
    ref = O.Partition(g, C, M).coverage()
    assert np.array_equal(cov, ref)
    assert unc == int((ref == 0).sum())


@pytest.mark.parametrize("name", ["rmat11_sym", "doubled_sym", "rmat14_sym"])
@pytest.mark.parametrize("C,M", [(1, 2), (2, 3)])
@pytest.mark.parametrize("osched,sched", [(2, api.SCHED_FIXED), (1, api.SCHED_NDP_ENGINE), (0, api.SCHED_PLAIN)])
def test_cc_schedules(ctx, name, C, M, osched, sched):
    _say("cc ship", name, C, M, sched)
    g = graph(name)
    P = O.Partition(g, C, M)
    res = P.cc(sched=osched)
    ref, st = res[0], res[1]
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=C, M=M)
    S = api.State(G, api.CC)
    S.init(sched=sched, fixed_option=api.OPT_SHIP_EDGES)
    stats, tr = S.run(trace_cap=1000)
    assert np.array_equal(S.result(), ref)
    assert stats["iterations"] == int(st[0])
    assert stats["bytes_exchanged"] == int(st[1])
    if osched:
        want = np.where(res[2] == 1, api.OPT_SHIP_EDGES, api.OPT_PUSH_SPARSE)
        assert np.array_equal(_options(tr), want)
        assert stats["option_counts"][api.OPT_SHIP_EDGES] == int(res[2].sum())
    S.free()


@pytest.mark.parametrize("C,M", [(1, 2), (2, 2)])
@pytest.mark.parametrize("kw", [{}, {"weighted": True}, {"weighted": True, "delta": 32}])
@pytest.mark.parametrize("osched,sched", [(2, api.SCHED_FIXED), (1, api.SCHED_NDP_ENGINE)])
def test_sssp_schedules(ctx, C, M, kw, osched, sched):
    _say("sssp ship", C, M, kw, sched)
    g = graph("doubled_dir_w")
    P = O.Partition(g, C, M)
    ref, st, votes = P.sssp(0, sched=osched, **kw)
    G = api.Graph.from_csr(ctx, g.rp, g.col, g.w, C=C, M=M)
    S = api.State(G, api.SSSP)
    S.init(source=0, max_iter=100000, sched=sched, fixed_option=api.OPT_SHIP_EDGES, **kw)
    stats, tr = S.run(trace_cap=100000)
    assert np.array_equal(S.result(), ref)
    assert stats["iterations"] == int(st[0])
    assert stats["bytes_exchanged"] == int(st[1])
    want = np.where(votes == 1, api.OPT_SHIP_EDGES, api.OPT_PUSH_SPARSE)
    assert np.array_equal(_options(tr), want)
    if kw.get("weighted"):
        assert np.array_equal(S.result(), O.dijkstra(g, 0))
    S.free()


def test_fixed_shipping_falls_back_when_uncovered(ctx):
    _say("ship fallback")
    g = graph("rmat11_sym")
    ref, st = O.Partition(g, 2, 2).cc()
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=2, M=2)
    assert G.coverage()[1] > 0
    S = api.State(G, api.CC)
    S.init(sched=api.SCHED_FIXED, fixed_option=api.OPT_SHIP_EDGES)
    stats = S.run()
    assert np.array_equal(S.result(), ref)
    assert stats["option_counts"][api.OPT_SHIP_EDGES] == 0
    assert stats["fallbacks"] == stats["iterations"]
    S.free()


def test_stepwise_shipping(ctx):
    _say("ship stepwise")
    g = graph("rmat11_sym")
    # uncovered frontier: the traversal refuses to ship
    G2 = api.Graph.from_csr(ctx, g.rp, g.col, C=2, M=2)
    S = api.State(G2, api.CC)
    S.init()
    L = api.lib()
    assert L.grd_broadcast(S.h, api.OPT_SHIP_EDGES) == 0
    assert L.grd_traverse(S.h, api.OPT_SHIP_EDGES) == 4  # GRD_ERR_STATE
    assert b"covered" in L.grd_last_error()
    S.free()
    # covered (C = 1): alternating shipping and offloaded supersteps
    G1 = api.Graph.from_csr(ctx, g.rp, g.col, C=1, M=3)
    S = api.State(G1, api.CC)
    S.init()
    n_iter = 0
    while S.step(api.OPT_SHIP_EDGES if n_iter % 2 == 0 else api.OPT_PULL_DENSE):
        n_iter += 1
    n_iter += 1
    ref, st = O.Partition(g, 1, 3).cc()
    assert np.array_equal(S.result(), ref) and n_iter == int(st[0])
    S.free()


def test_snapshot_restore_with_shipping(ctx):
    _say("ship snapshot")
    g = graph("doubled_dir_w")
    G = api.Graph.from_csr(ctx, g.rp, g.col, g.w, C=1, M=2)
    S = api.State(G, api.SSSP)
    S.init(source=0, weighted=True, sched=api.SCHED_FIXED, fixed_option=api.OPT_SHIP_EDGES)
    S.step(api.OPT_SHIP_EDGES)
    S.step(api.OPT_SHIP_EDGES)
    S.snapshot()
    S.step(api.OPT_SHIP_EDGES)
    S.restore()
    S.run()
    assert np.array_equal(S.result(), O.dijkstra(g, 0))
    S.free()


@pytest.mark.parametrize("C,M", [(1, 1), (1, 4)])
def test_pr_shipping_reference_semantics(ctx, C, M):
    _say("pr ship", C, M)
    g = O.rmat(12, 16, seed=7)
    P = O.Partition(g, C, M)
    ref64, st = P.pr(fp64=True)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=C, M=M)
    S = api.State(G, api.PR)
    S.init(sched=api.SCHED_FIXED, fixed_option=api.OPT_SHIP_EDGES)
    stats = S.run()
    out = S.result().astype(np.float64)
    l1 = np.abs(out / out.sum() - ref64 / ref64.sum()).sum()
    ref32, _ = P.pr(fp64=False)
    l1_32 = np.abs(ref32 / ref32.sum() - ref64 / ref64.sum()).sum()
    _say("   L1 normalised", l1, "bound", max(1e-6, 4 * l1_32), "iters", stats["iterations"], int(st[0]))
    assert l1 <= max(1e-6, 4 * l1_32)
    assert abs(stats["iterations"] - int(st[0])) <= 3
    assert stats["option_counts"][api.OPT_SHIP_EDGES] == stats["iterations"]
    S.free()


@pytest.mark.parametrize("sched", [api.SCHED_HEURISTIC, api.SCHED_LEARNED])
def test_heuristic_and_learned_ship_unless_offload_only(ctx, sched):
    _say("ship refine", sched)
    g = graph("doubled_dir_w")
    w = np.fromfile(os.path.join(ROOT, "grudon_amd", "selector_bf16.bin"), dtype="<u2")
    ctx.selector_load(w, min_margin=0.0)
    ref = O.dijkstra(g, 0)
    G = api.Graph.from_csr(ctx, g.rp, g.col, g.w, C=1, M=2)
    S = api.State(G, api.SSSP)
    S.init(source=0, weighted=True, delta=32, max_iter=100000, sched=sched)
    st = S.run()
    assert np.array_equal(S.result(), ref)
    # every push decision ran as edge shipping (C = 1: always covered)
    assert st["option_counts"][api.OPT_PUSH_SPARSE] == st["option_counts"][api.OPT_PUSH_DENSE] == 0
    if sched == api.SCHED_HEURISTIC:  # the first frontier is one vertex: a push
        assert st["option_counts"][api.OPT_SHIP_EDGES] > 0
    S.init(source=0, weighted=True, delta=32, max_iter=100000, sched=sched, offload_only=True)
    st = S.run()
    assert np.array_equal(S.result(), ref)
    assert st["option_counts"][api.OPT_SHIP_EDGES] == 0
    S.free()


@pytest.mark.parametrize("osched,sched", [(2, api.SCHED_FIXED), (1, api.SCHED_NDP_ENGINE)])
def test_shipping_with_partition_file(ctx, osched, sched):
    """Mirrors from a partition file (not gid % M), two compute partitions,
    every vertex covered: the shipping schedules against the oracle."""
    _say("ship partition file", sched)
    g = graph("doubled_sym")
    mirror = ((np.arange(g.n, dtype=np.uint64) * 2654435761) % 3).astype(np.uint32)
    P = O.Partition(g, 2, 3, mirror=mirror)
    ref, st, votes = P.cc(sched=osched)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=2, M=3, mirror=mirror)
    assert G.coverage()[1] == 0
    S = api.State(G, api.CC)
    S.init(sched=sched, fixed_option=api.OPT_SHIP_EDGES)
    stats, tr = S.run(trace_cap=1000)
    assert np.array_equal(S.result(), ref)
    assert stats["iterations"] == int(st[0]) and stats["bytes_exchanged"] == int(st[1])
    want = np.where(votes == 1, api.OPT_SHIP_EDGES, api.OPT_PUSH_SPARSE)
    assert np.array_equal(_options(tr), want)
    S.free()
