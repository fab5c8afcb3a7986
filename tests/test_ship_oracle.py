"""CPU tests that pin the oracle's edge-shipping schedule (NO_OFFLOAD), its
coverage vector and the NDPEngine vote (no GPU).

Pins: coverage of hand-made graphs worked out by hand; the exchange bytes of
a path graph worked out by hand for the offloaded and the shipping
superstep; and the invariants the schedule must keep: CC labels and SSSP
distances (and their iteration counts) do not depend on the schedule, and
equal union-find components and Dijkstra distances.
"""
import numpy as np
import pytest

import oracle_lib as O


def test_coverage_by_hand():
    # arcs: 0->2, 0->4, 1->2, 1->3, 2->0, 3->5, 5->4; vertex 4 has none
    g = O.from_edges(6, [(0, 2), (0, 4), (1, 3), (1, 2), (2, 0), (3, 5), (5, 4)])
    assert O.Partition(g, 1, 2).coverage().tolist() == [1, 1, 1, 1, 1, 1]
    # C = 2: masters by parity; 1->2 and 5->4 leave their master, 3->5 does not
    assert O.Partition(g, 2, 2).coverage().tolist() == [1, 0, 1, 1, 1, 0]
    # C = 3: every vertex with an arc has one to another master
    assert O.Partition(g, 3, 1).coverage().tolist() == [0, 0, 0, 0, 1, 0]


def test_parity_doubled_graph_is_covered():
    g = O.parity_doubled(O.rmat(8, 8, seed=2))
    assert O.Partition(g, 2, 3).coverage().all()
    assert not O.Partition(O.rmat(8, 8, seed=2), 2, 3).coverage().all()


def test_path_bytes_by_hand():
    # SSSP hops from 0 on 0 -> 1 -> 2, C = M = 1: one pair with T = 3 (one
    # bitvector word, 8 bytes).  Offloaded superstep: broadcast 8 + 4 per
    # frontier value, reduce 8 + 4 per dirty proxy: 24 + 24 + 20 = 68.
    # Shipping superstep: broadcast as before plus 4 bytes per gid of
    # [src, out-neighbours], no reduce: (12 + 8) + (12 + 8) + (12 + 4) = 56.
    g = O.from_edges(3, [(0, 1), (1, 2)])
    P = O.Partition(g, 1, 1)
    d0, st0 = P.sssp(0)
    d2, st2, v2 = P.sssp(0, sched=2)
    assert d0.tolist() == d2.tolist() == [0, 1, 2]
    assert int(st0[0]) == int(st2[0]) == 3
    assert int(st0[1]) == 68 and int(st2[1]) == 56
    assert v2.tolist() == [1, 1, 1]
    # NDPEngine: n / 20 = 0, so a frontier of one is above the threshold
    d1, st1, v1 = P.sssp(0, sched=1)
    assert v1.tolist() == [0, 0, 0] and int(st1[1]) == 68
    # weighted: 4 more bytes per shipped arc (2 arcs)
    gw = O.Graph(g.rp, g.col, np.array([5, 7], dtype=np.uint32))
    dw, stw, _ = O.Partition(gw, 1, 1).sssp(0, weighted=True, sched=2)
    assert dw.tolist() == [0, 5, 12] and int(stw[1]) == 56 + 8


@pytest.mark.parametrize("C,M", [(1, 1), (1, 3), (2, 2), (3, 2)])
@pytest.mark.parametrize("sched", [1, 2])
def test_cc_independent_of_schedule(C, M, sched):
    for g in (O.rmat(11, 8, seed=3, symmetrize=True), O.parity_doubled(O.rmat(10, 8, seed=6, symmetrize=True))):
        P = O.Partition(g, C, M)
        ref, st = P.cc()
        out, st2, votes = P.cc(sched=sched)
        assert np.array_equal(out, ref) and int(st2[0]) == int(st[0])
        lab = np.zeros(g.n, dtype=np.uint32)
        O.lib().orc_cc_unionfind(g.n, g.rp, g.col, lab)
        assert np.array_equal(out, lab)
        cov = P.coverage().all()
        if not cov:
            # an uncovered frontier always votes for the offloaded superstep
            # (the first frontier is every vertex)
            assert votes[0] == 0


@pytest.mark.parametrize("C,M", [(1, 2), (2, 2)])
@pytest.mark.parametrize("kw", [{}, {"weighted": True}, {"weighted": True, "delta": 32}])
def test_sssp_independent_of_schedule(C, M, kw):
    g = O.parity_doubled(O.rmat(10, 8, seed=4).with_weights(5))
    P = O.Partition(g, C, M)
    ref, st = P.sssp(0, **kw)
    for sched in (1, 2):
        out, st2, votes = P.sssp(0, sched=sched, **kw)
        assert np.array_equal(out, ref) and int(st2[0]) == int(st[0])
        if kw.get("weighted"):
            assert np.array_equal(out, O.dijkstra(g, 0))
        if C == 2 and sched == 1:
            # the odd partition's frontier is always empty: every vote is NDP
            assert not votes.any()
        if sched == 2:
            assert votes.all()


def test_ndp_vote_threshold_on_one_partition():
    # C = 1: the vote ships exactly when 0 < |frontier| <= n / 20; the
    # frontier entering iteration i is the trace entry of iteration i - 1
    g = O.rmat(11, 8, seed=9, symmetrize=True)
    P = O.Partition(g, 1, 2)
    _, st, tr, votes = P.cc(trace=True, sched=1)
    before = np.concatenate([[g.n], tr[:-1]])
    assert np.array_equal(votes.astype(bool), (before > 0) & (before <= g.n // 20))
    assert votes.any() and not votes.all()


def test_pr_shipping_close_to_offloaded():
    g = O.rmat(10, 8, seed=5)
    P = O.Partition(g, 1, 2)
    ref, st = P.pr(fp64=True)
    out, st2, votes = P.pr(fp64=True, sched=2)
    # FP64 sums in another order: equal to rounding
    assert np.abs(out - ref).sum() <= 1e-9 * np.abs(ref).sum()
    assert int(st2[0]) == int(st[0]) and votes.all()


def _random_graph(rng, n, p):
    edges = [(u, v) for u in range(n) for v in range(n) if u != v and rng.random() < p]
    return O.from_edges(n, edges)


@pytest.mark.parametrize("seed", range(12))
def test_random_small_graphs_all_schedules(seed):
    """Brute force on small random graphs: every schedule gives the BFS
    distances (hop SSSP) and the union-find components (CC on the
    symmetrized graph), with the offloaded schedule's iteration count."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    g = _random_graph(rng, n, float(rng.uniform(0.02, 0.3)))
    es = [(u, int(v)) for u in range(n) for v in g.col[g.rp[u]:g.rp[u + 1]]]
    gs = O.from_edges(n, es + [(v, u) for u, v in es])
    C, M = int(rng.integers(1, 4)), int(rng.integers(1, 4))
    bfs = np.zeros(n, dtype=np.uint32)
    O.lib().orc_bfs(n, g.rp, g.col, 0, bfs)
    lab = np.zeros(n, dtype=np.uint32)
    O.lib().orc_cc_unionfind(n, gs.rp, gs.col, lab)
    P, Ps = O.Partition(g, C, M), O.Partition(gs, C, M)
    d0, st0 = P.sssp(0)
    c0, cst0 = Ps.cc()
    assert np.array_equal(d0, bfs) and np.array_equal(c0, lab)
    for sched in (1, 2):
        d, st, _ = P.sssp(0, sched=sched)
        c, cst, _ = Ps.cc(sched=sched)
        assert np.array_equal(d, bfs) and int(st[0]) == int(st0[0])
        assert np.array_equal(c, lab) and int(cst[0]) == int(cst0[0])
