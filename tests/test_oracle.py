"""CPU tests that pin the oracle (no GPU).

Each oracle function is checked against something it does not share code
with: the golden fixtures of tests/golden (textbook definitions), scipy /
networkx library routines, closed forms (BFS eccentricity + 1 supersteps),
or an invariant the reference's verify() states.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "tiny_graphs.json")))
INF = 0xFFFFFFFF
PARTS = [(1, 1), (1, 2), (1, 3), (2, 2), (3, 2)]


def gold_graph(name, weighted=False):
    g = GOLD[name]
    return O.Graph(g["row_ptr"], g["col"], g["weights"] if weighted else None), g


# ---------------- golden fixtures ----------------

@pytest.mark.parametrize("name", sorted(GOLD))
@pytest.mark.parametrize("C,M", PARTS)
def test_golden_cc(name, C, M):
    g, ref = gold_graph(name)
    lab, st = O.Partition(g, C, M).cc()
    assert lab.tolist() == ref["cc_label"]
    assert int(st[0]) == ref["cc_iters"]


@pytest.mark.parametrize("name", sorted(GOLD))
@pytest.mark.parametrize("C,M", PARTS)
def test_golden_sssp_hops(name, C, M):
    g, ref = gold_graph(name)
    d, st = O.Partition(g, C, M).sssp(0)
    assert d.tolist() == ref["hops"]
    assert int(st[0]) == ref["hop_iters"]
    assert O.bfs(g, 0).tolist() == ref["hops"]


@pytest.mark.parametrize("name", sorted(GOLD))
@pytest.mark.parametrize("delta", [0, 3])
def test_golden_sssp_weighted(name, delta):
    g, ref = gold_graph(name, weighted=True)
    d, _ = O.Partition(g, 1, 2).sssp(0, weighted=True, delta=delta)
    assert d.tolist() == ref["dijkstra"]
    assert O.dijkstra(g, 0).tolist() == ref["dijkstra"]


@pytest.mark.parametrize("name", sorted(GOLD))
def test_golden_pr_pull(name):
    g, ref = gold_graph(name)
    r, _, _ = O.pr_pull(g, iters=0, tol=1e-15)
    assert np.abs(r - np.array(ref["pr_pull"])).sum() < 1e-9
    r2, _, _ = O.pr_pull(g, iters=0, tol=1e-15, nthreads=2)
    assert np.abs(r2 - np.array(ref["pr_pull"])).sum() < 1e-9


@pytest.mark.parametrize("name", sorted(GOLD))
@pytest.mark.parametrize("C,M", [(1, 1), (2, 3)])
def test_golden_pr_delta_fixpoint(name, C, M):
    """The reference delta PageRank converges to x = b + d A^T D^-1 x; the
    activation threshold 1e-6 leaves at most a few residuals of that size."""
    g, ref = gold_graph(name)
    x64, st = O.Partition(g, C, M).pr(fp64=True)
    x = np.array(ref["pr_delta"])
    assert np.abs(x64 - x).max() < 1e-5
    assert st[3] == 0  # converged before MAX_ITERATIONS


# ---------------- closed forms on generated graphs ----------------

@pytest.fixture(scope="module")
def rmat_sym():
    return O.rmat(11, 8, seed=4, symmetrize=True)


@pytest.fixture(scope="module")
def rmat_dir():
    return O.rmat(11, 8, seed=6)


def _adj(g, w=None):
    data = np.ones(g.m) if w is None else w.astype(float)
    return sp.csr_matrix((data, g.col, g.rp), shape=(g.n, g.n))


@pytest.mark.parametrize("C,M", [(1, 1), (1, 4), (2, 3)])
def test_cc_matches_connected_components(rmat_sym, C, M):
    lab, _ = O.Partition(rmat_sym, C, M).cc()
    ncomp, comp = cg.connected_components(_adj(rmat_sym), directed=False)
    mins = np.full(ncomp, np.iinfo(np.int64).max)
    np.minimum.at(mins, comp, np.arange(rmat_sym.n))
    assert np.array_equal(lab, mins[comp])
    assert np.array_equal(O.cc_unionfind(rmat_sym), mins[comp])
    # cc.cpp:151-176 invariant: every edge joins equal labels
    src = np.repeat(np.arange(rmat_sym.n), np.diff(rmat_sym.rp.astype(np.int64)))
    assert np.array_equal(lab[src], lab[rmat_sym.col])


def test_cc_directed_backward_reachability(rmat_dir):
    lab, _ = O.Partition(rmat_dir, 1, 3).cc()
    # label(v) = min{u : u reaches v}: BFS forward from ascending sources,
    # the first source to reach v labels it
    A = _adj(rmat_dir)
    expect = np.full(rmat_dir.n, INF, dtype=np.int64)
    for s in range(rmat_dir.n):
        if expect[s] != INF:
            continue
        d = cg.breadth_first_order(A, s, directed=True, return_predecessors=False)
        d = d[expect[d] == INF]
        expect[d] = s
    assert np.array_equal(lab.astype(np.int64), expect)


@pytest.mark.parametrize("C,M", [(1, 1), (1, 4), (3, 2)])
def test_sssp_hop_is_bfs(rmat_dir, C, M):
    d, st = O.Partition(rmat_dir, C, M).sssp(0)
    ref = cg.shortest_path(_adj(rmat_dir), unweighted=True, indices=0)
    ref = np.where(np.isinf(ref), INF, ref).astype(np.int64)
    assert np.array_equal(d.astype(np.int64), ref)
    assert int(st[0]) == int(ref[ref != INF].max()) + 1  # eccentricity + 1 supersteps


@pytest.mark.parametrize("delta", [0, 16, 300])
def test_sssp_weighted_is_dijkstra(rmat_sym, delta):
    gw = rmat_sym.with_weights(3)
    src = 7
    d, _ = O.Partition(gw, 1, 4).sssp(src, weighted=True, delta=delta)
    ref = cg.dijkstra(_adj(gw, gw.w), indices=src)
    ref = np.where(np.isinf(ref), INF, ref).astype(np.int64)
    assert np.array_equal(d.astype(np.int64), ref)


def test_weights_symmetric_in_range(rmat_sym):
    gw = rmat_sym.with_weights(9)
    assert gw.w.min() >= 1 and gw.w.max() <= 255
    W = _adj(gw, gw.w)
    assert (W != W.T).nnz == 0


def test_pr_pull_matches_networkx(rmat_dir):
    import networkx as nx

    G = nx.DiGraph()
    G.add_nodes_from(range(rmat_dir.n))
    src = np.repeat(np.arange(rmat_dir.n), np.diff(rmat_dir.rp.astype(np.int64)))
    G.add_edges_from(zip(src.tolist(), rmat_dir.col.tolist()))
    ref = nx.pagerank(G, alpha=0.85, tol=1e-14, max_iter=10000)
    ref = np.array([ref[i] for i in range(rmat_dir.n)])
    r, _, _ = O.pr_pull(rmat_dir, iters=0, tol=1e-14, nthreads=4)
    assert np.abs(r - ref).sum() < 1e-9


def test_pr_delta_fp32_close_to_fp64(rmat_dir):
    a, st32 = O.Partition(rmat_dir, 1, 2).pr(fp64=False)
    b, st64 = O.Partition(rmat_dir, 1, 2).pr(fp64=True)
    assert np.abs(a / a.sum() - b / b.sum()).sum() < 1e-5
    assert abs(int(st32[0]) - int(st64[0])) <= 3


def test_bsp_deterministic_across_threads(rmat_sym, rmat_dir):
    P = O.Partition(rmat_sym, 1, 4)
    a, sa = P.cc(nthreads=1)
    b, sb = P.cc(nthreads=4)
    assert np.array_equal(a, b) and sa[0] == sb[0]
    Q = O.Partition(rmat_dir, 2, 3)
    c, sc = Q.sssp(0, nthreads=1)
    d, sd = Q.sssp(0, nthreads=4)
    assert np.array_equal(c, d) and sc[0] == sd[0]


def test_max_iterations_cap(rmat_dir):
    _, st = O.Partition(rmat_dir, 1, 2).pr(max_iter=3)
    assert int(st[0]) == 3 and int(st[3]) == 1


# ---------------- generators ----------------

@pytest.mark.parametrize("sym", [False, True])
def test_rmat_structure(sym):
    g = O.rmat(10, 16, seed=2, symmetrize=sym)
    rp = g.rp.astype(np.int64)
    src = np.repeat(np.arange(g.n), np.diff(rp))
    assert np.all(src != g.col)  # no self-loops
    keys = src * g.n + g.col
    assert np.all(np.diff(keys) > 0)  # sorted, no duplicates
    if sym:
        A = _adj(g)
        assert (A != A.T).nnz == 0


def test_rmat_quadrant_frequencies():
    # level 0 of every raw edge picks a quadrant with p = .57/.19/.19/.05:
    # the top bit of (src, dst) over raw edges before dedup follows it; after
    # dedup the lower-left quadrant (src top bit 0) still dominates
    g = O.rmat(12, 16, seed=11)
    half = g.n // 2
    rp = g.rp.astype(np.int64)
    src = np.repeat(np.arange(g.n), np.diff(rp))
    q = (src >= half) * 2 + (g.col >= half)
    frac = np.bincount(q, minlength=4) / len(q)
    assert frac[0] > frac[1] > frac[3] and frac[0] > frac[2] > frac[3]


def test_er_mean_degree():
    g = O.er(20000, 10.0, seed=5)
    assert abs(g.m / g.n - 10.0) < 0.2
    A = _adj(g)
    assert (A != A.T).nnz == 0


def test_hash_known_answers():
    # SplitMix64 finaliser (Steele, Lea, Flood 2014): the first outputs of
    # the generator seeded with 0 are mix(0 + gamma), mix(2 gamma), ...
    gamma = 0x9E3779B97F4A7C15
    def mix(z):
        z = (z + gamma) & (2**64 - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return z ^ (z >> 31)
    # published first output of SplitMix64 with seed 0
    assert mix(0) == 0xE220A8397B1DCDAF
    assert O.lib().orc_hash(0, mix(0)) == mix(0)  # mix(mix(0) ^ mix(0)) = mix(0)


# ---- the full-size checks of bench.py, pinned here ----

def test_pull_plan_equals_threaded_pull():
    """The reusable pull plan is the same recurrence, summed in the same
    order, as orc_pr_pull_mt (itself pinned by networkx above), and its
    iteration loop is timed."""
    g = O.rmat(12, 16, seed=7)
    P = O.PullPlan(g.rp, g.col, 4)
    assert P.arcs == g.m
    r, it, res, sec = P.run(iters=0, tol=1e-8)
    r2, it2, res2 = O.pr_pull(g, iters=0, tol=1e-8, nthreads=3)
    assert it == it2 and np.array_equal(r, r2) and res == res2
    assert sec > 0
    r3, it3, _, _ = P.run(iters=5)
    assert it3 == 5 and np.array_equal(r3, O.pr_pull(g, iters=5, nthreads=2)[0])
    P.free()


@pytest.mark.parametrize("scale,sym", [(10, True), (13, True), (13, False)])
def test_parallel_unionfind_equals_sequential(scale, sym):
    g = O.rmat(scale, 8, seed=scale, symmetrize=sym)
    ref = O.cc_unionfind(g)
    for nt in (1, 4, 8):
        assert np.array_equal(O.cc_unionfind_mt(g.rp, g.col, nt), ref)


def test_parallel_unionfind_min_gid_by_hand():
    # components {0, 3}, {1, 2, 4}, {5}
    g = O.from_edges(6, [(3, 0), (4, 2), (2, 1), (1, 4)])
    assert O.cc_unionfind_mt(g.rp, g.col, 2).tolist() == [0, 1, 1, 0, 1, 5]


def test_sssp_check_accepts_exact_and_rejects_planted_errors():
    g = O.rmat(12, 8, seed=3, symmetrize=True).with_weights(2)
    src = int(np.argmax(g.out_degree()))
    d = O.dijkstra(g, src)
    assert O.sssp_violations(g.rp, g.col, g.w, d, src, 4) == 0
    hop = O.bfs(g, src)
    assert O.sssp_violations(g.rp, g.col, None, hop, src, 4) == 0
    reached = np.flatnonzero((d != 0xFFFFFFFF) & (np.arange(g.n) != src))
    v = int(reached[len(reached) // 2])
    for delta in (-1, 1):  # too short (no tight arc) and too long (an arc relaxes it)
        bad = d.copy()
        bad[v] = int(bad[v]) + delta
        assert O.sssp_violations(g.rp, g.col, g.w, bad, src, 4) > 0, delta
    bad = d.copy()
    bad[v] = 0xFFFFFFFF  # reachable vertex reported unreached
    assert O.sssp_violations(g.rp, g.col, g.w, bad, src, 4) > 0
    bad = d.copy()
    bad[src] = 1
    assert O.sssp_violations(g.rp, g.col, g.w, bad, src, 4) > 0
    # the wrong weights (hop distances against weighted arcs) are refused
    assert O.sssp_violations(g.rp, g.col, g.w, hop, src, 4) > 0
