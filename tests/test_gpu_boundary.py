"""GPU tests of the drop-in boundary beyond the bulk superstep.

- aggregate(lid, value) per element (GraphAlgorithm.cpp:565-578 calls it
  with rTranslationTable[m][bit], a compute-local id): a whole CC / SSSP run
  at C = 2, M = 3 whose reduce is replaced by the oracle's reduce messages
  applied element by element through grd_aggregate_one equals the oracle
  after every superstep;
- 8-byte properties (CC<uint64_t>, SSSP<uint64_t>; PageRank<double> refused);
- snapshot / restore of the pull PageRank (its FP64 dangling mass lives in
  the device counters);
- the CLI's --value-bytes.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from grudon_amd import api

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "grudon_amd", "bin", "grudon")
U32_MAX = 0xFFFFFFFF
U64_MAX = 0xFFFFFFFFFFFFFFFF


def _say(*a):
    print(*a, flush=True)


def _widen(x):
    x = x.astype(np.uint64)
    x[x == U32_MAX] = U64_MAX
    return x


@pytest.mark.parametrize("algo,sym", [(api.CC, True), (api.SSSP, False)])
def test_per_element_aggregate_follows_oracle(ctx, algo, sym):
    """Every superstep: broadcast + traverse on the device, then the oracle's
    reduce message of each (c, m) pair applied element by element with
    aggregate(lid, value) on compute partition c (lid = gid / C), then the
    update; the state equals the oracle's after the same superstep."""
    C, M = 2, 3
    _say("per-element aggregate", algo, C, M)
    g = O.rmat(11, 8, seed=13, symmetrize=sym)
    P = O.Partition(g, C, M)
    oalgo = O.ORC_CC if algo == api.CC else O.ORC_SSSP
    final, st = P.cc() if algo == api.CC else P.sssp(0)
    n_iter = int(st[0])
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=C, M=M)
    S = api.State(G, algo)
    S.init(source=0)
    L = api.lib()
    applied = 0
    for it in range(n_iter):
        assert L.grd_broadcast(S.h, api.OPT_PUSH_SPARSE) == 0
        assert L.grd_traverse(S.h, api.OPT_PUSH_SPARSE) == 0
        for c in range(C):
            for m in range(M):
                gids, vals = P.pair_exchange(oalgo, it, 1, c, m)
                for gid, v in zip(gids.tolist(), vals.tolist()):
                    assert gid % C == c
                    S.aggregate_one(c, gid // C, v)
                    applied += 1
        fr = S.step_update()
        want = P.cc(max_iter=it + 1)[0] if algo == api.CC else P.sssp(0, max_iter=it + 1)[0]
        assert np.array_equal(S.result(), want), it
        assert (fr == 0) == (it == n_iter - 1), it
    assert applied > 0
    assert np.array_equal(S.result(), final)
    S.free()
    G.free()


def test_per_element_aggregate_equals_bulk_reduce_pr(ctx):
    """Reference PageRank (float add): the first superstep's reduce applied
    element by element equals the bulk grd_reduce, up to the order of the
    float additions."""
    C, M = 2, 3
    _say("per-element aggregate pr")
    g = O.rmat(11, 8, seed=17)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=C, M=M)
    A = api.State(G, api.PR)
    B = api.State(G, api.PR)
    A.init()
    B.init()
    L = api.lib()
    for S in (A, B):
        assert L.grd_broadcast(S.h, api.OPT_PUSH_SPARSE) == 0
        assert L.grd_traverse(S.h, api.OPT_PUSH_SPARSE) == 0
    assert L.grd_reduce(A.h, api.OPT_PUSH_SPARSE) == 0
    n = 0
    for c in range(C):
        for m in range(M):
            gids, bits = A.pair_exchange(1, c, m)
            for gid, b in zip(gids.tolist(), bits.view(np.float32).tolist()):
                B.aggregate_one(c, gid // C, b)
                n += 1
    fa, fb = A.step_update(), B.step_update()
    assert n > 0 and fa == fb
    ra, rb = A.result(), B.result()
    np.testing.assert_allclose(rb, ra, rtol=1e-6, atol=1e-9)
    A.free()
    B.free()
    G.free()


def test_aggregate_argument_checks(ctx):
    g = O.rmat(8, 8, seed=3)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=2, M=2)
    S = api.State(G, api.CC)
    S.init()
    with pytest.raises(api.GrdError, match="compute partition"):
        S.aggregate_one(2, 0, 1)
    with pytest.raises(api.GrdError, match="local id"):
        S.aggregate_one(1, g.n // 2, 1)  # gid = n/2 * 2 + 1 >= n
    S.aggregate_one(1, g.n // 2 - 1, 0)  # gid n - 1: fine
    S.free()
    G.free()


@pytest.mark.parametrize("C,M", [(1, 2), (2, 3)])
def test_wide_properties(ctx, C, M):
    """CC<uint64_t> and SSSP<uint64_t>: the oracle's labels and distances
    widened (unreached = 2^64 - 1), the same iteration counts, 8 bytes per
    exchanged value."""
    _say("wide", C, M)
    gs = O.rmat(11, 8, seed=23, symmetrize=True)
    gd = O.rmat(11, 8, seed=29).with_weights(5)
    for g, algo, kw in [(gs, api.CC, {}), (gd, api.SSSP, {"source": 3}),
                        (gd, api.SSSP, {"source": 3, "weighted": True, "delta": 32})]:
        P = O.Partition(g, C, M)
        ref, st = P.cc() if algo == api.CC else P.sssp(kw["source"], weighted=kw.get("weighted", False),
                                                       delta=kw.get("delta", 0))
        G = api.Graph.from_csr(ctx, g.rp, g.col, g.w, C=C, M=M)
        stats = {}
        for vb in (4, 8):
            S = api.State(G, algo, value_bytes=vb)
            S.init(**kw)
            stats[vb] = S.run()
            r = S.result()
            if vb == 8:
                assert r.dtype == np.uint64
                assert np.array_equal(r, _widen(ref))
                if algo == api.SSSP:
                    assert (r == U64_MAX).any()  # some vertices are unreached
            else:
                assert np.array_equal(r, ref)
            S.free()
        assert stats[4]["iterations"] == stats[8]["iterations"] == int(st[0])
        b4, b8 = stats[4]["bytes_exchanged"], stats[8]["bytes_exchanged"]
        assert b8 > b4 and (b8 - b4) % 4 == 0
        G.free()


def test_wide_exchange_and_aggregate(ctx):
    g = O.rmat(10, 8, seed=31, symmetrize=True)
    P = O.Partition(g, 2, 2)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=2, M=2)
    S = api.State(G, api.CC, value_bytes=8)
    S.init()
    S.step(api.OPT_PUSH_SPARSE)
    gids, vals = S.pair_exchange(1, 1, 0)
    og, ov = P.pair_exchange(O.ORC_CC, 0, 1, 1, 0)
    assert vals.dtype == np.uint64
    assert np.array_equal(gids, og) and np.array_equal(vals, _widen(ov))
    # a value beyond 32 bits cannot improve a label
    S.aggregate_one(0, 0, U64_MAX)
    S.free()
    G.free()


def test_pagerank_double_is_refused(ctx):
    g = O.rmat(8, 8, seed=3)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=1, M=1)
    for algo in (api.PR, api.PR_PULL):
        with pytest.raises(api.GrdError, match="PageRank<double>"):
            api.State(G, algo, value_bytes=8)
    with pytest.raises(api.GrdError, match="4 or 8"):
        api.State(G, api.CC, value_bytes=2)
    G.free()


def test_pr_pull_snapshot_restore_is_exact(ctx):
    """Snapshot at iteration k, run three more, restore, run to the end: the
    ranks, iteration count and residual equal an uninterrupted run bit for
    bit (the FP64 dangling mass of the next update is restored too)."""
    _say("pr pull snapshot")
    g = O.rmat(12, 16, seed=41)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=1, M=3)
    A = api.State(G, api.PR_PULL)
    A.init(tolerance=1e-7)
    sa = A.run()
    B = api.State(G, api.PR_PULL)
    B.init(tolerance=1e-7)
    for _ in range(4):
        B.step(api.OPT_PULL_DENSE)
    B.snapshot()
    for _ in range(3):
        B.step(api.OPT_PULL_DENSE)
    B.restore()
    sb = B.run()
    assert sb["iterations"] == sa["iterations"]
    assert sb["residual"] == sa["residual"]
    assert sb["bytes_exchanged"] == sa["bytes_exchanged"]
    assert np.array_equal(B.result().view(np.uint32), A.result().view(np.uint32))
    A.free()
    B.free()
    G.free()


def test_cli_value_bytes(tmp_path):
    g = O.rmat(10, 8, seed=3, symmetrize=True)
    ref, _ = O.Partition(g, 1, 2).cc()
    res = tmp_path / "cc64.bin"
    p = subprocess.run([CLI, "-g", "rmat:10:8:3:sym", "-c", "1", "-m", "2", "-a", "cc", "--value-bytes", "8",
                        "--verify", "--result", str(res)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    assert "CC verify: 0 edges" in p.stdout
    assert np.array_equal(np.fromfile(res, dtype=np.uint64), ref.astype(np.uint64))
    p = subprocess.run([CLI, "-g", "rmat:8:8:3", "-c", "1", "-m", "1", "-a", "pr", "--value-bytes", "8"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 2 and "PageRank<double>" in p.stdout, p.stdout
