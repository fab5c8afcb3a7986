"""GPU tests of the batched form, the selector kernels, the step-wise
operator API and the C++ CLI."""
import os
import subprocess

import numpy as np
import pytest

import mlp_oracle as M
import oracle_lib as O
from grudon_amd import api

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def small_batch_graphs():
    gs = [O.rmat(s, ef, seed=100 + i) for i, (s, ef) in enumerate([(10, 8), (11, 16), (9, 32), (12, 8)])]
    gs += [O.er(n, d, seed=200 + i) for i, (n, d) in enumerate([(3000, 4.0), (5000, 16.0), (1500, 40.0)])]
    gs.append(O.from_edges(5, [(0, 1), (1, 2), (3, 4)]))
    return gs


@pytest.mark.parametrize("tile,gpc,K", [(2048, 8, 4), (1024, 4, 2), (4096, 16, 8)])
def test_batch_pagerank_fp64_bound(ctx, small_batch_graphs, tile, gpc, K):
    _say("batch pr", tile, gpc, K)
    B = api.Batch(ctx, [(g.rp, g.col) for g in small_batch_graphs])
    B.tune(tile, gpc, K)
    iters, ms = B.pagerank(tolerance=1e-6, max_iter=200)
    for i, g in enumerate(small_batch_graphs):
        ref, _, _ = O.pr_pull(g, iters=int(iters[i]))
        l1 = np.abs(B.result(i, api.PR_PULL).astype(np.float64) - ref).sum()
        assert l1 <= 1e-6, (i, l1)
        # the graph stopped at its own convergence point
        _, it_ref, _ = O.pr_pull(g, iters=0, tol=1e-6)
        assert abs(int(iters[i]) - it_ref) <= 1


@pytest.mark.parametrize("tile", [1024, 2048])
def test_batch_cc_bit_exact(ctx, small_batch_graphs, tile):
    _say("batch cc", tile)
    B = api.Batch(ctx, [(g.rp, g.col) for g in small_batch_graphs])
    B.tune(tile, 8, 4)
    iters, _ = B.cc()
    for i, g in enumerate(small_batch_graphs):
        ref, st = O.Partition(g, 1, 1).cc()
        assert np.array_equal(B.result(i, api.CC), ref), i
        assert int(iters[i]) == int(st[0]), i


def test_batch_autotune_keeps_results(ctx, small_batch_graphs):
    _say("batch autotune")
    B = api.Batch(ctx, [(g.rp, g.col) for g in small_batch_graphs])
    chosen = B.autotune()
    assert chosen["tile"] in (1024, 2048, 4096)
    iters, _ = B.cc()
    for i, g in enumerate(small_batch_graphs):
        ref, _ = O.Partition(g, 1, 1).cc()
        assert np.array_equal(B.result(i, api.CC), ref)


def test_selector_gradient_matches_fp64(ctx):
    _say("selector gradient")
    rng = np.random.default_rng(5)
    w = M.init(2).astype(np.float32)
    x = rng.normal(size=(300, 16)).astype(np.float32)
    y = rng.integers(0, 4, size=300).astype(np.uint32)
    sw = rng.uniform(0.2, 3.0, size=300).astype(np.float32)
    loss, g = ctx.selector_grad(w, x, y, sw)
    loss64, g64 = M.loss_grad(w.astype(np.float64), x, y.astype(np.int64), sw)
    assert abs(loss - loss64) <= 1e-4 * abs(loss64)
    err = np.abs(g - g64).max() / max(1e-12, np.abs(g64).max())
    assert err < 1e-4, err


def test_selector_training_learns_separable_rule(ctx):
    _say("selector training")
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, size=(4000, 16)).astype(np.float32)
    # option = (pull if f2 > .3) | (dense if f1 > .5)
    y = ((x[:, 2] > 0.3).astype(np.uint32) | ((x[:, 1] > 0.5).astype(np.uint32) << 1))
    w, loss = ctx.selector_train(x, y, epochs=40, batch=128, lr=3e-3, seed=3)
    logits = ctx.selector_infer(x)
    acc = float((logits.argmax(1) == y).mean())
    _say("   loss", loss, "acc", acc)
    assert acc > 0.9
    # inference with the BF16 weights agrees with the FP64 oracle's forward
    wf = np.asarray(w, dtype=np.uint16).astype(np.uint32) << 16
    ref, _ = M.forward(wf.view(np.float32).astype(np.float64), x[:64].astype(np.float64))
    assert np.abs(logits[:64] - ref).max() < 1e-3 * max(1.0, np.abs(ref).max())


def test_stepwise_api_matches_run(ctx):
    _say("stepwise")
    g = O.rmat(11, 16, seed=9, symmetrize=True)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=1, M=3)
    S = api.State(G, api.CC)
    S.init(sched=api.SCHED_HEURISTIC)
    n_iter = 0
    while True:
        f = S.step(api.OPT_PULL_DENSE if n_iter % 2 else api.OPT_PUSH_SPARSE)
        n_iter += 1
        if f == 0:
            break
    ref, st = O.Partition(g, 1, 3).cc()
    assert np.array_equal(S.result(), ref) and n_iter == int(st[0])


def test_stepwise_direction_must_match_broadcast(ctx):
    _say("stepwise direction")
    g = O.rmat(9, 8, seed=4, symmetrize=True)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=1, M=2)
    S = api.State(G, api.CC)
    S.init()
    L = api.lib()
    assert L.grd_traverse(S.h, api.OPT_PUSH_SPARSE) == 4  # no broadcast yet: GRD_ERR_STATE
    assert L.grd_broadcast(S.h, api.OPT_PUSH_SPARSE) == 0
    assert L.grd_traverse(S.h, api.OPT_PULL_SPARSE) == 4
    assert b"direction" in L.grd_last_error()
    # the matching direction still completes the superstep correctly
    assert L.grd_traverse(S.h, api.OPT_PUSH_DENSE) == 0
    S.free()


def test_cli_runs_all_algorithms(ctx, tmp_path):
    _say("cli")
    cli = os.path.join(ROOT, "grudon_amd", "bin", "grudon")
    g = O.rmat(11, 8, seed=3, symmetrize=True)
    res = tmp_path / "cc.bin"
    p = subprocess.run([cli, "-g", "rmat:11:8:3:sym", "-c", "1", "-m", "2", "-a", "cc", "--verify",
                        "--result", str(res), "--sched", "heuristic"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    assert "CC verify: 0 edges" in p.stdout
    ref, st = O.Partition(g, 1, 2).cc()
    assert np.array_equal(np.fromfile(res, dtype=np.uint32), ref)
    assert "STAT, CC, Iterations, %d" % int(st[0]) in p.stdout
    for algo in ("pr", "sssp", "prpull"):
        p = subprocess.run([cli, "-g", "rmat:10:8:3", "-c", "1", "-m", "2", "-a", algo, "--verify"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
        assert p.returncode == 0, p.stdout
    gr = tmp_path / "g.gr"
    p = subprocess.run([cli, "-g", "rmat:10:8:3:sym:5", "-c", "1", "-m", "1", "-a", "sssp", "--weighted",
                        "--write-gr", str(gr)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout
    res2 = tmp_path / "sssp.bin"
    p = subprocess.run([cli, "-g", str(gr), "-c", "2", "-m", "3", "-a", "sssp", "--weighted", "--verify",
                        "--result", str(res2)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout
    gw = O.rmat(10, 8, seed=3, symmetrize=True).with_weights(5)
    assert np.array_equal(np.fromfile(res2, dtype=np.uint32), O.dijkstra(gw, 0))
    # learned schedule with the packaged weights and a partition file
    pf = tmp_path / "parts.txt"
    mirror = (np.arange(g.n) * 5 + 1) % 3
    pf.write_text("".join("%d\n" % x for x in mirror))
    res3 = tmp_path / "cc3.bin"
    p = subprocess.run([cli, "-g", "rmat:11:8:3:sym", "-c", "2", "-m", "3", "-p", str(pf), "-a", "cc", "--sched",
                        "learned", "--selector", os.path.join(ROOT, "grudon_amd", "selector_bf16.bin"), "--verify",
                        "--result", str(res3)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout
    ref3, st3 = O.Partition(g, 2, 3, mirror=mirror.astype(np.uint32)).cc()
    assert np.array_equal(np.fromfile(res3, dtype=np.uint32), ref3)
    assert "STAT, CC, Iterations, %d" % int(st3[0]) in p.stdout
    # the reference's offload vote, and edge shipping every superstep
    for sched in ("ndp", "fixed:4"):
        res4 = tmp_path / ("cc_%s.bin" % sched.replace(":", ""))
        p = subprocess.run([cli, "-g", "rmat:11:8:3:sym", "-c", "1", "-m", "2", "-a", "cc", "--sched", sched,
                            "--verify", "--result", str(res4)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert p.returncode == 0, p.stdout
        assert np.array_equal(np.fromfile(res4, dtype=np.uint32), ref)
        assert "STAT, CC, Iterations, %d" % int(st[0]) in p.stdout


def test_gather_probe_runs(ctx):
    """The bench's measurement library (not part of libgrudon_hip.so)."""
    _say("gather probe")
    ms, rate = api.probe_gather(0, 1 << 16, 1 << 20)
    assert ms > 0 and rate > 1e8
    with pytest.raises(api.GrdError):
        api.probe_gather(0, 0, 1 << 20)
    assert api.probe_stream(0, 1 << 24, 2) > 0
    assert api.probe_stream(0, 1 << 24, 2, write=True) > 0


def test_learned_run_decisions_match_host_features(ctx):
    """The learned run computes each next decision on the device from the
    counters; it must equal the MLP applied to the host-side features that
    the trace records (ties within 1e-3 of the top logit excepted)."""
    _say("learned decisions")
    w = np.fromfile(os.path.join(ROOT, "grudon_amd", "selector_bf16.bin"), dtype="<u2")
    ctx.selector_load(w, min_margin=0.0)
    for sym, algo, kw in [(True, api.CC, {}), (False, api.SSSP, {"source": 1, "weighted": True, "delta": 64}),
                          (False, api.PR, {"max_iter": 40})]:
        g = O.rmat(13, 16, seed=21, symmetrize=sym)
        if algo == api.SSSP:
            g = g.with_weights(3)
        G = api.Graph.from_csr(ctx, g.rp, g.col, g.w, C=1, M=2)
        S = api.State(G, algo)
        S.init(sched=api.SCHED_LEARNED, **kw)
        st, tr = S.run(trace_cap=2000)
        feats = np.array([np.frombuffer(bytes(r.feat), dtype=np.float32) for r in tr])
        logits = ctx.selector_infer(feats)
        for r, lg in zip(tr, logits):
            top = np.sort(lg)[-2:]
            if top[1] - top[0] > 1e-3:
                # a push decision runs as edge shipping (C = 1: always covered)
                want = int(np.argmax(lg))
                want = api.OPT_SHIP_EDGES if want & 1 == 0 else want
                assert r.option == want, (algo, r.iteration)
        S.free()


def test_batch_pagerank_repeats_bit_exact(ctx, small_batch_graphs):
    """Residual and dangling mass are reduced per vertex tile and summed in
    tile order: two runs give the same iterations and bit-identical ranks."""
    _say("batch pr repeat")
    B = api.Batch(ctx, [(g.rp, g.col) for g in small_batch_graphs])
    try:
        B.tune(1024, 16, 4)
        it1, _ = B.pagerank(tolerance=1e-7, max_iter=300)
        r1 = [B.result(i, api.PR_PULL).view(np.uint32) for i in range(B.count)]
        it2, _ = B.pagerank(tolerance=1e-7, max_iter=300)
        r2 = [B.result(i, api.PR_PULL).view(np.uint32) for i in range(B.count)]
        assert np.array_equal(it1, it2)
        assert all(np.array_equal(a, b) for a, b in zip(r1, r2))
    finally:
        B.free()
