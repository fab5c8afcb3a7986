"""Build the schedule selector: label iterations by timing, train on the GPU.

    python -m grudon_amd.train_selector --budget 240 --out grudon_amd/selector_bf16.bin
    python -m grudon_amd.train_selector --budget 900 --seed 2 --records r2.npz --no-train
    python -m grudon_amd.train_selector --load-records r1.npz,r2.npz --out grudon_amd/selector_bf16.bin

1. Records.  Seeded RMAT (scale 16-22, edge factor 8-32) and Erdos-Renyi
   graphs (n 2^16-2^21, mean degree 4-64), partitioned with M in {1, 2, 4},
   run CC (symmetric), SSSP hop count and weighted delta-stepping SSSP
   (weights [1,255], several sources) and the reference delta PageRank once
   per schedule option (push/pull x sparse/dense, and edge shipping) with
   per-iteration device timing.  CC and SSSP supersteps are deterministic,
   so the runs visit identical states and iteration i of each run is the
   same decision; PageRank iterations are kept only where the runs still
   agree on the frontier size.  A record is (the 16 features before the
   iteration, the measured times); its label is the fastest class.  The
   learned schedule runs a push decision as edge shipping whenever the
   frontier is covered (always, with one compute partition), so both push
   classes are charged the shipping time when it was measured.
2. Training.  grd_selector_train runs the hand-written forward / backward /
   Adam kernels (mlp.hip) on the records and returns BF16 weights.
3. The weights are written as 5,508 little-endian uint16 values.
"""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from grudon_amd import api  # noqa: E402

OPTIONS = [api.OPT_PUSH_SPARSE, api.OPT_PULL_SPARSE, api.OPT_PUSH_DENSE, api.OPT_PULL_DENSE, api.OPT_SHIP_EDGES]


def class_costs(costs):
    """Cost of each of the four selector classes: records with a measured
    edge-shipping time (fifth column) charge it to both push classes."""
    if costs.shape[1] == 5:
        return costs[:, [4, 1, 4, 3]]
    return costs


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_options(S, init_kw, max_iter):
    """Run the same problem once per option; return traces."""
    traces = []
    for o in OPTIONS:
        S.init(sched=api.SCHED_FIXED, fixed_option=o, time_iters=True, max_iter=max_iter, **init_kw)
        _, tr = S.run(trace_cap=max_iter + 1)
        traces.append(tr)
    return traces


def records_from(traces, exact):
    feats, costs = [], []
    n = min(len(t) for t in traces)
    for i in range(n):
        fr = [t[i].frontier for t in traces]
        if not exact and len(set(fr)) != 1:
            break  # PageRank runs diverged in their float order
        if exact and len(set(fr)) != 1:
            raise RuntimeError("deterministic supersteps disagree across options")
        feats.append(np.frombuffer(bytes(traces[0][i].feat), dtype=np.float32).copy())
        costs.append([t[i].ms for t in traces])
    return feats, costs


def collect(ctx, budget_s, seed=1, max_scale=22, pr_max_scale=20):
    rng = np.random.default_rng(seed)
    feats, costs = [], []
    t0 = time.time()
    g_count = 0
    while time.time() - t0 < budget_s:
        kind = rng.choice(["rmat", "er"], p=[0.7, 0.3])
        M = int(rng.choice([1, 2, 4]))
        wseed = int(rng.integers(1, 1 << 30))
        if kind == "rmat":
            scale = int(rng.integers(16, max_scale + 1))
            ef = int(rng.choice([8, 16, 32]))
            gseed = int(rng.integers(1, 1 << 30))
            Gd = api.Graph.rmat(ctx, scale, ef, seed=gseed, wseed=wseed, C=1, M=M)
            Gs = api.Graph.rmat(ctx, scale, ef, seed=gseed, symmetrize=True, wseed=wseed, C=1, M=M)
        else:
            n = int(2 ** rng.uniform(16, min(21, max_scale)))
            deg = float(rng.uniform(4, 64))
            gseed = int(rng.integers(1, 1 << 30))
            Gs = api.Graph.er(ctx, n, deg, seed=gseed, wseed=wseed, C=1, M=M)
            Gd = Gs
        deg_d = Gd.out_degree()
        sources = np.flatnonzero(deg_d > 0)
        jobs = [(Gs, api.CC, {}, 200, True)]
        for s in rng.choice(sources, size=min(3, len(sources)), replace=False):
            jobs.append((Gd, api.SSSP, {"source": int(s)}, 400, True))
        for s in rng.choice(sources, size=min(2, len(sources)), replace=False):
            delta = int(rng.choice([0, 32, 64, 128, 256]))
            jobs.append((Gd, api.SSSP, {"source": int(s), "weighted": True, "delta": delta}, 2000, True))
        if Gs is not Gd:
            deg_s = Gs.out_degree()
            srcs = np.flatnonzero(deg_s > 0)
            for s in rng.choice(srcs, size=2, replace=False):
                jobs.append((Gs, api.SSSP, {"source": int(s)}, 400, True))
            s = int(rng.choice(srcs))
            jobs.append((Gs, api.SSSP, {"source": s, "weighted": True, "delta": int(rng.choice([32, 64, 128]))},
                         2000, True))
        if kind == "rmat" and Gd.n <= (1 << pr_max_scale):
            jobs.append((Gd, api.PR, {}, 60, False))
        for G, algo, kw, mi, exact in jobs:
            S = api.State(G, algo)
            f, c = records_from(run_options(S, kw, mi), exact)
            feats += f
            costs += c
            S.free()
        g_count += 1
        if Gs is not Gd:
            Gs.free()
        Gd.free()
        if g_count % 10 == 0:
            log("graphs %d records %d (%.0fs)" % (g_count, len(feats), time.time() - t0))
    return np.array(feats, dtype=np.float32), np.array(costs, dtype=np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--budget", type=float, default=240.0, help="seconds of record collection")
    ap.add_argument("--max-scale", type=int, default=22)
    ap.add_argument("--pr-max-scale", type=int, default=20, help="largest RMAT scale of the PageRank jobs")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_bf16.bin"))
    ap.add_argument("--records", default="", help="also save the records (.npz)")
    ap.add_argument("--load-records", default="", help="train from saved records instead of collecting")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--seed", type=int, default=1, help="seed of the graph / source draw")
    ap.add_argument("--no-train", action="store_true", help="only collect (and save) records")
    args = ap.parse_args()
    ctx = api.Context(0)
    if args.load_records:
        parts = [np.load(f) for f in args.load_records.split(",")]
        feats = np.concatenate([d["feat"] for d in parts])
        costs = np.concatenate([d["cost"] for d in parts])
    else:
        feats, costs = collect(ctx, args.budget, seed=args.seed, max_scale=args.max_scale,
                               pr_max_scale=args.pr_max_scale)
        if args.records:
            np.savez_compressed(args.records, feat=feats, cost=costs)
        if args.no_train:
            log("records %d saved to %s" % (len(feats), args.records))
            return
    raw = costs
    costs = class_costs(raw)
    labels = costs.argmin(axis=1).astype(np.uint32)
    log("records %d, label histogram %s" % (len(labels), np.bincount(labels, minlength=4).tolist()))
    # hold out 10% to report the selector's regret
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(labels))
    cut = int(len(labels) * 0.9)
    tr, te = perm[:cut], perm[cut:]
    w, loss = ctx.selector_train(feats[tr], labels[tr], costs[tr], epochs=args.epochs, batch=256, lr=2e-3, seed=7)
    pred = ctx.selector_infer(feats[te]).argmax(axis=1)
    best = costs[te].min(axis=1).sum()
    chosen = costs[te][np.arange(len(te)), pred].sum()
    plain = raw[te][:, api.OPT_PUSH_SPARSE].sum()  # the offloaded push, every iteration
    log("final loss %.4f  held-out accuracy %.3f  time: chosen %.1f ms, oracle-best %.1f ms, plain %.1f ms"
        % (loss, float((pred == labels[te]).mean()), chosen, best, plain))
    np.asarray(w, dtype="<u2").tofile(args.out)
    log("wrote", args.out)


if __name__ == "__main__":
    main()
