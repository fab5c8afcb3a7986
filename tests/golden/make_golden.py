"""Writes tests/golden/tiny_graphs.json: small hand graphs with expected
results computed from textbook definitions, independently of the oracle.

The reference ships no golden vectors or fixtures for this path (its only
test is a template stub), so the fixtures pin the oracle to the mathematics:
  cc_label    min gid over the backward-reachable set (CC push semantics,
              src/graph_algorithms/cc.cpp); on symmetric input the min gid of
              the connected component
  cc_iters    synchronous (Jacobi) min-label rounds that change something,
              plus the final superstep with an empty frontier
  hops        BFS distances from vertex 0 (sssp.cpp:69 adds 1 per edge)
  hop_iters   eccentricity of vertex 0 + 1
  dijkstra    weighted distances from vertex 0 (scipy)
  pr_pull     networkx.pagerank, alpha 0.85, uniform dangling mass
  pr_delta    fixed point x = b + 0.85 * A^T D^-1 x of the reference's delta
              PageRank (b = 0.15 for vertices with out-edges, 0 for sinks),
              solved with numpy.linalg.solve

Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0xFFFFFFFF

GRAPHS = {
    # name: (n, edge list)
    "path4": (4, [(0, 1), (1, 2), (2, 3)]),
    "star_out": (6, [(0, i) for i in range(1, 6)]),
    "star_in": (6, [(i, 0) for i in range(1, 6)]),
    "cycle5": (5, [(i, (i + 1) % 5) for i in range(5)]),
    "two_components_sym": (7, [(0, 1), (1, 0), (1, 2), (2, 1), (4, 5), (5, 4), (5, 6), (6, 5), (4, 6), (6, 4)]),
    "sink_heavy": (8, [(0, 1), (0, 2), (0, 3), (1, 4), (2, 4), (3, 5), (6, 0), (7, 6)]),
    "diamond_sym": (5, [(0, 1), (1, 0), (0, 2), (2, 0), (1, 3), (3, 1), (2, 3), (3, 2), (3, 4), (4, 3)]),
    "reverse_path": (5, [(4, 3), (3, 2), (2, 1), (1, 0)]),
}


def csr(n, edges):
    e = sorted(set((int(a), int(b)) for a, b in edges if a != b))
    rp = [0] * (n + 1)
    for a, _ in e:
        rp[a + 1] += 1
    for i in range(n):
        rp[i + 1] += rp[i]
    return rp, [b for _, b in e], e


def weight(u, v):
    # small deterministic weights for the fixtures (symmetric in u, v)
    a, b = min(u, v), max(u, v)
    return 1 + (a * 7 + b * 13) % 9


def cc_backward_min(n, e):
    preds = [[] for _ in range(n)]
    for a, b in e:
        preds[b].append(a)
    out = []
    for v in range(n):
        seen, stack = {v}, [v]
        while stack:
            x = stack.pop()
            for p in preds[x]:
                if p not in seen:
                    seen.add(p)
                    stack.append(p)
        out.append(min(seen))
    return out


def cc_jacobi_iters(n, e):
    lab = list(range(n))
    rounds = 0
    while True:
        new = lab[:]
        for a, b in e:
            new[b] = min(new[b], lab[a])
        if new == lab:
            return rounds + 1
        lab = new
        rounds += 1


def bfs(n, e, s):
    adj = [[] for _ in range(n)]
    for a, b in e:
        adj[a].append(b)
    d = [INF] * n
    d[s] = 0
    q = [s]
    for x in q:
        for y in adj[x]:
            if d[y] == INF:
                d[y] = d[x] + 1
                q.append(y)
    return d


def main():
    import networkx as nx
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra

    out = {}
    for name, (n, edges) in GRAPHS.items():
        rp, col, e = csr(n, edges)
        w = [weight(a, b) for a, b in e]
        hops = bfs(n, e, 0)
        finite = [h for h in hops if h != INF]
        A = csr_matrix((np.array(w, dtype=float), np.array(col), np.array(rp)), shape=(n, n))
        dj = dijkstra(A, indices=0)
        dj = [INF if np.isinf(x) else int(x) for x in dj]
        G = nx.DiGraph()
        G.add_nodes_from(range(n))
        G.add_edges_from(e)
        pr = nx.pagerank(G, alpha=0.85, tol=1e-14, max_iter=10000)
        deg = np.diff(np.array(rp))
        M = np.eye(n)
        for a, b in e:
            M[b, a] -= 0.85 / deg[a]
        bvec = np.where(deg > 0, 0.15, 0.0)
        x = np.linalg.solve(M, bvec)
        out[name] = {
            "n": n, "row_ptr": rp, "col": col, "weights": w,
            "cc_label": cc_backward_min(n, e), "cc_iters": cc_jacobi_iters(n, e),
            "hops": hops, "hop_iters": max(finite) + 1,
            "dijkstra": dj,
            "pr_pull": [pr[i] for i in range(n)],
            "pr_delta": [float(v) for v in x],
        }
    with open(os.path.join(HERE, "tiny_graphs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", len(out), "graphs")


if __name__ == "__main__":
    main()
