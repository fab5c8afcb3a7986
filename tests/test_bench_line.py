"""bench.py's GTEPS arithmetic (no GPU): a PageRank step counts arcs times
iterations, a CC / SSSP step arcs times solves."""
import types

import bench


class _FakeState:
    def __init__(self, iters):
        self.iters = iters

    def init(self, **kw):
        pass

    def run(self):
        return {"iterations": self.iters, "kernel_ms": 1.0, "kernel_launches": 1, "kernel_bytes": 8,
                "edges_examined": 10}


def test_timed_leg_counts_iterations_for_pagerank(monkeypatch):
    args = types.SimpleNamespace(steps=4, warmup=1)
    clock = iter([0.0, 2.0])
    monkeypatch.setattr(bench.time, "perf_counter", lambda: next(clock))
    g, ms, it, _, _, secs = bench._timed_leg(None, _FakeState(21), args, {}, arcs=1000, per_iteration=True)
    assert secs == 2.0 and it == 21
    assert abs(g - 1000 * 21 * 4 / 2.0 / 1e9) < 1e-15
    clock = iter([0.0, 2.0])
    g, *_ = bench._timed_leg(None, _FakeState(7), args, {}, arcs=1000, sources=list(range(16)))
    assert abs(g - 1000 * 16 * 4 / 2.0 / 1e9) < 1e-15
