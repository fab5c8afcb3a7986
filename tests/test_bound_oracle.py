"""CPU tests of the bound that lets a CC / SSSP pull leave rows out, on the
oracle's per-pair exchange of the reference's offloaded supersteps (no GPU).

b(it) = the smallest value broadcast in superstep `it` plus the smallest arc
increment (0 for CC labels, 1 for hop counts, the smallest arc weight for
weighted SSSP).  The pull leaves out rows whose value is at or below b, and
rows whose hot partial reached b; that is exact when
  - every value a reduce carries in superstep `it` (a candidate that improved
    a memory proxy) is at least b(it), and
  - the smallest broadcast value never decreases from one superstep to the
    next, so a row left out once stays left out.
Both are checked here on every superstep of CC, hop, weighted and
delta-stepping SSSP runs, with one and with several compute partitions.
"""
import numpy as np
import pytest

import oracle_lib as O


def _supersteps(P, algo, C, M, kw):
    """(smallest broadcast value, smallest reduce value) of every superstep
    with a non-empty broadcast."""
    if algo == O.ORC_CC:
        _, st = P.cc()
    else:
        _, st = P.sssp(kw.get("source", 0), weighted=kw.get("weighted", False), delta=kw.get("delta", 0))
    out = []
    for it in range(int(st[0])):
        bmin, rmin = None, None
        for c in range(C):
            for m in range(M):
                _, bv = P.pair_exchange(algo, it, 0, c, m, **kw)
                _, rv = P.pair_exchange(algo, it, 1, c, m, **kw)
                if bv.size:
                    bmin = int(bv.min()) if bmin is None else min(bmin, int(bv.min()))
                if rv.size:
                    rmin = int(rv.min()) if rmin is None else min(rmin, int(rv.min()))
        if bmin is not None:
            out.append((bmin, rmin))
    return out


CASES = [
    ("cc", {}, True),
    ("hop", {"source": 0}, False),
    ("weighted", {"source": 0, "weighted": True}, True),
    ("delta", {"source": 0, "weighted": True, "delta": 32}, True),
]


@pytest.mark.parametrize("C,M", [(1, 2), (2, 3)])
@pytest.mark.parametrize("name,kw,sym", CASES)
def test_reduce_values_at_or_above_the_bound(name, kw, sym, C, M):
    g = O.rmat(10, 8, seed=17, symmetrize=sym)
    if kw.get("weighted"):
        g = g.with_weights(5)
    algo = O.ORC_CC if name == "cc" else O.ORC_SSSP
    inc = 0 if name == "cc" else (int(g.w.min()) if kw.get("weighted") else 1)
    P = O.Partition(g, C, M)
    steps = _supersteps(P, algo, C, M, kw)
    assert len(steps) >= 3
    prev = -1
    for it, (bmin, rmin) in enumerate(steps):
        bound = min(bmin + inc, 0xFFFFFFFF)
        assert rmin is None or rmin >= bound, (name, it, rmin, bound)
        assert bmin >= prev, (name, it, bmin, prev)
        prev = bmin


def test_the_bound_is_tight_somewhere():
    """The check above could pass with a bound that is too low; in a hop
    count from a single source the bound is met exactly: superstep 0
    broadcasts the source (0) and its out-neighbours receive 1."""
    g = O.from_edges(4, [(0, 1), (0, 2), (1, 3)])
    P = O.Partition(g, 1, 1)
    steps = _supersteps(P, O.ORC_SSSP, 1, 1, {"source": 0})
    assert steps[0] == (0, 1)
    assert steps[1] == (1, 2)
