"""The partitioned layout built by device kernels equals the host build, array
for array (grd_graph_digest), for directed, symmetric, weighted and
partition-file inputs and several (C, M); invalid inputs are refused."""
import numpy as np
import pytest

import oracle_lib as O
from grudon_amd import api

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _both(ctx, g, C, M, mirror=None):
    out = []
    for on_device in (True, False):
        ctx.set_layout(on_device)
        G = api.Graph.from_csr(ctx, g.rp, g.col, g.w, C=C, M=M, mirror=mirror)
        out.append((G.digest(), G))
    ctx.set_layout(True)
    return out


GRAPHS = {
    "rmat12": lambda: O.rmat(12, 16, seed=31),
    "rmat12_sym_w": lambda: O.rmat(12, 8, seed=32, symmetrize=True).with_weights(4),
    "er": lambda: O.er(3000, 9.0, seed=33),
    "sinks": lambda: O.from_edges(9, [(0, 1), (0, 2), (2, 3), (5, 1), (5, 8), (8, 0)]),
}


@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("C,M", [(1, 1), (1, 4), (2, 3), (3, 2)])
def test_device_layout_equals_host(ctx, name, C, M):
    _say("layout", name, C, M)
    g = GRAPHS[name]()
    (dd, Gd), (dh, Gh) = _both(ctx, g, C, M)
    assert np.array_equal(dd, dh), [i for i in range(16) if dd[i] != dh[i]]
    # and it runs: CC labels equal the oracle's
    if name != "rmat12":
        S = api.State(Gd, api.CC)
        S.init(sched=api.SCHED_HEURISTIC)
        S.run()
        ref, _ = O.Partition(g, C, M).cc()
        assert np.array_equal(S.result(), ref)
        S.free()


def test_device_layout_partition_file(ctx):
    _say("layout mirror file")
    g = O.rmat(11, 8, seed=34)
    mirror = (np.arange(g.n, dtype=np.uint32) * 7 + 3) % 3
    (dd, Gd), (dh, Gh) = _both(ctx, g, 2, 3, mirror=mirror)
    assert np.array_equal(dd, dh)
    for c in range(2):
        for m in range(3):
            assert np.array_equal(Gd.table(c, m), O.Partition(g, 2, 3, mirror=mirror).query(1, c, m))


def test_device_layout_rejects_bad_input(ctx):
    _say("layout errors")
    g = O.rmat(9, 4, seed=35)
    bad = g.col.copy()
    bad[5] = g.n + 7
    with pytest.raises(api.GrdError, match="column"):
        api.Graph.from_csr(ctx, g.rp, bad, C=1, M=2)
    with pytest.raises(api.GrdError, match="mirror"):
        api.Graph.from_csr(ctx, g.rp, g.col, C=1, M=2, mirror=np.full(g.n, 2, dtype=np.uint32))
    rp = g.rp.copy()
    rp[3], rp[4] = rp[4], rp[3] + 0
    if rp[3] > rp[4]:
        with pytest.raises(api.GrdError, match="monotone"):
            api.Graph.from_csr(ctx, rp, g.col, C=1, M=2)
