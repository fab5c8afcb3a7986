"""GPU parity of the exchange protocol itself: in every superstep, each
(compute, memory) pair's broadcast and reduce carry the same vertices with
the same packed values as the oracle's bitvectors and property buffers
(the device orders bits by out-degree, so the sets are compared by gid).
CC and SSSP are deterministic per superstep, so the bar is equality.  A pull
that leaves out the rows that cannot improve (the default) sends the
oracle's reduce message less those rows: a subset with equal values (the
bound only grows from superstep to superstep, so a row left out once is
left out after, and every other row sees the same running update).
"""
import numpy as np
import pytest

import oracle_lib as O
from grudon_amd import api

pytestmark = pytest.mark.gpu

OPTIONS = [api.OPT_PUSH_SPARSE, api.OPT_PULL_SPARSE, api.OPT_PUSH_DENSE, api.OPT_PULL_DENSE]


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("every_row", [True, False])
@pytest.mark.parametrize("algo", [api.CC, api.SSSP])
@pytest.mark.parametrize("C,M", [(1, 2), (2, 3)])
@pytest.mark.parametrize("option", OPTIONS)
def test_pair_exchange_equals_oracle(ctx, algo, C, M, option, every_row):
    _say("exchange", algo, C, M, option, every_row)
    g = O.rmat(10, 8, seed=13, symmetrize=(algo == api.CC))
    oalgo = O.ORC_CC if algo == api.CC else O.ORC_SSSP
    P = O.Partition(g, C, M)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=C, M=M)
    S = api.State(G, algo)
    S.init(source=0, pull_every_row=every_row)
    exact = every_row or not (option & 1)
    L = api.lib()
    it = 0
    while True:
        assert L.grd_broadcast(S.h, option) == 0
        for c in range(C):
            for m in range(M):
                got = S.pair_exchange(0, c, m)
                ref = P.pair_exchange(oalgo, it, 0, c, m)
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), ("bcast", it, c, m)
        assert L.grd_traverse(S.h, option) == 0
        assert L.grd_reduce(S.h, option) == 0
        for c in range(C):
            for m in range(M):
                got = S.pair_exchange(1, c, m)
                ref = P.pair_exchange(oalgo, it, 1, c, m)
                if exact:
                    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), ("reduce", it, c, m)
                else:
                    want = dict(zip(ref[0].tolist(), ref[1].tolist()))
                    assert all(want.get(x) == y for x, y in zip(got[0].tolist(), got[1].tolist())), \
                        ("reduce", it, c, m)
        f = S.step_update()
        it += 1
        if f == 0:
            break
    S.free()


def test_shipping_superstep_has_no_reduce(ctx):
    _say("exchange ship")
    g = O.rmat(9, 8, seed=3, symmetrize=True)
    G = api.Graph.from_csr(ctx, g.rp, g.col, C=1, M=2)
    S = api.State(G, api.CC)
    S.init()
    L = api.lib()
    assert L.grd_broadcast(S.h, api.OPT_SHIP_EDGES) == 0
    gids, vals = S.pair_exchange(0, 0, 1)
    ref = O.Partition(g, 1, 2).pair_exchange(O.ORC_CC, 0, 0, 0, 1)
    assert np.array_equal(gids, ref[0]) and np.array_equal(vals, ref[1])
    assert L.grd_traverse(S.h, api.OPT_SHIP_EDGES) == 0
    assert L.grd_reduce(S.h, api.OPT_SHIP_EDGES) == 0
    with pytest.raises(api.GrdError, match="no such exchange"):
        S.pair_exchange(1, 0, 0)
    S.free()
