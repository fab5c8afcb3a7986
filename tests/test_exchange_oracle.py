"""CPU tests that pin the oracle's per-pair exchange capture: the broadcast
and reduce bitvectors with their packed values (no GPU).

Pins: a path graph worked out by hand; the first CC broadcast, which must
carry every vertex to its mirror with its own gid as the label; and the
capture's sizes, which must add up to the run's exchanged-byte count.
"""
import numpy as np
import pytest

import oracle_lib as O


def test_path_exchange_by_hand():
    # SSSP hops from 0 on 0 -> 1 -> 2, C = M = 1
    P = O.Partition(O.from_edges(3, [(0, 1), (1, 2)]), 1, 1)
    want = {(0, 0): ([0], [0]), (0, 1): ([1], [1]), (1, 0): ([1], [1]), (1, 1): ([2], [2]),
            (2, 0): ([2], [2]), (2, 1): ([], []), (3, 0): ([], [])}
    for (it, d), (gids, vals) in want.items():
        g, v = P.pair_exchange(O.ORC_SSSP, it, d, 0, 0, source=0)
        assert g.tolist() == gids and v.tolist() == vals, (it, d)


@pytest.mark.parametrize("C,M", [(1, 3), (2, 2), (3, 2)])
def test_first_cc_broadcast_carries_every_label(C, M):
    g = O.rmat(9, 8, seed=4, symmetrize=True)
    P = O.Partition(g, C, M)
    mirror = np.arange(g.n) % M
    for c in range(C):
        seen = []
        for m in range(M):
            gids, vals = P.pair_exchange(O.ORC_CC, 0, 0, c, m)
            assert np.array_equal(gids, vals)  # CC starts from label = gid
            assert (gids % C == c).all() and (mirror[gids] == m).all()
            seen.append(gids)
        assert np.array_equal(np.sort(np.concatenate(seen)), np.arange(c, g.n, C))


@pytest.mark.parametrize("C,M", [(1, 2), (2, 3)])
def test_capture_sizes_add_up_to_the_byte_count(C, M):
    g = O.rmat(9, 8, seed=6)
    P = O.Partition(g, C, M)
    _, st = P.sssp(0)
    words = sum(8 * ((int(P.query(0, c, m, fetch=False)) + 63) // 64) for c in range(C) for m in range(M))
    total = 0
    for it in range(int(st[0])):
        total += 2 * words
        for c in range(C):
            for m in range(M):
                for d in (0, 1):
                    total += 4 * len(P.pair_exchange(O.ORC_SSSP, it, d, c, m)[0])
    assert total == int(st[1])
