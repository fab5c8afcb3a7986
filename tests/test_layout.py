"""Partition layout of the oracle against hand-derived tables.

6-vertex graph, C = 1, M = 2 (mirror = gid % 2):
    0->1, 0->2, 1->3, 2->4, 3->5, 4->0, 5->1
Memory 0 owns {0, 2, 4}; its out-edges reach 1 (mirror 1) -> cross mirror.
Memory 1 owns {1, 3, 5}; its out-edges reach 0 (from 5->1? no: 5->1 is
local) ... 1->3 local, 3->5 local, 5->1 local: no cross mirrors.
Translation tables (DistributedGraph.cpp:644-660, :701-734):
    (c=0, m=0): own on 0 ascending = [0, 2, 4], then cross on 0 = [1]
    (c=0, m=1): own on 1 = [1, 3, 5], no cross
"""
import numpy as np

import oracle_lib as O

EDGES = [(0, 1), (0, 2), (1, 3), (2, 4), (3, 5), (4, 0), (5, 1)]


def test_hand_tables():
    g = O.from_edges(6, EDGES)
    P = O.Partition(g, 1, 2)
    assert P.query(1, 0, 0).tolist() == [0, 2, 4, 1]
    assert P.query(1, 0, 1).tolist() == [1, 3, 5]
    assert P.query(2, 0, 0).tolist() == [0, 2, 4, 1]
    assert P.query(2, 1, 0).tolist() == [1, 3, 5]
    # memory 0: own lids 0..2, cross lid 3; local CSR in memory lids
    assert P.query(3, 0).tolist() == [0, 2, 4, 1]
    assert P.query(4, 0, fetch=False) == 3
    assert P.query(5, 0).tolist() == [0, 2, 3, 4]       # rows of gid 0, 2, 4
    assert P.query(6, 0).tolist() == [3, 1, 2, 0]       # 0->1(x), 0->2, 2->4, 4->0


def test_partition_file_mirror():
    g = O.from_edges(6, EDGES)
    mirror = np.array([1, 1, 0, 0, 1, 0], dtype=np.uint32)
    P = O.Partition(g, 1, 2, mirror=mirror)
    # memory 0 owns {2, 3, 5}; 2->4 and 3->5 (local), 5->1 crosses -> {4, 1}
    assert P.query(3, 0).tolist() == [2, 3, 5, 1, 4]
    # M == 1 ignores the file (Partitioner.hpp:73-85)
    P1 = O.Partition(g, 1, 1, mirror=np.zeros(6, dtype=np.uint32))
    assert P1.query(3, 0).tolist() == list(range(6))


def test_tables_are_bijections():
    g = O.rmat(10, 8, seed=3)
    C, M = 2, 3
    P = O.Partition(g, C, M)
    owners = {}
    for c in range(C):
        for m in range(M):
            t = P.query(1, c, m)
            assert np.array_equal(t, P.query(2, m, c))
            assert np.all(t % C == c)
            own = t[(t % M) == m]
            assert np.all(np.diff(own) > 0)  # own part ascending
            for v in t:
                owners.setdefault(int(v), []).append(m)
    # every vertex has exactly one own entry (its mirror) and any number of cross ones
    for v, ms in owners.items():
        assert (v % M) in ms
