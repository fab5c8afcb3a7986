"""Writes the graph-file fixtures under tests/golden/formats/ byte by byte.

The files follow the published Galois binary graph layout (the reference's
input, read by galois::graphs::readGraph at src/DistributedGraph.cpp:49):

    u64 version (1 or 2), u64 sizeof(edge data), u64 nodes, u64 edges,
    u64 out-index[nodes]     (the end of each node's arcs),
    destinations[edges]      (u32 in version 1, u64 in version 2),
    version 1 only: one u32 of padding when edges is odd,
    edge data[edges]         (u32 when sizeof(edge data) is 4),

and the MatrixMarket coordinate text format (1-based "row col [value]"
lines after a size line; '%' starts a comment; "symmetric" stores each
off-diagonal entry once).  The writer below shares no code with the
library's own .gr writer.  The expected CSR of each good file is written by
hand in EXPECTED (row order of the file kept, duplicates and self-loops
kept), and the bad files must be refused.

    python tests/golden/make_format_fixtures.py
"""
import json
import os
import struct

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "formats")


def gr(version, esz, rows, weights=None, pad=True, ends=None, dst_override=None):
    """rows: list of destination lists; weights: parallel lists (esz 4)."""
    n = len(rows)
    dsts = [d for r in rows for d in r]
    if dst_override is not None:
        dsts = dst_override
    ne = len(dsts)
    if ends is None:
        ends, e = [], 0
        for r in rows:
            e += len(r)
            ends.append(e)
    b = struct.pack("<4Q", version, esz, n, ne)
    b += struct.pack("<%dQ" % n, *ends)
    if version == 1:
        b += struct.pack("<%dI" % ne, *dsts)
        if ne % 2 and pad:
            b += struct.pack("<I", 0)
    else:
        b += struct.pack("<%dQ" % ne, *dsts)
    if esz == 4:
        b += struct.pack("<%dI" % ne, *[x for r in weights for x in r])
    elif esz == 8:
        b += struct.pack("<%dQ" % ne, *range(ne))
    return b


FILES = {
    # version 1, u32 edge data, 7 arcs (odd: padded), a self-loop and a
    # duplicate arc kept as in the file
    "v1_odd_w.gr": gr(1, 4, [[1, 4], [1], [], [0, 2, 0], [3]], [[10, 20], [5], [], [7, 8, 9], [1]]),
    # version 1, no edge data, 3 arcs (padding present)
    "v1_nodata.gr": gr(1, 0, [[1, 2], [], [3], []]),
    # version 2: u64 destinations, u32 edge data, no padding
    "v2_u64_w.gr": gr(2, 4, [[5], [0, 2], [], [4], [1, 3], [0]], [[3], [4, 6], [], [2], [9, 1], [255]]),
    # refused: a version-2 destination of 2^32 + 2 (would alias vertex 2 if
    # truncated to 32 bits)
    "bad_v2_big_dst.gr": gr(2, 0, [[1], [2], []], dst_override=[1, (1 << 32) + 2]),
    # refused: 8-byte edge data
    "bad_esz8.gr": gr(1, 8, [[1], []]),
    # refused: the last out-index (2) is short of the 3 arcs in the header
    "bad_short_end.gr": gr(1, 0, [[1], [2], [0]], ends=[1, 2, 2]),
    # refused: a destination beyond the nodes
    "bad_dst_range.gr": gr(1, 0, [[3], [], []]),
    # refused: the file ends inside the destinations
    "bad_truncated.gr": gr(1, 0, [[1, 2], [0], [1]])[:-6],
    "general.mtx": (
        "%%MatrixMarket matrix coordinate integer general\n"
        "% a comment line\n"
        "%\n"
        "4 4 5\n"
        "1 2 7\n"
        "3 1 2\n"
        "% a comment between entries\n"
        "1 4 3\n"
        "4 4 1\n"
        "2 3 9\n").encode(),
    "symmetric_pattern.mtx": (
        "%%MatrixMarket matrix coordinate pattern symmetric\n"
        "%\n"
        "3 3 3\n"
        "2 1\n"
        "3 1\n"
        "3 3\n").encode(),
    # refused: an entry outside the matrix
    "bad_range.mtx": (
        "%%MatrixMarket matrix coordinate pattern general\n"
        "3 3 2\n"
        "1 2\n"
        "4 1\n").encode(),
}

EXPECTED = {
    "v1_odd_w.gr": {"n": 5, "rp": [0, 2, 3, 3, 6, 7], "col": [1, 4, 1, 0, 2, 0, 3], "w": [10, 20, 5, 7, 8, 9, 1]},
    "v1_nodata.gr": {"n": 4, "rp": [0, 2, 2, 3, 3], "col": [1, 2, 3], "w": None},
    "v2_u64_w.gr": {"n": 6, "rp": [0, 1, 3, 3, 4, 6, 7], "col": [5, 0, 2, 4, 1, 3, 0], "w": [3, 4, 6, 2, 9, 1, 255]},
    # (0,1,7) (2,0,2) (0,3,3) (3,3,1) (1,2,9) ordered by source, file order
    # within a source
    "general.mtx": {"n": 4, "rp": [0, 2, 3, 4, 5], "col": [1, 3, 2, 0, 3], "w": [7, 3, 9, 2, 1]},
    # (1,0)+(0,1), (2,0)+(0,2), (2,2) once
    "symmetric_pattern.mtx": {"n": 3, "rp": [0, 2, 3, 5], "col": [1, 2, 0, 0, 2], "w": None},
}
BAD = ["bad_v2_big_dst.gr", "bad_esz8.gr", "bad_short_end.gr", "bad_dst_range.gr", "bad_truncated.gr",
       "bad_range.mtx"]


def main():
    os.makedirs(HERE, exist_ok=True)
    for name, data in FILES.items():
        with open(os.path.join(HERE, name), "wb") as f:
            f.write(data)
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump({"good": EXPECTED, "bad": BAD}, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
