"""Input formats against hand-built byte fixtures (tests/golden/formats,
written by make_format_fixtures.py from the Galois .gr layout and the
MatrixMarket text format, not by the library's writer).  CPU only:
grd_graph_read makes no device call."""
import ctypes as ct
import json
import os

import numpy as np
import pytest

from grudon_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FMT = os.path.join(ROOT, "tests", "golden", "formats")
SPEC = json.load(open(os.path.join(FMT, "expected.json")))
GRD_ERR_IO = 5


def read(path):
    L = api.lib()
    n, m, hw = ct.c_uint32(), ct.c_uint64(), ct.c_int()
    rc = L.grd_graph_read(path.encode(), ct.byref(n), ct.byref(m), None, None, None, ct.byref(hw))
    if rc:
        return rc, None
    rp = np.zeros(n.value + 1, dtype=np.uint32)
    col = np.zeros(max(m.value, 1), dtype=np.uint32)
    w = np.zeros(max(m.value, 1), dtype=np.uint32)
    rc = L.grd_graph_read(path.encode(), ct.byref(n), ct.byref(m), rp.ctypes.data, col.ctypes.data, w.ctypes.data,
                          ct.byref(hw))
    assert rc == 0
    return 0, (n.value, rp, col[: m.value], w[: m.value] if hw.value else None)


@pytest.mark.parametrize("name", sorted(SPEC["good"]))
def test_fixture_reads_to_expected_csr(name):
    rc, g = read(os.path.join(FMT, name))
    assert rc == 0, api.lib().grd_last_error()
    n, rp, col, w = g
    want = SPEC["good"][name]
    assert n == want["n"]
    assert rp.tolist() == want["rp"]
    assert col.tolist() == want["col"]
    if want["w"] is None:
        assert w is None
    else:
        assert w is not None and w.tolist() == want["w"]


@pytest.mark.parametrize("name", SPEC["bad"])
def test_bad_fixture_is_refused(name):
    rc, _ = read(os.path.join(FMT, name))
    assert rc == GRD_ERR_IO, (rc, api.lib().grd_last_error())


def test_fixtures_are_current(tmp_path):
    """The committed fixtures are what the generator writes."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("mff", os.path.join(ROOT, "tests", "golden",
                                                                       "make_format_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    for name, data in mod.FILES.items():
        assert open(os.path.join(FMT, name), "rb").read() == data, name


def test_library_writer_round_trips_fixture(tmp_path):
    """grd_graph_write_gr of a fixture's CSR reads back to the same CSR, and
    its bytes equal the hand-built version-1 file (same layout)."""
    rc, (n, rp, col, w) = read(os.path.join(FMT, "v1_odd_w.gr"))
    out = str(tmp_path / "rt.gr")
    L = api.lib()
    assert L.grd_graph_write_gr(out.encode(), n, rp.ctypes.data, col.ctypes.data, w.ctypes.data) == 0
    assert open(out, "rb").read() == open(os.path.join(FMT, "v1_odd_w.gr"), "rb").read()
