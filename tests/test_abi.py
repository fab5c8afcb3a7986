"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/grudon_hip.h declares, and the CLI keeps the
reference's argument handling (src/main.cpp:34-88).  No GPU calls."""
import ctypes as ct
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "grudon_hip.h")
LIB = os.environ.get("GRUDON_HIP_LIB") or os.path.join(ROOT, "grudon_amd", "libgrudon_hip.so")
CLI = os.path.join(ROOT, "grudon_amd", "bin", "grudon")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(grd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_core_entry_points():
    syms = declared_symbols()
    for s in ["grd_create", "grd_graph_from_csr", "grd_state_create", "grd_init", "grd_broadcast", "grd_traverse",
              "grd_reduce", "grd_update", "grd_run", "grd_download", "grd_selector_train", "grd_batch_pagerank"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_lists_every_symbol():
    from grudon_amd import api

    assert sorted(api.EXPORTS) == sorted(s for s in declared_symbols())


def test_last_error_is_a_string():
    lib = ct.CDLL(LIB)
    lib.grd_last_error.restype = ct.c_char_p
    assert isinstance(lib.grd_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device():
    lib = ct.CDLL(LIB)
    lib.grd_create.argtypes = [ct.c_int, ct.c_void_p]
    assert lib.grd_create(0, None) == 1  # GRD_ERR_ARG before touching the runtime
    lib.grd_graph_info.argtypes = [ct.c_void_p, ct.c_void_p]
    assert lib.grd_graph_info(None, None) == 1
    lib.grd_state_stats.argtypes = [ct.c_void_p, ct.c_void_p]
    assert lib.grd_state_stats(None, None) == 1


def _cli(*args):
    p = subprocess.run([CLI] + list(args), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    return p.returncode, p.stdout + p.stderr


def test_cli_help():
    rc, out = _cli("-h")
    assert rc == 0 and "--graph-path" in out and "-a, --algorithm" in out


@pytest.mark.parametrize("args,msg", [
    ((), "Please provide a graph path"),
    (("-g", "x.gr"), "Please provide the number of master nodes"),
    (("-g", "x.gr", "-c", "1"), "Please provide the number of mirror nodes"),
    (("-g", "x.gr", "-c", "1", "-m", "1", "--nope"), "Error parsing options"),
    (("-g", "x.gr", "-c", "one", "-m", "1"), "Error parsing options"),
    (("-g", "x.gr", "-c", "1", "-m", "1", "-a", "bfs"), "unknown algorithm"),
    (("-g", "x.gr", "-c", "1", "-m", "1", "--source", "4294967296"), "failed to parse for option --source"),
    (("-g", "x.gr", "-c", "1", "-m", "1", "--delta", "-3"), "failed to parse for option --delta"),
    (("-g", "x.gr", "-c", "1", "-m", "1", "--sched", "fixed:abc"), "fixed schedule option must be 0..4"),
    (("-g", "x.gr", "-c", "1", "-m", "1", "--sched", "fixed:9"), "fixed schedule option must be 0..4"),
    (("-g", "x.gr", "-c", "1", "-m", "1", "--value-bytes", "6"), "--value-bytes must be 4 or 8"),
])
def test_cli_argument_errors(args, msg):
    rc, out = _cli(*args)
    assert rc == 1 and msg in out
