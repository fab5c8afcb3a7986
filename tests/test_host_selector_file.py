"""The CLI's selector-weight reader (grudon_amd/host/SelectorFile.hpp), built
with g++ and checked against numpy's little-endian read of the same bytes:
hand-made values whose two bytes differ, the packaged weights, and an
odd-sized file it must refuse (no GPU)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "grudon_amd", "host")
WEIGHTS = os.path.join(ROOT, "grudon_amd", "selector_bf16.bin")

PROGRAM = r"""
#include <cstdio>
#include <exception>
#include "SelectorFile.hpp"
int main(int argc, char **argv)
{
  try {
    const std::vector<uint16_t> w = read_selector_weights(argv[1]);
    std::printf("%zu\n", w.size());
    for (uint16_t x : w) std::printf("%u\n", (unsigned)x);
  } catch (const std::exception &e) {
    std::printf("error %s\n", e.what());
    return 3;
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def reader(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not on this machine")
    d = tmp_path_factory.mktemp("selfile")
    src, exe = d / "read.cpp", d / "read"
    src.write_text(PROGRAM)
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", HOST, str(src), "-o", str(exe)], check=True)
    return str(exe)


def _read(reader, path):
    p = subprocess.run([reader, str(path)], capture_output=True, text=True)
    return p.returncode, p.stdout.split()


def test_reads_little_endian_pairs(reader, tmp_path):
    vals = np.array([0x0001, 0xABCD, 0xFF00, 0x1234, 0x8000, 0x00FF], dtype="<u2")
    f = tmp_path / "w.bin"
    vals.tofile(f)
    rc, out = _read(reader, f)
    assert rc == 0
    assert int(out[0]) == len(vals)
    assert [int(x) for x in out[1:]] == vals.tolist()


def test_reads_the_packaged_weights(reader):
    rc, out = _read(reader, WEIGHTS)
    ref = np.fromfile(WEIGHTS, dtype="<u2")
    assert rc == 0
    assert int(out[0]) == ref.size == 5508
    assert np.array_equal(np.array([int(x) for x in out[1:]], dtype=np.uint16), ref)


def test_refuses_an_odd_byte_count(reader, tmp_path):
    f = tmp_path / "odd.bin"
    f.write_bytes(b"\x01\x02\x03")
    rc, out = _read(reader, f)
    assert rc == 3 and out[0] == "error"
