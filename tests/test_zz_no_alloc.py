"""The allocation-free selector inference (grd_selector_infer_noalloc)
against grd_selector_infer, through the C-ABI: bit-identical logits on two
calls in a row, the second with fewer samples, and the size limit refused."""
import ctypes as ct
import os

import numpy as np
import pytest

from grudon_amd import api

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAX_SAMPLES = 4096


def _noalloc(ctx, feat):
    L = api.lib()
    f = L.grd_selector_infer_noalloc
    f.restype = ct.c_int
    f.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_uint64, ct.c_void_p]
    x = np.ascontiguousarray(feat, dtype=np.float32)
    out = np.full((len(x), 4), np.nan, dtype=np.float32)
    rc = f(ctx.h, x.ctypes.data, len(x), out.ctypes.data)
    return rc, out


def test_noalloc_inference_equals_allocating_entry_point():
    ctx = api.Context(0)
    try:
        w = np.fromfile(os.path.join(ROOT, "grudon_amd", "selector_bf16.bin"), dtype="<u2")
        ctx.selector_load(w, min_margin=0.0)
        rng = np.random.default_rng(7)
        for n in (MAX_SAMPLES, 37):  # a full call, then a smaller one over the same scratch
            feat = rng.random((n, 16), dtype=np.float32)
            rc, got = _noalloc(ctx, feat)
            assert rc == 0, api.lib().grd_last_error()
            want = ctx.selector_infer(feat)
            assert np.array_equal(got.view(np.uint32), np.asarray(want, dtype=np.float32).view(np.uint32)), n
        rc, _ = _noalloc(ctx, rng.random((MAX_SAMPLES + 1, 16), dtype=np.float32))
        assert rc == 1 and b"at most" in api.lib().grd_last_error()
    finally:
        ctx.close()
