"""ctypes binding of libgrudon_hip.so (C-ABI in include/grudon_hip.h).

This is the Python face of the HIP library: it raises if the library is not
built or a call fails; there is no CPU fallback.
"""
import atexit
import ctypes as ct
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# GRUDON_HIP_LIB selects another build of the same library (the sanitizer
# build of grudon_amd/build.py --sanitize runs the CPU tests against one)
LIB_PATH = os.environ.get("GRUDON_HIP_LIB") or os.path.join(PKG, "libgrudon_hip.so")

PR, CC, SSSP, PR_PULL = 0, 1, 2, 3
OPT_PUSH_SPARSE, OPT_PULL_SPARSE, OPT_PUSH_DENSE, OPT_PULL_DENSE, OPT_SHIP_EDGES = 0, 1, 2, 3, 4
SCHED_PLAIN, SCHED_HEURISTIC, SCHED_LEARNED, SCHED_FIXED, SCHED_NDP_ENGINE = 0, 1, 2, 3, 4
RMAT_ABC = (0.57, 0.19, 0.19)
MLP_PARAMS = 5508
SELECTOR_MARGIN = 1.0  # GRD_SELECTOR_MARGIN of include/grudon_hip.h: the packaged weights' overrule margin

# every extern "C" symbol declared in include/grudon_hip.h
EXPORTS = [
    "grd_last_error", "grd_create", "grd_destroy", "grd_set_host_threads", "grd_set_layout", "grd_set_pull_tiling", "grd_graph_digest",
    "grd_generate_rmat", "grd_generate_er", "grd_edge_weights", "grd_graph_from_csr", "grd_graph_rmat",
    "grd_graph_er", "grd_graph_load", "grd_graph_read", "grd_graph_write_gr", "grd_graph_free", "grd_graph_info",
    "grd_graph_table", "grd_graph_out_degree", "grd_graph_coverage",
    "grd_state_create", "grd_state_create_typed", "grd_state_free", "grd_init", "grd_broadcast", "grd_traverse", "grd_reduce",
    "grd_update", "grd_aggregate_one", "grd_run", "grd_select_option", "grd_state_stats", "grd_download",
    "grd_features", "grd_snapshot", "grd_pair_exchange",
    "grd_restore", "grd_selector_load", "grd_selector_train", "grd_selector_infer", "grd_selector_infer_noalloc",
    "grd_selector_grad",
    "grd_batch_create", "grd_batch_free", "grd_batch_tune", "grd_batch_autotune", "grd_batch_pagerank",
    "grd_batch_cc", "grd_batch_download",
]


class Params(ct.Structure):
    _fields_ = [("source", ct.c_uint32), ("weighted", ct.c_uint32), ("delta", ct.c_uint32),
                ("max_iter", ct.c_uint32), ("damping", ct.c_double), ("tolerance", ct.c_double),
                ("sched", ct.c_uint32), ("fixed_option", ct.c_uint32), ("time_iters", ct.c_uint32),
                ("time_kernel", ct.c_uint32), ("offload_only", ct.c_uint32), ("pull_every_row", ct.c_uint32)]


class IterRec(ct.Structure):
    _fields_ = [("iteration", ct.c_uint32), ("option", ct.c_uint32), ("frontier", ct.c_uint64),
                ("frontier_edges", ct.c_uint64), ("dirty", ct.c_uint64), ("next_frontier", ct.c_uint64),
                ("ms", ct.c_double), ("feat", ct.c_float * 16)]


class RunStats(ct.Structure):
    _fields_ = [("iterations", ct.c_uint64), ("bytes_exchanged", ct.c_uint64), ("edges_examined", ct.c_uint64),
                ("hit_max_iter", ct.c_uint64), ("ms", ct.c_double), ("residual", ct.c_double),
                ("option_counts", ct.c_uint64 * 5), ("fallbacks", ct.c_uint64), ("kernel_ms", ct.c_double),
                ("kernel_launches", ct.c_uint64), ("kernel_bytes", ct.c_uint64)]

    def as_dict(self):
        return {"iterations": int(self.iterations), "bytes_exchanged": int(self.bytes_exchanged),
                "edges_examined": int(self.edges_examined), "hit_max_iter": int(self.hit_max_iter),
                "ms": float(self.ms), "residual": float(self.residual),
                "option_counts": [int(x) for x in self.option_counts], "fallbacks": int(self.fallbacks),
                "kernel_ms": float(self.kernel_ms), "kernel_launches": int(self.kernel_launches),
                "kernel_bytes": int(self.kernel_bytes)}


_lib = None


class GrdError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GrdError("libgrudon_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ct.CDLL(LIB_PATH)
        L.grd_last_error.restype = ct.c_char_p
        vp = ct.c_void_p
        pp = ct.POINTER(ct.c_void_p)
        sigs = {
            "grd_create": [ct.c_int, pp], "grd_destroy": [vp], "grd_set_host_threads": [vp, ct.c_int],
            "grd_set_layout": [vp, ct.c_int], "grd_graph_digest": [vp, vp],
            "grd_set_pull_tiling": [vp, ct.c_int, ct.c_uint32, ct.c_uint32],
            "grd_generate_rmat": [vp, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_uint64,
                                  ct.c_int, vp, vp, ct.c_uint64, ct.POINTER(ct.c_uint64)],
            "grd_generate_er": [vp, ct.c_uint32, ct.c_double, ct.c_uint64, vp, vp, ct.c_uint64,
                                ct.POINTER(ct.c_uint64)],
            "grd_graph_from_csr": [vp, ct.c_uint32, vp, vp, vp, ct.c_int, ct.c_int, vp, pp],
            "grd_graph_rmat": [vp, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_uint64, ct.c_int,
                               ct.c_uint64, ct.c_int, ct.c_int, pp],
            "grd_graph_er": [vp, ct.c_uint32, ct.c_double, ct.c_uint64, ct.c_uint64, ct.c_int, ct.c_int, pp],
            "grd_graph_load": [vp, ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int, pp],
            "grd_graph_free": [vp], "grd_graph_info": [vp, vp], "grd_graph_table": [vp, ct.c_int, ct.c_int, vp,
                                                                                   ct.POINTER(ct.c_uint64)],
            "grd_graph_out_degree": [vp, vp], "grd_graph_coverage": [vp, vp, ct.POINTER(ct.c_uint64)],
            "grd_edge_weights": [vp, ct.c_uint32, vp, vp, ct.c_uint64, vp],
            "grd_graph_read": [ct.c_char_p, ct.POINTER(ct.c_uint32), ct.POINTER(ct.c_uint64), vp, vp, vp,
                               ct.POINTER(ct.c_int)],
            "grd_graph_write_gr": [ct.c_char_p, ct.c_uint32, vp, vp, vp],
            "grd_select_option": [vp, ct.POINTER(ct.c_int)], "grd_state_stats": [vp, ct.POINTER(RunStats)],
            "grd_state_create": [vp, ct.c_int, pp], "grd_state_create_typed": [vp, ct.c_int, ct.c_uint32, pp],
            "grd_state_free": [vp], "grd_init": [vp, ct.POINTER(Params)],
            "grd_broadcast": [vp, ct.c_int], "grd_traverse": [vp, ct.c_int], "grd_reduce": [vp, ct.c_int],
            "grd_pair_exchange": [vp, ct.c_int, ct.c_int, ct.c_int, vp, vp, ct.POINTER(ct.c_uint64)],
            "grd_update": [vp, ct.POINTER(ct.c_uint64)], "grd_aggregate_one": [vp, ct.c_int, ct.c_uint32, vp],
            "grd_run": [vp, ct.POINTER(RunStats), vp, ct.c_uint32], "grd_download": [vp, vp],
            "grd_features": [vp, vp], "grd_snapshot": [vp], "grd_restore": [vp],
            "grd_selector_load": [vp, vp, ct.c_uint64, ct.c_float],
            "grd_selector_train": [vp, vp, vp, vp, ct.c_uint64, ct.c_uint32, ct.c_uint32, ct.c_float, ct.c_uint64,
                                   vp, ct.POINTER(ct.c_float)],
            "grd_selector_infer": [vp, vp, ct.c_uint64, vp],
            "grd_selector_infer_noalloc": [vp, vp, ct.c_uint64, vp],
            "grd_selector_grad": [vp, vp, vp, vp, vp, ct.c_uint64, ct.POINTER(ct.c_float), vp],
            "grd_batch_create": [vp, ct.c_uint32, vp, vp, vp, vp, pp], "grd_batch_free": [vp],
            "grd_batch_tune": [vp, ct.c_uint32, ct.c_uint32, ct.c_uint32], "grd_batch_autotune": [vp, vp],
            "grd_batch_pagerank": [vp, ct.c_double, ct.c_double, ct.c_uint32, vp, ct.POINTER(ct.c_double)],
            "grd_batch_cc": [vp, ct.c_uint32, vp, ct.POINTER(ct.c_double)],
            "grd_batch_download": [vp, ct.c_uint32, ct.c_int, vp],
        }
        for name, args in sigs.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = ct.c_int
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise GrdError("%s failed (status %d): %s" % (what, rc, lib().grd_last_error().decode()))


def _ptr(a):
    return None if a is None else a.ctypes.data


_LIVE_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    """Frees every context still open when the interpreter exits (its states,
    graphs and batches first), while the library and this module are intact,
    rather than from finalizers run during module teardown."""
    for c in list(_LIVE_CONTEXTS):
        try:
            c.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        h = ct.c_void_p()
        _check(lib().grd_create(device, ct.byref(h)), "grd_create")
        self.h = h
        self.device = device
        self._deps = []  # weak references to the graphs, states and batches made on it
        _LIVE_CONTEXTS.add(self)

    def _track(self, obj):
        self._deps.append(weakref.ref(obj))

    def close(self):
        """Frees what is still alive on the context (newest first: states
        before their graphs), then the context."""
        if self.h:
            for ref in reversed(self._deps):
                obj = ref()
                if obj is not None:
                    obj.free()
            self._deps = []
            lib().grd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_host_threads(self, n):
        _check(lib().grd_set_host_threads(self.h, n), "grd_set_host_threads")

    def set_pull_tiling(self, mode=3, tier_bits=0, tile_arcs=0):
        """How pull PageRank states created afterwards pull: 0 the row-block
        pull, 1 LDS tiles (tiers of 2^tier_bits sources, tiles of about
        tile_arcs arcs), 2 L2 segments (tiers of 2^tier_bits), 3 (the
        default) the hot / cold split (each partition's 2^tier_bits hottest
        sources from an LDS window, default 2^15; tile_arcs must be 0)."""
        _check(lib().grd_set_pull_tiling(self.h, int(mode), tier_bits, tile_arcs), "grd_set_pull_tiling")

    def set_layout(self, on_device):
        """Build partitioned layouts with device kernels (True) or host threads."""
        _check(lib().grd_set_layout(self.h, int(bool(on_device))), "grd_set_layout")

    def generate_rmat(self, scale, ef=16, seed=1, symmetrize=False, abc=RMAT_ABC):
        n = 1 << scale
        cap = (ef << scale) * (2 if symmetrize else 1)
        rp = np.zeros(n + 1, dtype=np.uint32)
        col = np.zeros(max(cap, 1), dtype=np.uint32)
        m = ct.c_uint64()
        _check(lib().grd_generate_rmat(self.h, scale, ef, abc[0], abc[1], abc[2], seed, int(symmetrize),
                                       _ptr(rp), _ptr(col), cap, ct.byref(m)), "grd_generate_rmat")
        return rp, col[: m.value].copy()

    def generate_er(self, n, mean_degree, seed):
        cap = int(n * mean_degree / 2.0) * 2
        rp = np.zeros(n + 1, dtype=np.uint32)
        col = np.zeros(max(cap, 1), dtype=np.uint32)
        m = ct.c_uint64()
        _check(lib().grd_generate_er(self.h, n, mean_degree, seed, _ptr(rp), _ptr(col), cap, ct.byref(m)),
               "grd_generate_er")
        return rp, col[: m.value].copy()

    def edge_weights(self, rp, col, wseed):
        """Seeded arc weights in [1, 255], equal for (u, v) and (v, u)."""
        rp = np.ascontiguousarray(rp, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        w = np.zeros(max(len(col), 1), dtype=np.uint32)
        _check(lib().grd_edge_weights(self.h, len(rp) - 1, _ptr(rp), _ptr(col), wseed, _ptr(w)), "grd_edge_weights")
        return w[: len(col)]


    # selector
    def selector_load(self, w_bf16, min_margin=0.0):
        w = np.ascontiguousarray(w_bf16, dtype=np.uint16)
        _check(lib().grd_selector_load(self.h, _ptr(w), w.size, min_margin), "grd_selector_load")

    def selector_train(self, feat, label, cost=None, epochs=20, batch=256, lr=3e-3, seed=1):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        lb = np.ascontiguousarray(label, dtype=np.uint32)
        c = None if cost is None else np.ascontiguousarray(cost, dtype=np.float32)
        out = np.zeros(MLP_PARAMS, dtype=np.uint16)
        loss = ct.c_float()
        _check(lib().grd_selector_train(self.h, _ptr(f), _ptr(lb), _ptr(c), lb.size, epochs, batch, lr, seed,
                                        _ptr(out), ct.byref(loss)), "grd_selector_train")
        return out, loss.value

    def selector_infer(self, feat):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        n = f.shape[0]
        out = np.zeros((n, 4), dtype=np.float32)
        _check(lib().grd_selector_infer(self.h, _ptr(f), n, _ptr(out)), "grd_selector_infer")
        return out

    def selector_grad(self, w, feat, label, sample_weight=None):
        w = np.ascontiguousarray(w, dtype=np.float32)
        f = np.ascontiguousarray(feat, dtype=np.float32)
        lb = np.ascontiguousarray(label, dtype=np.uint32)
        sw = None if sample_weight is None else np.ascontiguousarray(sample_weight, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ct.c_float()
        _check(lib().grd_selector_grad(self.h, _ptr(w), _ptr(f), _ptr(lb), _ptr(sw), lb.size, ct.byref(loss),
                                       _ptr(g)), "grd_selector_grad")
        return loss.value, g


class Graph:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        self._states = []  # weak references: a graph frees its states first
        ctx._track(self)
        info = np.zeros(16, dtype=np.uint64)
        _check(lib().grd_graph_info(self.h, _ptr(info)), "grd_graph_info")
        self.n, self.m, self.C, self.M = int(info[0]), int(info[1]), int(info[2]), int(info[3])
        self.slots, self.proxies, self.cross = int(info[4]), int(info[5]), int(info[6])
        self.weighted = bool(info[7])
        self.device_bytes = int(info[8])

    def info(self):
        """The 16 words of grd_graph_info (the tiled pull's sizes in 11-14)."""
        out = np.zeros(16, dtype=np.uint64)
        _check(lib().grd_graph_info(self.h, _ptr(out)), "grd_graph_info")
        return [int(x) for x in out]

    @classmethod
    def from_csr(cls, ctx, rp, col, w=None, C=1, M=1, mirror=None):
        rp = np.ascontiguousarray(rp, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        mirror = None if mirror is None else np.ascontiguousarray(mirror, dtype=np.uint32)
        h = ct.c_void_p()
        _check(lib().grd_graph_from_csr(ctx.h, len(rp) - 1, _ptr(rp), _ptr(col), _ptr(w), C, M, _ptr(mirror),
                                        ct.byref(h)), "grd_graph_from_csr")
        return cls(ctx, h)

    @classmethod
    def rmat(cls, ctx, scale, ef=16, seed=1, symmetrize=False, wseed=0, C=1, M=1, abc=RMAT_ABC):
        h = ct.c_void_p()
        _check(lib().grd_graph_rmat(ctx.h, scale, ef, abc[0], abc[1], abc[2], seed, int(symmetrize), wseed, C, M,
                                    ct.byref(h)), "grd_graph_rmat")
        return cls(ctx, h)

    @classmethod
    def er(cls, ctx, n, mean_degree, seed, wseed=0, C=1, M=1):
        h = ct.c_void_p()
        _check(lib().grd_graph_er(ctx.h, n, mean_degree, seed, wseed, C, M, ct.byref(h)), "grd_graph_er")
        return cls(ctx, h)

    @classmethod
    def load(cls, ctx, path, partition_file=None, C=1, M=1):
        h = ct.c_void_p()
        _check(lib().grd_graph_load(ctx.h, path.encode(), (partition_file or "").encode(), C, M, ct.byref(h)),
               "grd_graph_load")
        return cls(ctx, h)

    def table(self, c, m):
        cnt = ct.c_uint64()
        _check(lib().grd_graph_table(self.h, c, m, None, ct.byref(cnt)), "grd_graph_table")
        out = np.zeros(max(cnt.value, 1), dtype=np.uint32)
        _check(lib().grd_graph_table(self.h, c, m, _ptr(out), ct.byref(cnt)), "grd_graph_table")
        return out[: cnt.value]

    def digest(self):
        """16 digests of the partitioned layout (for comparing two builds)."""
        out = np.zeros(16, dtype=np.uint64)
        _check(lib().grd_graph_digest(self.h, _ptr(out)), "grd_graph_digest")
        return out

    def out_degree(self):
        out = np.zeros(self.n, dtype=np.uint32)
        _check(lib().grd_graph_out_degree(self.h, _ptr(out)), "grd_graph_out_degree")
        return out

    def coverage(self):
        """(covered[gid] as uint8, number of vertices not covered)."""
        out = np.zeros(self.n, dtype=np.uint8)
        cnt = ct.c_uint64()
        _check(lib().grd_graph_coverage(self.h, _ptr(out), ct.byref(cnt)), "grd_graph_coverage")
        return out, int(cnt.value)

    def free(self):
        if self.h:
            for ref in self._states:
                st = ref()
                if st is not None:
                    st.free()
            self._states = []
            lib().grd_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class State:
    def __init__(self, graph, algo, value_bytes=4):
        """value_bytes 8: the uint64 property width of CC<uint64_t> /
        SSSP<uint64_t> (results and exchanged values come back as uint64)."""
        self.graph = graph
        self.algo = algo
        self.value_bytes = value_bytes
        h = ct.c_void_p()
        _check(lib().grd_state_create_typed(graph.h, algo, value_bytes, ct.byref(h)), "grd_state_create")
        self.h = h
        graph._states.append(weakref.ref(self))
        graph.ctx._track(self)

    def _vdtype(self):
        if self.algo in (PR, PR_PULL):
            return np.float32
        return np.uint64 if self.value_bytes == 8 else np.uint32

    def aggregate_one(self, compute_partition, lid, value):
        """aggregate(lid, value) on compute partition compute_partition."""
        v = np.array([value], dtype=self._vdtype())
        _check(lib().grd_aggregate_one(self.h, compute_partition, lid, _ptr(v)), "grd_aggregate_one")

    def init(self, source=0, weighted=False, delta=0, max_iter=1000, damping=0.85, tolerance=1e-6,
             sched=SCHED_PLAIN, fixed_option=0, time_iters=False, time_kernel=False, offload_only=False,
             pull_every_row=False):
        p = Params(source, int(weighted), delta, max_iter, damping, tolerance, sched, fixed_option,
                   int(time_iters), int(time_kernel), int(offload_only), int(pull_every_row))
        _check(lib().grd_init(self.h, ct.byref(p)), "grd_init")

    def run(self, trace_cap=0):
        st = RunStats()
        tr = (IterRec * trace_cap)() if trace_cap else None
        _check(lib().grd_run(self.h, ct.byref(st), tr, trace_cap), "grd_run")
        if trace_cap:
            n = min(int(st.iterations), trace_cap)
            return st.as_dict(), [tr[i] for i in range(n)]
        return st.as_dict()

    def step_update(self):
        """grd_update alone (after grd_broadcast / grd_traverse / grd_reduce);
        returns the next frontier size."""
        f = ct.c_uint64()
        _check(lib().grd_update(self.h, ct.byref(f)), "grd_update")
        return f.value

    def step(self, option):
        _check(lib().grd_broadcast(self.h, option), "grd_broadcast")
        _check(lib().grd_traverse(self.h, option), "grd_traverse")
        _check(lib().grd_reduce(self.h, option), "grd_reduce")
        f = ct.c_uint64()
        _check(lib().grd_update(self.h, ct.byref(f)), "grd_update")
        return f.value

    def result(self):
        out = np.zeros(self.graph.n, dtype=self._vdtype())
        _check(lib().grd_download(self.h, _ptr(out)), "grd_download")
        return out

    def pair_exchange(self, direction, c, m):
        """(gids ascending, value bits) of pair (c, m)'s broadcast (direction
        0) or reduce (direction 1) in the current superstep: uint32 bits, or
        uint64 values for an 8-byte state."""
        cnt = ct.c_uint64()
        _check(lib().grd_pair_exchange(self.h, direction, c, m, None, None, ct.byref(cnt)), "grd_pair_exchange")
        gids = np.zeros(max(cnt.value, 1), dtype=np.uint32)
        vals = np.zeros(max(cnt.value, 1), dtype=np.uint64 if self.value_bytes == 8 else np.uint32)
        _check(lib().grd_pair_exchange(self.h, direction, c, m, _ptr(gids), _ptr(vals), ct.byref(cnt)),
               "grd_pair_exchange")
        return gids[: cnt.value], vals[: cnt.value]

    def features(self):
        f = np.zeros(16, dtype=np.float32)
        _check(lib().grd_features(self.h, _ptr(f)), "grd_features")
        return f

    def snapshot(self):
        _check(lib().grd_snapshot(self.h), "grd_snapshot")

    def restore(self):
        _check(lib().grd_restore(self.h), "grd_restore")

    def free(self):
        if self.h:
            lib().grd_state_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many graphs processed together: one persistent launch per phase,
    iterations replayed from a captured hipGraph."""

    def __init__(self, ctx, graphs):
        """graphs: list of (row_ptr, col) numpy arrays."""
        self.ctx = ctx
        self._rp = [np.ascontiguousarray(g[0], dtype=np.uint32) for g in graphs]
        self._col = [np.ascontiguousarray(g[1], dtype=np.uint32) for g in graphs]
        cnt = len(graphs)
        self.n = np.array([len(r) - 1 for r in self._rp], dtype=np.uint32)
        self.m = np.array([len(c) for c in self._col], dtype=np.uint64)
        rps = (ct.c_void_p * cnt)(*[r.ctypes.data for r in self._rp])
        cols = (ct.c_void_p * cnt)(*[c.ctypes.data if len(c) else 0 for c in self._col])
        h = ct.c_void_p()
        _check(lib().grd_batch_create(ctx.h, cnt, _ptr(self.n), _ptr(self.m), rps, cols, ct.byref(h)),
               "grd_batch_create")
        self.h = h
        ctx._track(self)
        self.count = cnt

    def tune(self, tile=0, grid_per_cu=0, K=0):
        _check(lib().grd_batch_tune(self.h, tile, grid_per_cu, K), "grd_batch_tune")

    def autotune(self):
        out = np.zeros(3, dtype=np.uint32)
        _check(lib().grd_batch_autotune(self.h, _ptr(out)), "grd_batch_autotune")
        return {"tile": int(out[0]), "grid_per_cu": int(out[1]), "K": int(out[2])}

    def pagerank(self, damping=0.85, tolerance=1e-6, max_iter=1000):
        it = np.zeros(self.count, dtype=np.uint32)
        ms = ct.c_double()
        _check(lib().grd_batch_pagerank(self.h, damping, tolerance, max_iter, _ptr(it), ct.byref(ms)),
               "grd_batch_pagerank")
        return it, ms.value

    def cc(self, max_iter=1000):
        it = np.zeros(self.count, dtype=np.uint32)
        ms = ct.c_double()
        _check(lib().grd_batch_cc(self.h, max_iter, _ptr(it), ct.byref(ms)), "grd_batch_cc")
        return it, ms.value

    def result(self, i, algo):
        out = np.zeros(int(self.n[i]), dtype=np.float32 if algo == PR_PULL else np.uint32)
        _check(lib().grd_batch_download(self.h, i, algo, _ptr(out)), "grd_batch_download")
        return out

    def free(self):
        if self.h:
            lib().grd_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- measurement library (bench only) ----

PROBE_PATH = os.path.join(PKG, "libgrudon_probe.so")
_probe = None


def probe_lib():
    """libgrudon_probe.so: the random-gather ceiling and a streaming read of
    known size (measurement only; not part of the operator library)."""
    global _probe
    if _probe is None:
        if not os.path.exists(PROBE_PATH):
            raise GrdError("libgrudon_probe.so is not built (%s); run __graft_entry__.build()" % PROBE_PATH)
        L = ct.CDLL(PROBE_PATH)
        L.grdm_last_error.restype = ct.c_char_p
        L.grdm_probe_gather.restype = ct.c_int
        L.grdm_probe_gather.argtypes = [ct.c_int, ct.c_uint64, ct.c_uint64, ct.POINTER(ct.c_double),
                                        ct.POINTER(ct.c_double)]
        L.grdm_probe_stream.restype = ct.c_int
        L.grdm_probe_stream.argtypes = [ct.c_int, ct.c_uint64, ct.c_int, ct.c_int, ct.POINTER(ct.c_double)]
        _probe = L
    return _probe


def probe_gather(device, table_entries, n_gathers=1 << 28):
    """Random-gather ceiling: (ms per pass, gathers per second)."""
    ms, rate = ct.c_double(), ct.c_double()
    L = probe_lib()
    if L.grdm_probe_gather(device, table_entries, n_gathers, ct.byref(ms), ct.byref(rate)) != 0:
        raise GrdError("grdm_probe_gather: " + L.grdm_last_error().decode())
    return ms.value, rate.value


def probe_stream(device, nbytes, launches=3, write=False):
    """Streaming read (or write) of nbytes, launches times: mean ms per launch."""
    ms = ct.c_double()
    L = probe_lib()
    if L.grdm_probe_stream(device, nbytes, launches, int(write), ct.byref(ms)) != 0:
        raise GrdError("grdm_probe_stream: " + L.grdm_last_error().decode())
    return ms.value
