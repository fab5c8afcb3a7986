"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The oracle is the checker, never the
thing measured as the product.
"""
import ctypes as ct
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GRUDON_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")

_P32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_P64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_PF64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

ORC_PR, ORC_CC, ORC_SSSP = 0, 1, 2
RMAT_ABC = (0.57, 0.19, 0.19)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run __graft_entry__.build() first (%s)" % LIB_PATH)
        L = ct.CDLL(LIB_PATH)
        L.orc_hash.restype = ct.c_uint64
        L.orc_hash.argtypes = [ct.c_uint64, ct.c_uint64]
        L.orc_rmat.restype = ct.c_int64
        L.orc_rmat.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_uint64, ct.c_int,
                               _P32, _P32, ct.c_int64]
        L.orc_er.restype = ct.c_int64
        L.orc_er.argtypes = [ct.c_uint32, ct.c_double, ct.c_uint64, _P32, _P32, ct.c_int64]
        L.orc_weights.restype = None
        L.orc_weights.argtypes = [ct.c_uint32, _P32, _P32, ct.c_uint64, _P32]
        L.orc_csr_from_pairs.restype = ct.c_int64
        L.orc_csr_from_pairs.argtypes = [ct.c_uint32, _P64, ct.c_int64, _P32, _P32]
        L.orc_part_build.restype = ct.c_void_p
        L.orc_part_build.argtypes = [ct.c_uint32, _P32, _P32, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p]
        L.orc_part_free.restype = None
        L.orc_part_free.argtypes = [ct.c_void_p]
        L.orc_part_query.restype = ct.c_int64
        L.orc_part_query.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]
        L.orc_bsp_u32.restype = ct.c_int
        L.orc_bsp_u32.argtypes = [ct.c_void_p, ct.c_int, ct.c_uint32, ct.c_int, ct.c_uint32, ct.c_int, ct.c_int,
                                  _P32, _P64, ct.c_void_p]
        L.orc_bsp_pr.restype = ct.c_int
        L.orc_bsp_pr.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, _PF64, _P64, ct.c_void_p]
        L.orc_bsp_u32_sched.restype = ct.c_int
        L.orc_bsp_u32_sched.argtypes = [ct.c_void_p, ct.c_int, ct.c_uint32, ct.c_int, ct.c_uint32, ct.c_int,
                                        ct.c_int, ct.c_int, _P32, _P64, ct.c_void_p, ct.c_void_p]
        L.orc_bsp_pr_sched.restype = ct.c_int
        L.orc_bsp_pr_sched.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _PF64, _P64,
                                       ct.c_void_p, ct.c_void_p]
        L.orc_bsp_pair_exchange.restype = ct.c_int64
        L.orc_bsp_pair_exchange.argtypes = [ct.c_void_p, ct.c_int, ct.c_uint32, ct.c_int, ct.c_uint32, ct.c_int,
                                            ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p]
        L.orc_bfs.restype = None
        L.orc_bfs.argtypes = [ct.c_uint32, _P32, _P32, ct.c_uint32, _P32]
        L.orc_dijkstra.restype = None
        L.orc_dijkstra.argtypes = [ct.c_uint32, _P32, _P32, _P32, ct.c_uint32, _P32]
        L.orc_cc_unionfind.restype = None
        L.orc_cc_unionfind.argtypes = [ct.c_uint32, _P32, _P32, _P32]
        L.orc_pr_pull.restype = ct.c_int
        L.orc_pr_pull.argtypes = [ct.c_uint32, _P32, _P32, ct.c_double, ct.c_int, ct.c_double, _PF64,
                                  ct.POINTER(ct.c_double)]
        L.orc_pull_plan_build.restype = ct.c_void_p
        L.orc_pull_plan_build.argtypes = [ct.c_uint32, _P32, _P32, ct.c_int]
        L.orc_pull_plan_run.restype = ct.c_int
        L.orc_pull_plan_run.argtypes = [ct.c_void_p, ct.c_double, ct.c_int, ct.c_double, _PF64,
                                        ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
        L.orc_pull_plan_arcs.restype = ct.c_uint64
        L.orc_pull_plan_arcs.argtypes = [ct.c_void_p]
        L.orc_pull_plan_free.restype = None
        L.orc_pull_plan_free.argtypes = [ct.c_void_p]
        L.orc_cc_unionfind_mt.restype = None
        L.orc_cc_unionfind_mt.argtypes = [ct.c_uint32, _P32, _P32, _P32, ct.c_int]
        L.orc_sssp_check_mt.restype = ct.c_uint64
        L.orc_sssp_check_mt.argtypes = [ct.c_uint32, _P32, _P32, ct.c_void_p, _P32, ct.c_uint32, ct.c_int]
        L.orc_pr_pull_mt.restype = ct.c_int
        L.orc_pr_pull_mt.argtypes = [ct.c_uint32, _P32, _P32, ct.c_double, ct.c_int, ct.c_double, _PF64,
                                     ct.POINTER(ct.c_double), ct.c_int]
        _lib = L
    return _lib


class Graph:
    """Global CSR (row_ptr u32[n+1], col u32[m], optional weights u32[m])."""

    def __init__(self, rp, col, w=None):
        self.rp = np.ascontiguousarray(rp, dtype=np.uint32)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        self.n = len(self.rp) - 1
        self.m = len(self.col)

    def with_weights(self, wseed):
        w = np.zeros(self.m, dtype=np.uint32)
        lib().orc_weights(self.n, self.rp, self.col, wseed, w)
        return Graph(self.rp, self.col, w)

    def out_degree(self):
        return np.diff(self.rp.astype(np.int64))


def rmat(scale, ef=16, seed=1, symmetrize=False, abc=RMAT_ABC):
    n = 1 << scale
    raw = (ef << scale) * (2 if symmetrize else 1)
    rp = np.zeros(n + 1, dtype=np.uint32)
    col = np.zeros(max(raw, 1), dtype=np.uint32)
    m = lib().orc_rmat(scale, ef, abc[0], abc[1], abc[2], seed, 1 if symmetrize else 0, rp, col, raw)
    assert m >= 0
    return Graph(rp, col[:m].copy())


def er(n, mean_degree, seed):
    cap = int(n * mean_degree / 2.0) * 2
    rp = np.zeros(n + 1, dtype=np.uint32)
    col = np.zeros(max(cap, 1), dtype=np.uint32)
    m = lib().orc_er(n, mean_degree, seed, rp, col, cap)
    assert m >= 0
    return Graph(rp, col[:m].copy())


def parity_doubled(g):
    """Two copies of g on the even and the odd gids (u -> v becomes 2u -> 2v
    and 2u+1 -> 2v+1): with C = 2 every arc stays on its master, so every
    vertex is covered."""
    deg = g.out_degree()
    src = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    rank = np.arange(g.m, dtype=np.int64) - g.rp.astype(np.int64)[src]
    rp2 = np.zeros(2 * g.n + 1, dtype=np.int64)
    rp2[1:] = np.cumsum(np.repeat(deg, 2))
    col2 = np.zeros(2 * g.m, dtype=np.uint32)
    w2 = None if g.w is None else np.zeros(2 * g.m, dtype=np.uint32)
    for par in (0, 1):
        idx = rp2[2 * src + par] + rank
        col2[idx] = 2 * g.col.astype(np.int64) + par
        if w2 is not None:
            w2[idx] = g.w
    return Graph(rp2.astype(np.uint32), col2, w2)


def from_edges(n, edges):
    """CSR from an (src, dst) edge list: sorted, self-loops and duplicates dropped."""
    keys = np.array([(int(s) << 32) | int(d) for s, d in edges] or [0], dtype=np.uint64)
    cnt = len(edges)
    rp = np.zeros(n + 1, dtype=np.uint32)
    col = np.zeros(max(cnt, 1), dtype=np.uint32)
    m = lib().orc_csr_from_pairs(n, keys, cnt, rp, col)
    return Graph(rp, col[:m].copy())


class Partition:
    """Restated DistributedGraph layout for C compute and M memory partitions."""

    def __init__(self, g, C, M, mirror=None):
        self.g = g
        self.C, self.M = C, M
        self._mirror = None if mirror is None else np.ascontiguousarray(mirror, dtype=np.uint32)
        wptr = None if g.w is None else g.w.ctypes.data
        mptr = None if self._mirror is None else self._mirror.ctypes.data
        self.h = lib().orc_part_build(g.n, g.rp, g.col, wptr, C, M, mptr)
        if not self.h:
            raise ValueError("invalid partition")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_part_free(self.h)
            self.h = None

    def query(self, what, i=0, j=0, fetch=True):
        cnt = lib().orc_part_query(self.h, what, i, j, None)
        if not fetch:
            return cnt
        out = np.zeros(max(cnt, 1), dtype=np.uint32)
        lib().orc_part_query(self.h, what, i, j, out.ctypes.data)
        return out[:cnt]

    def pair_exchange(self, algo, it, direction, c, m, source=0, weighted=False, delta=0):
        """(gids ascending, uint32 values) of pair (c, m)'s broadcast
        (direction 0) or reduce (direction 1) at iteration `it` of the
        offloaded CC / SSSP run."""
        T = int(self.query(0, c, m, fetch=False))
        gids = np.zeros(max(T, 1), dtype=np.uint32)
        vals = np.zeros(max(T, 1), dtype=np.uint32)
        k = lib().orc_bsp_pair_exchange(self.h, algo, source, 1 if weighted else 0, delta, it, direction, c, m,
                                        gids.ctypes.data, vals.ctypes.data)
        assert k >= 0
        return gids[:k], vals[:k]

    def coverage(self):
        """covered[gid] (1 = every out-neighbour shares the master)."""
        return self.query(10).astype(np.uint8)

    def _u32(self, algo, source, weighted, delta, max_iter, nthreads, trace, sched):
        out = np.zeros(self.g.n, dtype=np.uint32)
        st = np.zeros(4, dtype=np.uint64)
        tr = np.zeros(max_iter + 1, dtype=np.uint32) if trace else None
        votes = np.zeros(max_iter + 1, dtype=np.uint8)
        rc = lib().orc_bsp_u32_sched(self.h, algo, source, 1 if weighted else 0, delta, max_iter, nthreads, sched,
                                     out, st, None if tr is None else tr.ctypes.data, votes.ctypes.data)
        assert rc == 0
        it = int(st[0])
        res = (out, st, tr[:it]) if trace else (out, st)
        return res + (votes[:it],) if sched else res

    def cc(self, max_iter=1000, nthreads=1, trace=False, sched=0):
        """sched 0: offloaded superstep every iteration; 1: NDPEngine vote;
        2: edge shipping whenever the frontier is covered.  With sched, the
        per-iteration shipping votes are appended to the result."""
        return self._u32(ORC_CC, 0, False, 0, max_iter, nthreads, trace, sched)

    def sssp(self, source=0, weighted=False, delta=0, max_iter=100000, nthreads=1, trace=False, sched=0):
        return self._u32(ORC_SSSP, source, weighted, delta, max_iter, nthreads, trace, sched)

    def pr(self, fp64=False, max_iter=1000, nthreads=1, trace=False, sched=0):
        out = np.zeros(self.g.n, dtype=np.float64)
        st = np.zeros(4, dtype=np.uint64)
        tr = np.zeros(max_iter + 1, dtype=np.uint32) if trace else None
        votes = np.zeros(max_iter + 1, dtype=np.uint8)
        rc = lib().orc_bsp_pr_sched(self.h, 1 if fp64 else 0, max_iter, nthreads, sched, out, st,
                                    None if tr is None else tr.ctypes.data, votes.ctypes.data)
        assert rc == 0
        it = int(st[0])
        res = (out, st, tr[:it]) if trace else (out, st)
        return res + (votes[:it],) if sched else res


def bfs(g, src):
    out = np.zeros(g.n, dtype=np.uint32)
    lib().orc_bfs(g.n, g.rp, g.col, src, out)
    return out


def dijkstra(g, src):
    out = np.zeros(g.n, dtype=np.uint32)
    lib().orc_dijkstra(g.n, g.rp, g.col, g.w, src, out)
    return out


def cc_unionfind(g):
    out = np.zeros(g.n, dtype=np.uint32)
    lib().orc_cc_unionfind(g.n, g.rp, g.col, out)
    return out


def pr_pull(g, damping=0.85, iters=0, tol=1e-10, nthreads=0):
    """FP64 power iteration; nthreads > 0 uses the threaded pull form."""
    r = np.zeros(g.n, dtype=np.float64)
    res = ct.c_double(0.0)
    if nthreads > 0:
        it = lib().orc_pr_pull_mt(g.n, g.rp, g.col, damping, iters, tol, r, ct.byref(res), nthreads)
    else:
        it = lib().orc_pr_pull(g.n, g.rp, g.col, damping, iters, tol, r, ct.byref(res))
    return r, it, res.value


class PullPlan:
    """The threaded FP64 pull PageRank with its transposed graph built once:
    run() repeats a solve from r0 = 1/n and returns (r, iterations,
    last residual, seconds of the iteration loop)."""

    def __init__(self, rp, col, nthreads):
        self.rp = np.ascontiguousarray(rp, dtype=np.uint32)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.n = len(self.rp) - 1
        self.h = lib().orc_pull_plan_build(self.n, self.rp, self.col, nthreads)

    @property
    def arcs(self):
        return int(lib().orc_pull_plan_arcs(self.h))

    def run(self, iters=0, tol=1e-6, damping=0.85):
        r = np.zeros(self.n, dtype=np.float64)
        res, sec = ct.c_double(0.0), ct.c_double(0.0)
        it = lib().orc_pull_plan_run(self.h, damping, iters, tol, r, ct.byref(res), ct.byref(sec))
        return r, it, res.value, sec.value

    def free(self):
        if getattr(self, "h", None):
            lib().orc_pull_plan_free(self.h)
            self.h = None

    def __del__(self):
        self.free()


def cc_unionfind_mt(rp, col, nthreads):
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    out = np.zeros(len(rp) - 1, dtype=np.uint32)
    lib().orc_cc_unionfind_mt(len(rp) - 1, rp, col, out, nthreads)
    return out


def sssp_violations(rp, col, w, dist, src, nthreads):
    """Violated Bellman conditions of dist from src (0 = exact distances)."""
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    dist = np.ascontiguousarray(dist, dtype=np.uint32)
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.uint32)
        wp = w.ctypes.data
    return int(lib().orc_sssp_check_mt(len(rp) - 1, rp, col, wp, dist, src, nthreads))
