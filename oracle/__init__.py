"""CPU oracle for the dynamic-graph hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker and the CPU baseline.  The product
(meerkat_amd/) never imports it.  The algorithms are restated in oracle.c
(plain C + OpenMP); see its header for the reference lines each follows and
tests/test_oracle.py for how each function is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_lib = None


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_csr_build.argtypes = [_u32, _u64, _vp, _vp, _vp, _vp, _vp, _vp]
        L.orc_csr_build.restype = ctypes.c_int64
        L.orc_csr_transpose.argtypes = [_u32, _vp, _vp, _vp, _vp]
        L.orc_csr_transpose.restype = None
        L.orc_bfs.argtypes = [_u32, _vp, _vp, _u32, _vp]
        L.orc_bfs.restype = None
        L.orc_bfs_par.argtypes = [_u32, _vp, _vp, _u32, _vp, ctypes.c_int]
        L.orc_bfs_par.restype = _u32
        L.orc_dijkstra.argtypes = [_u32, _vp, _vp, _vp, _u32, _vp]
        L.orc_dijkstra.restype = ctypes.c_int
        L.orc_sssp_par.argtypes = [_u32, _vp, _vp, _vp, _u32, _vp, ctypes.c_int]
        L.orc_sssp_par.restype = ctypes.c_int64
        L.orc_pagerank.argtypes = [_u32, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _u32,
                                   _vp, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.orc_pagerank.restype = _u32
        L.orc_tc.argtypes = [_u32, _vp, _vp, ctypes.c_int]
        L.orc_tc.restype = _u64
        L.orc_cc.argtypes = [_u32, _vp, _vp, _vp]
        L.orc_cc.restype = _u32
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


class CSR:
    """Sorted, de-duplicated adjacency (row_off uint64[n+1], col, w)."""

    def __init__(self, n, row_off, col, w=None):
        self.n, self.row_off, self.col, self.w = n, row_off, col, w

    @property
    def m(self):
        return int(self.row_off[-1])

    def degrees(self):
        return np.diff(self.row_off).astype(np.uint32)

    def keys(self):
        src = np.repeat(np.arange(self.n, dtype=np.uint64), np.diff(self.row_off).astype(np.int64))
        return (src << np.uint64(32)) | self.col.astype(np.uint64)


def csr(n, src, dst, w=None):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    w = None if w is None else np.ascontiguousarray(w, np.uint32)
    m = len(src)
    row = np.empty(n + 1, np.uint64)
    col = np.empty(max(m, 1), np.uint32)
    wc = np.empty(max(m, 1), np.uint32) if w is not None else None
    got = lib().orc_csr_build(n, m, _p(src), _p(dst), _p(w), _p(row), _p(col), _p(wc))
    if got < 0:
        raise ValueError("edge endpoint out of range")
    return CSR(n, row, col[:got].copy(), None if wc is None else wc[:got].copy())


def transpose(g):
    rrow = np.empty(g.n + 1, np.uint64)
    rcol = np.empty(max(g.m, 1), np.uint32)
    lib().orc_csr_transpose(g.n, _p(g.row_off), _p(g.col), _p(rrow), _p(rcol))
    return CSR(g.n, rrow, rcol[:g.m].copy())


def bfs(g, source):
    d = np.empty(g.n, np.uint32)
    lib().orc_bfs(g.n, _p(g.row_off), _p(g.col), source, _p(d))
    return d


def bfs_par(g, source, threads=0):
    d = np.empty(g.n, np.uint32)
    lib().orc_bfs_par(g.n, _p(g.row_off), _p(g.col), source, _p(d), threads)
    return d


def dijkstra(g, source):
    d = np.empty(g.n, np.uint32)
    rc = lib().orc_dijkstra(g.n, _p(g.row_off), _p(g.col), _p(g.w), source, _p(d))
    if rc != 0:
        raise OverflowError("distance exceeds uint32")
    return d


def sssp_par(g, source, threads=0):
    """Multithreaded label-correcting SSSP (the CPU baseline); same distances
    as dijkstra()."""
    d = np.empty(g.n, np.uint32)
    rc = lib().orc_sssp_par(g.n, _p(g.row_off), _p(g.col), _p(g.w), source, _p(d), threads)
    if rc < 0:
        raise OverflowError("distance exceeds uint32")
    return d


def pagerank(gin, outdeg, damping=0.85, margin=1e-14, max_iter=1000, init=None, threads=0):
    """FP64 power iteration on the in-adjacency `gin`; returns (pr, iters, delta)."""
    pr = (np.full(gin.n, 1.0 / gin.n) if init is None
          else np.array(init, dtype=np.float64, copy=True))
    outdeg = np.ascontiguousarray(outdeg, np.uint32)
    delta = ctypes.c_double()
    it = lib().orc_pagerank(gin.n, _p(gin.row_off), _p(gin.col), _p(outdeg), damping, margin,
                            max_iter, _p(pr), ctypes.byref(delta), threads)
    return pr, it, delta.value


def tc(g, threads=0):
    return int(lib().orc_tc(g.n, _p(g.row_off), _p(g.col), threads))


def cc(g):
    labels = np.empty(g.n, np.uint32)
    k = lib().orc_cc(g.n, _p(g.row_off), _p(g.col), _p(labels))
    return labels, int(k)


def num_threads():
    return int(lib().orc_num_threads())


def check_parents(g, dist, parent, source, weights=False):
    """Parent validity: dist[p] + w(p, v) == dist[v] and (p, v) is an edge,
    for every reached v != source.  Returns the number of violations."""
    bad = 0
    reached = np.nonzero(dist != 0xFFFFFFFF)[0]
    for v in reached:
        if v == source:
            continue
        p = parent[v]
        if p >= g.n:
            bad += 1
            continue
        lo, hi = int(g.row_off[p]), int(g.row_off[p + 1])
        j = lo + np.searchsorted(g.col[lo:hi], v)
        if j >= hi or g.col[j] != v:
            bad += 1
            continue
        w = int(g.w[j]) if weights else 1
        if int(dist[p]) + w != int(dist[v]):
            bad += 1
    return bad


def check_parents_fast(g, dist, parent, source, weights=False):
    """Vectorised check_parents for large graphs."""
    v = np.nonzero(dist != 0xFFFFFFFF)[0]
    v = v[v != source]
    if len(v) == 0:
        return 0
    p = parent[v].astype(np.int64)
    in_range = p < g.n
    p = np.where(in_range, p, 0)  # ids past n are violations; read vertex 0 instead
    keys = g.keys()
    if len(keys) == 0:
        return len(v)
    q = (p.astype(np.uint64) << np.uint64(32)) | v.astype(np.uint64)
    j = np.searchsorted(keys, q)
    j = np.minimum(j, len(keys) - 1)
    present = keys[j] == q
    w = g.w[j].astype(np.int64) if weights else np.ones(len(v), np.int64)
    ok = in_range & present & (dist[p].astype(np.int64) + w == dist[v].astype(np.int64))
    return int((~ok).sum())
