"""Host-side mirror of the reference's algorithm classes over the C ABI.

Each class keeps the reference's method names and protocol: mutate the graph
first (``InsertEdges`` / ``DeleteEdges``), then call ``Incremental`` /
``Decremental`` with the same batch; every compute call blocks and returns
elapsed milliseconds (include/Algorithms/BFS.cuh:34-68, SSSP.cuh:34-69,
PageRank.cuh:32-67, TC.cuh:28-40), except BFS/SSSP Incremental /
Decremental on a handle set to deferred mode (``set_deferred``).
"""
import ctypes

import numpy as np

from . import _lib
from .graph import device_u32, dptr

_vp = ctypes.c_void_p
INF = 0xFFFFFFFF


class _Traversal:
    _prefix = None

    def __init__(self, graph, source):
        self.G = graph
        self.source = int(source)
        self._h = _vp()
        _lib.call(f"{self._prefix}_create", graph.handle, self.source, ctypes.byref(self._h))

    @property
    def handle(self):
        """The C-ABI handle (mk_bfs* / mk_sssp*)."""
        return self._h

    def Static(self, edges_n=0):
        ms = ctypes.c_float()
        _lib.call(f"{self._prefix}_static", self._h, ctypes.byref(ms))
        return ms.value

    def set_source(self, source):
        """Rebind the source; call Static() before reading distances."""
        _lib.call(f"{self._prefix}_set_source", self._h, int(source))
        self.source = int(source)

    def read(self):
        """(distances, parents) as uint32 numpy arrays; INF = 0xFFFFFFFF."""
        d = np.empty(self.G.n, np.uint32)
        p = np.empty(self.G.n, np.uint32)
        _lib.call(f"{self._prefix}_read", self._h, d.ctypes.data_as(_vp), p.ctypes.data_as(_vp))
        return d, p

    def Distances(self):
        return self.read()[0]

    def capture_step(self, del_src, del_dst, ins_src, ins_dst, ins_w=None):
        """Record one update step on fixed device batch buffers (torch
        tensors or device pointers), DeleteEdges -> Decremental ->
        InsertEdges -> Incremental, into a hipGraph; returns a CapturedStep
        whose launch() replays it (refill the buffers between replays)."""
        from .graph import device_u32, dptr

        dev = self.G.device
        bufs = [device_u32(b, dev) for b in (del_src, del_dst, ins_src, ins_dst, ins_w)]
        ds, dd, is_, id_, iw = bufs
        out = _vp()
        if self._prefix == "mk_sssp":
            _lib.call("mk_sssp_capture_step", self._h, dptr(ds), dptr(dd), len(ds), dptr(is_),
                      dptr(id_), dptr(iw), len(is_), ctypes.byref(out))
        else:
            _lib.call("mk_bfs_capture_step", self._h, dptr(ds), dptr(dd), len(ds), dptr(is_),
                      dptr(id_), len(is_), ctypes.byref(out))
        return CapturedStep(out, bufs)

    def Parents(self):
        return self.read()[1]

    def set_deferred(self, on=True):
        """Deferred calls: Incremental / Decremental return once queued on the
        graph's stream (returning 0.0: not measured) and the host does not
        wait for them; the handle's next call, read, last_stats or sync()
        settles the call, raising its MeerkatError if it failed."""
        _lib.call(f"{self._prefix}_set_deferred", self._h, int(on))

    def sync(self):
        """Settle a deferred call: raises its MeerkatError if it failed."""
        _lib.call(f"{self._prefix}_sync", self._h)

    def deferred_stats(self):
        """Deferred calls settled since create, and how many were recomputed
        from scratch at settling (their relax loop ran out of steps)."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.call(f"{self._prefix}_deferred_stats", self._h, ctypes.byref(a), ctypes.byref(b))
        return {"settled": a.value, "recomputed": b.value}

    def last_stats(self):
        it, slabs, edges = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        _lib.call(f"{self._prefix}_last_stats", self._h, ctypes.byref(it), ctypes.byref(slabs),
                  ctypes.byref(edges))
        return {"iterations": it.value, "slabs_read": slabs.value, "edges_read": edges.value}

    def close(self):
        if self._h:
            _lib.call(f"{self._prefix}_destroy", self._h)
            self._h = _vp()

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h:
                getattr(_lib.lib(), f"{self._prefix}_destroy")(self._h)
        except Exception:
            pass


class CapturedStep:
    """A captured update step (mk_step): launch() queues a replay on the
    graph's stream, sync() waits for the queued replays and raises on a
    failed batch or loop, close() releases it (before its handle)."""

    def __init__(self, h, buffers):
        self._h = h
        # the device buffers the replays read (refill them in place between
        # replays: buffers[0..4] = delete src, dst, insert src, dst, weights)
        self.buffers = buffers
        self._pending = []

    def launch(self):
        _lib.call("mk_step_launch", self._h)

    def load(self, del_src, del_dst, ins_src, ins_dst, ins_w=None):
        """Queue copies of the next batches (device tensors of the captured
        sizes) into the step's buffers, ordered between the replays."""
        from .graph import dptr

        # the sources stay alive until the copies are known to have run
        self._pending.append((del_src, del_dst, ins_src, ins_dst, ins_w))
        _lib.call("mk_step_load", self._h, dptr(del_src), dptr(del_dst), dptr(ins_src),
                  dptr(ins_dst), dptr(ins_w))

    def sync(self):
        try:
            _lib.call("mk_step_sync", self._h)
        finally:
            self._pending = []

    def close(self):
        if self._h:
            h, self._h = self._h, _vp()
            _lib.call("mk_step_destroy", h)

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h:
                _lib.lib().mk_step_destroy(self._h)
        except Exception:
            pass


class BFS(_Traversal):
    """Levels from a source vertex (include/Algorithms/BFS.cuh:13-69)."""

    _prefix = "mk_bfs"

    def Incremental(self, src, dst):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_bfs_incremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        self._batch = (s, d)  # a deferred call reads it until it is settled
        return ms.value

    def Decremental(self, src, dst):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_bfs_decremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        self._batch = (s, d)
        return ms.value


class SSSP(_Traversal):
    """Shortest distances with uint32 weights (include/Algorithms/SSSP.cuh:11-69)."""

    _prefix = "mk_sssp"

    def Incremental(self, src, dst, weights=None):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        w = device_u32(weights, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_sssp_incremental", self._h, dptr(s), dptr(d), dptr(w), len(s),
                  ctypes.byref(ms))
        self._batch = (s, d, w)  # a deferred call reads it until it is settled
        return ms.value

    def Decremental(self, src, dst, weights=None):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        w = device_u32(weights, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_sssp_decremental", self._h, dptr(s), dptr(d), dptr(w), len(s),
                  ctypes.byref(ms))
        self._batch = (s, d, w)  # a deferred call reads it until it is settled
        return ms.value

    def Reweight(self, src, dst, weights):
        """Distances kept exact after ``G.UpdateEdges`` with the same batch
        (``weights``: its new weights): tree edges that became heavier are
        handled as deletions, lighter ones as insertions."""
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        w = device_u32(weights, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_sssp_reweight", self._h, dptr(s), dptr(d), dptr(w), len(s),
                  ctypes.byref(ms))
        self._batch = (s, d, w)
        return ms.value


class PageRank:
    """FP32 PageRank with FP64 reductions (include/Algorithms/PageRank.cuh:9-68).

    ``graph`` holds reversed edges (reference style; pass ``out_degree``) or
    was created with ``reverse=True`` (its in-edge mirror is used and the
    out-degrees are exact)."""

    def __init__(self, graph, damping=0.85, error_margin=1e-7, out_degree=None, max_iter=100):
        self.G = graph
        self._h = _vp()
        od = None
        if out_degree is not None:
            od = np.ascontiguousarray(out_degree, dtype=np.uint32)
            if len(od) != graph.n:
                raise ValueError("out_degree must have n entries")
        _lib.call("mk_pagerank_create", graph.handle, float(damping), float(error_margin),
                  int(max_iter), None if od is None else od.ctypes.data_as(_vp),
                  ctypes.byref(self._h))

    def Static(self):
        ms = ctypes.c_float()
        _lib.call("mk_pagerank_static", self._h, ctypes.byref(ms))
        return ms.value

    def Incremental(self, src, dst):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_pagerank_incremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        return ms.value

    def Decremental(self, src, dst):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_pagerank_decremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        return ms.value

    def Values(self):
        out = np.empty(self.G.n, np.float32)
        _lib.call("mk_pagerank_read", self._h, out.ctypes.data_as(_vp))
        return out

    def last_stats(self):
        it, delta = ctypes.c_uint32(), ctypes.c_double()
        _lib.call("mk_pagerank_last_stats", self._h, ctypes.byref(it), ctypes.byref(delta))
        return {"iterations": it.value, "delta": delta.value}

    def close(self):
        if self._h:
            _lib.call("mk_pagerank_destroy", self._h)
            self._h = _vp()

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h:
                _lib.lib().mk_pagerank_destroy(self._h)
        except Exception:
            pass


class CC:
    """Connected components with min-vertex labels by union-find (the
    reference's include/Util/UnionFind.cuh primitives; no reference driver)."""

    def __init__(self, graph):
        self.G = graph
        self._h = _vp()
        _lib.call("mk_cc_create", graph.handle, ctypes.byref(self._h))

    def Static(self):
        ms = ctypes.c_float()
        _lib.call("mk_cc_static", self._h, ctypes.byref(ms))
        return ms.value

    def Incremental(self, src, dst):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_cc_incremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        return ms.value

    def Decremental(self, src, dst):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_cc_decremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        return ms.value

    def Labels(self):
        out = np.empty(self.G.n, np.uint32)
        _lib.call("mk_cc_labels", self._h, out.ctypes.data_as(_vp))
        return out

    def close(self):
        if self._h:
            _lib.call("mk_cc_destroy", self._h)
            self._h = _vp()

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h:
                _lib.lib().mk_cc_destroy(self._h)
        except Exception:
            pass


class TC:
    """Triangle counting on an undirected graph stored in both directions
    (include/Algorithms/TC.cuh:9-41).  Batches must hold both directions."""

    def __init__(self, graph):
        self.G = graph
        self._h = _vp()
        _lib.call("mk_tc_create", graph.handle, ctypes.byref(self._h))

    def Static(self, edges_n=0):
        ms = ctypes.c_float()
        _lib.call("mk_tc_static", self._h, ctypes.byref(ms))
        return ms.value

    def Incremental(self, src, dst, alloc=None):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_tc_incremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        return ms.value

    def Decremental(self, src, dst, alloc=None):
        s, d = device_u32(src, self.G.device), device_u32(dst, self.G.device)
        ms = ctypes.c_float()
        _lib.call("mk_tc_decremental", self._h, dptr(s), dptr(d), len(s), ctypes.byref(ms))
        return ms.value

    def GetCount(self):
        c = ctypes.c_uint64()
        _lib.call("mk_tc_count", self._h, ctypes.byref(c))
        return c.value

    def close(self):
        if self._h:
            _lib.call("mk_tc_destroy", self._h)
            self._h = _vp()

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h:
                _lib.lib().mk_tc_destroy(self._h)
        except Exception:
            pass


class MultiSource:
    """Up to 64 BFS (sssp=False) or SSSP queries answered by one traversal."""

    MAX_SOURCES = 64

    def __init__(self, graph, sssp=False):
        self.G = graph
        self.sssp = bool(sssp)
        self._h = _vp()
        _lib.call("mk_msq_create", graph.handle, int(self.sssp), ctypes.byref(self._h))

    def run(self, sources):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        ms = ctypes.c_float()
        _lib.call("mk_msq_run", self._h, src.ctypes.data_as(_vp), len(src), ctypes.byref(ms))
        return ms.value

    def distances(self, slot):
        out = np.empty(self.G.n, np.uint32)
        _lib.call("mk_msq_read", self._h, int(slot), out.ctypes.data_as(_vp))
        return out

    # continuous batching: a session admits sources into free slots while
    # it runs; each slot finishes on its own step
    def open(self):
        _lib.call("mk_msq_open", self._h)

    def admit(self, sources, slots):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        sl = np.ascontiguousarray(slots, dtype=np.uint32)
        assert len(src) == len(sl)
        _lib.call("mk_msq_admit", self._h, src.ctypes.data_as(_vp), sl.ctypes.data_as(_vp),
                  len(src))

    def launch(self, steps=1):
        _lib.call("mk_msq_launch", self._h, int(steps))

    def ready(self):
        """Whether the launched steps have run (poll() would not wait)."""
        r = ctypes.c_int()
        _lib.call("mk_msq_ready", self._h, ctypes.byref(r))
        return bool(r.value)

    def poll(self):
        """Wait for the launched steps; returns (active slot mask, steps so far).
        Raises MeerkatError (MK_ERR_STATE) when a BFS session's level log
        has filled up: its active slots are incomplete and belong in a new
        session (with a larger log); the finished slots stay readable until
        then."""
        a, st = ctypes.c_uint64(), ctypes.c_uint32()
        _lib.call("mk_msq_poll", self._h, ctypes.byref(a), ctypes.byref(st))
        return a.value, st.value

    def last_stats(self):
        st, e = ctypes.c_uint32(), ctypes.c_uint64()
        _lib.call("mk_msq_last_stats", self._h, ctypes.byref(st), ctypes.byref(e))
        return {"steps": st.value, "edges_read": e.value}

    def close(self):
        if self._h:
            _lib.call("mk_msq_destroy", self._h)
            self._h = _vp()

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h:
                _lib.lib().mk_msq_destroy(self._h)
        except Exception:
            pass
