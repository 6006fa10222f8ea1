"""Host-side mirror of the reference's graph API over the C ABI.

``SlabPool`` stands for ``AllocPolicy::DynamicAllocatorT`` and
``DynamicGraph`` for ``DynamicGraph<EdgePolicy, IsWeighted>``
(include/Graph/SlabGraph.cuh:343-506, 594-752): same construction
arguments (vertex count, allocator, load factor, degree hints, device) and
the same batched ``InsertEdges`` / ``DeleteEdges`` entry points, which take
device arrays.  Device arrays are torch tensors (int32 or uint32 on the GPU);
numpy arrays are copied to the device first.  torch is only the carrier of
device memory here; all work runs in libmeerkat_hip.so.
"""
import ctypes

import numpy as np

from . import _lib

_vp = ctypes.c_void_p


def _torch():
    import torch

    return torch


def device_u32(a, device=0):
    """A device int32 tensor holding the bits of a uint32 array (or the tensor
    itself if it already lives on the GPU)."""
    torch = _torch()
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            a = a.to(f"cuda:{device}")
        if a.dtype not in (torch.int32, torch.uint32):
            raise TypeError("device arrays must be int32/uint32")
        return a.contiguous()
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return torch.from_numpy(a.view(np.int32)).to(f"cuda:{device}")


def dptr(t):
    return None if t is None else _vp(t.data_ptr())


class SlabPool:
    """HBM-resident pool of 256-byte slabs with a bitmap allocator."""

    def __init__(self, nbytes=0, device=0):
        self._h = _vp()
        self.device = device
        _lib.call("mk_pool_create", nbytes, device, ctypes.byref(self._h))

    @property
    def handle(self):
        return self._h

    def stats(self):
        used, total = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.call("mk_pool_stats", self._h, ctypes.byref(used), ctypes.byref(total))
        return used.value, total.value

    def close(self):
        if self._h:
            _lib.call("mk_pool_destroy", self._h)
            self._h = _vp()

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h:
                _lib.lib().mk_pool_destroy(self._h)
        except Exception:
            pass


class DynamicGraph:
    """Per-vertex slab-hash adjacency: a set (unweighted) or a map from
    destination to uint32 weight (weighted)."""

    WEIGHTED = 1
    REVERSE = 2
    TRACK_UPDATES = 4

    def __init__(self, n, pool, load_factor=0.7, degree_hints=None, weighted=False, device=0,
                 reverse=False, in_degree_hints=None, track_updates=False):
        """reverse=True also keeps an in-edge mirror, updated by every
        InsertEdges/DeleteEdges, which decremental BFS/SSSP pull from.
        track_updates=True marks inserted keys for UpdatedEdges()."""
        self.n = int(n)
        self.pool = pool
        self.weighted = bool(weighted)
        self.reverse = bool(reverse)
        self.device = device
        self._h = _vp()
        self._owned = True

        def _hint(h):
            if h is None:
                return None
            h = np.ascontiguousarray(h, dtype=np.uint32)
            if len(h) != self.n:
                raise ValueError("degree hints must have n entries")
            return h

        hints, ihints = _hint(degree_hints), _hint(in_degree_hints)
        self._hints = (hints, ihints)
        self.track_updates = bool(track_updates)
        flags = ((self.WEIGHTED if self.weighted else 0) | (self.REVERSE if self.reverse else 0) |
                 (self.TRACK_UPDATES if self.track_updates else 0))
        _lib.call("mk_graph_create_ex", pool.handle, self.n, flags, float(load_factor),
                  None if hints is None else hints.ctypes.data_as(_vp),
                  None if ihints is None else ihints.ctypes.data_as(_vp), ctypes.byref(self._h))

    def mirror(self):
        """Read-only view of the in-edge mirror (inspection calls only)."""
        h = _vp()
        _lib.call("mk_graph_reverse", self._h, ctypes.byref(h))
        m = DynamicGraph.__new__(DynamicGraph)
        m.n, m.pool, m.weighted, m.reverse, m.device = self.n, self.pool, self.weighted, False, \
            self.device
        m._h, m._owned, m._hints, m._parent = h, False, (None, None), self
        return m

    @property
    def handle(self):
        return self._h

    # ---------------------------------------------------------- mutation
    def InsertEdges(self, src, dst, weights=None, stream=None):
        s, d = device_u32(src, self.device), device_u32(dst, self.device)
        w = device_u32(weights, self.device)
        _lib.call("mk_graph_insert", self._h, dptr(s), dptr(d), dptr(w), len(s), stream)
        self._batch = (s, d, w)  # kept alive until the next batch settles this one
        return s, d, w

    def DeleteEdges(self, src, dst, stream=None):
        s, d = device_u32(src, self.device), device_u32(dst, self.device)
        _lib.call("mk_graph_delete", self._h, dptr(s), dptr(d), len(s), stream)
        self._batch = (s, d)
        return s, d

    def UpdateEdges(self, src, dst, weights, stream=None):
        """Set the weight of every present edge of the batch (an absent edge
        is not inserted; last_status() tells which were present).  The
        reference's UpdateEdges (include/Graph/SlabGraphImpl.cuh:84-103,
        compiled out there) without its FilterMaps."""
        s, d = device_u32(src, self.device), device_u32(dst, self.device)
        w = device_u32(weights, self.device)
        _lib.call("mk_graph_update_weights", self._h, dptr(s), dptr(d), dptr(w), len(s), stream)
        self._batch = (s, d, w)
        return s, d, w

    update_edges = UpdateEdges

    def UpdateSlabPointers(self):
        """Edges inserted so far stop being reported by UpdatedEdges()
        (include/Graph/SlabGraphImpl.cuh:64-82)."""
        _lib.call("mk_graph_update_slab_pointers", self._h)

    def UpdatedEdges(self):
        """Device tensors (src, dst, w or None) of the edges inserted since
        the last UpdateSlabPointers and still present, in no particular
        order; the graph must have been created with track_updates=True."""
        torch = _torch()
        cnt = ctypes.c_uint64()
        _lib.call("mk_graph_updated_edges", self._h, None, None, None, 0, ctypes.byref(cnt))
        k = cnt.value
        dev = f"cuda:{self.device}"
        s = torch.empty(max(k, 1), dtype=torch.int32, device=dev)
        d = torch.empty(max(k, 1), dtype=torch.int32, device=dev)
        w = torch.empty(max(k, 1), dtype=torch.int32, device=dev) if self.weighted else None
        _lib.call("mk_graph_updated_edges", self._h, dptr(s), dptr(d), dptr(w), k,
                  ctypes.byref(cnt))
        if cnt.value != k:
            raise RuntimeError("updated edges changed between the two passes")
        return s[:k], d[:k], (None if w is None else w[:k])

    def set_deferred(self, on=True):
        """Deferred mutation checks: InsertEdges / DeleteEdges return once the
        batch is queued; its range and slab-pool outcome is raised by the
        graph's next call (or sync()), or by the BFS/SSSP call that follows
        it, which then leaves its handle invalid until Static()."""
        _lib.call("mk_graph_set_deferred", self._h, int(on))

    def sync(self):
        """Settle a deferred batch: raises its MeerkatError if it failed."""
        _lib.call("mk_graph_sync", self._h)

    insert_edges = InsertEdges
    delete_edges = DeleteEdges

    def last_status(self):
        """MK_EDGE_* code of every edge of the last insert/delete, as numpy."""
        p, n = _vp(), ctypes.c_uint32()
        _lib.call("mk_graph_last_status", self._h, ctypes.byref(p), ctypes.byref(n))
        host = np.empty(n.value, np.uint8)
        if n.value:
            _copy_d2h(host, p.value, n.value)
        return host

    def SearchEdges(self, src, dst):
        torch = _torch()
        s, d = device_u32(src, self.device), device_u32(dst, self.device)
        found = torch.empty(len(s), dtype=torch.uint8, device=s.device)
        w = torch.empty(len(s), dtype=torch.int32, device=s.device) if self.weighted else None
        _lib.call("mk_graph_search", self._h, dptr(s), dptr(d), len(s), dptr(found), dptr(w),
                  None)
        f = found.cpu().numpy().astype(bool)
        return (f, w.cpu().numpy().view(np.uint32)) if self.weighted else f

    # ---------------------------------------------------------- queries
    def num_edges(self):
        m = ctypes.c_uint64()
        _lib.call("mk_graph_num_edges", self._h, ctypes.byref(m))
        return m.value

    def degrees(self):
        out = np.empty(self.n, np.uint32)
        _lib.call("mk_graph_degrees", self._h, out.ctypes.data_as(_vp))
        return out

    def GetVertexBucketsOffsets(self):
        out = np.empty(self.n + 1, np.uint32)
        _lib.call("mk_graph_bucket_offsets", self._h, out.ctypes.data_as(_vp))
        return out

    def GetBucketsPerVertex(self):
        return np.diff(self.GetVertexBucketsOffsets())

    def count_slots(self):
        live = np.empty(self.n, np.uint32)
        tomb = np.empty(self.n, np.uint32)
        _lib.call("mk_graph_count_slots", self._h, live.ctypes.data_as(_vp),
                  tomb.ctypes.data_as(_vp))
        return live, tomb

    def export_csr(self):
        """(row_off uint64[n+1], col uint32[m], w uint32[m] or None), each row
        sorted by destination."""
        row = np.empty(self.n + 1, np.uint64)
        _lib.call("mk_graph_export_csr", self._h, row.ctypes.data_as(_vp), None, None)
        m = int(row[-1])
        col = np.empty(max(m, 1), np.uint32)
        w = np.empty(max(m, 1), np.uint32) if self.weighted else None
        _lib.call("mk_graph_export_csr", self._h, row.ctypes.data_as(_vp),
                  col.ctypes.data_as(_vp), None if w is None else w.ctypes.data_as(_vp))
        col, w = col[:m], (None if w is None else w[:m])
        # rows come out in slab order; sort each row by destination
        owner = np.repeat(np.arange(self.n, dtype=np.uint64), np.diff(row).astype(np.int64))
        order = np.lexsort((col, owner))
        col = col[order]
        if w is not None:
            w = w[order]
        return row, col, w

    def slab_stats(self):
        h, o, mx = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
        _lib.call("mk_graph_slab_stats", self._h, ctypes.byref(h), ctypes.byref(o),
                  ctypes.byref(mx))
        return {"head_slabs": h.value, "overflow_slabs": o.value, "max_chain": mx.value}

    def close(self):
        if self._h and self._owned:
            _lib.call("mk_graph_destroy", self._h)
        self._h = _vp()

    def __del__(self):
        try:
            if _lib.finalizer_ok() and self._h and self._owned:
                _lib.lib().mk_graph_destroy(self._h)
        except Exception:
            pass


def _copy_d2h(host, dev_addr, nbytes):
    """Copy nbytes of device memory at dev_addr into a numpy array."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    rc = _hip.hipMemcpy(host.ctypes.data_as(_vp), _vp(dev_addr), nbytes, 2)  # device->host
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")


_hip = None
