"""Seeded workload generators (host side, from libmeerkat_hip.so).

They stand in for the text edge lists and ``<prefix>.batch.<i>`` files the
reference drivers read (src/BFS.cu:23-36, 76-91).  All arrays are numpy
uint32; edge keys are uint64 ``src << 32 | dst``.
"""
import ctypes

import numpy as np

from . import _lib

RMAT_ABC = (0.57, 0.19, 0.19)  # d = 0.05


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _fetch(m, weighted, wseed):
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    w = np.empty(m, np.uint32) if weighted else None
    _lib.call("mk_gen_fetch", _p(src), _p(dst), _p(w), wseed)
    return src, dst, w


def rmat(scale, edge_factor=16, seed=1, undirected=False, weighted=False, wseed=None,
         abc=RMAT_ABC):
    """RMAT edge list over 2**scale vertices, sorted by (src, dst), unique,
    no self-loops.  Returns (n, src, dst, w-or-None)."""
    m = ctypes.c_uint64()
    a, b, c = abc
    _lib.call("mk_gen_rmat", scale, edge_factor, a, b, c, seed, int(undirected),
              ctypes.byref(m))
    src, dst, w = _fetch(m.value, weighted, seed + 1000003 if wseed is None else wseed)
    return 1 << scale, src, dst, w


def erdos_renyi(n, m, seed=5, weighted=False, wseed=None):
    """Directed G(n, m) with u != v, duplicates dropped."""
    out = ctypes.c_uint64()
    _lib.call("mk_gen_er", n, m, seed, ctypes.byref(out))
    src, dst, w = _fetch(out.value, weighted, seed + 1000003 if wseed is None else wseed)
    return n, src, dst, w


def keys_of(src, dst):
    return (src.astype(np.uint64) << np.uint64(32)) | dst.astype(np.uint64)


def split_keys(keys):
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(
        np.uint32)


def insert_batch(sorted_keys, n, count, seed, weighted=False, wseed=7):
    """`count` distinct new edges absent from `sorted_keys` (None: no check
    against existing edges)."""
    if sorted_keys is not None:
        sorted_keys = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    w = np.empty(count, np.uint32) if weighted else None
    _lib.call("mk_gen_insert_batch", _p(sorted_keys),
              0 if sorted_keys is None else len(sorted_keys), n, count, seed, _p(src), _p(dst),
              _p(w), wseed)
    return src, dst, w


def delete_batch(sorted_keys, count, seed):
    """`count` distinct edges sampled without replacement from `sorted_keys`."""
    sorted_keys = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    _lib.call("mk_gen_delete_batch", _p(sorted_keys), len(sorted_keys), count, seed, _p(src),
              _p(dst))
    return src, dst


def degree_hints(n, src):
    """Out-degree of every vertex in an edge list (the DynamicGraph hints)."""
    return np.bincount(src, minlength=n).astype(np.uint32)


def write_edge_file(path, n, src, dst):
    """Write a base graph in the text format the drivers read: "V E" then
    one "u v" line per edge (the reference's src/BFS.cu:27-36)."""
    src = np.asarray(src, np.uint32)
    dst = np.asarray(dst, np.uint32)
    with open(path, "w") as f:
        f.write(f"{int(n)} {len(src)}\n")
        np.savetxt(f, np.stack([src, dst], 1), fmt="%d")


def write_batch_file(prefix, i, src, dst):
    """Write update batch `i` as "<prefix>.batch.<i>": "n" then one "u v"
    line per edge (the reference's src/BFS.cu:76-91)."""
    src = np.asarray(src, np.uint32)
    dst = np.asarray(dst, np.uint32)
    with open(f"{prefix}.batch.{i}", "w") as f:
        f.write(f"{len(src)}\n")
        np.savetxt(f, np.stack([src, dst], 1), fmt="%d")
