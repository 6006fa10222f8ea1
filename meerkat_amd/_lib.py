"""ctypes binding of libmeerkat_hip.so (the C ABI declared in include/meerkat.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C meerkat_amd/csrc``).  There is no fallback: if the library is
missing or fails to load, every entry point raises.
"""
import atexit
import ctypes
import os
import re
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmeerkat_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "meerkat.h")

MK_OK = 0
MK_ERR_BAD_ARG = 1
MK_ERR_POOL_FULL = 2
MK_ERR_HIP = 3
MK_ERR_STATE = 4
MK_ERR_RANGE = 5

EDGE_UNCHANGED = 0
EDGE_APPLIED = 1
EDGE_SELF_LOOP = 2
EDGE_POOL_FULL = 3

_ERR_NAMES = {
    MK_ERR_BAD_ARG: "MK_ERR_BAD_ARG",
    MK_ERR_POOL_FULL: "MK_ERR_POOL_FULL",
    MK_ERR_HIP: "MK_ERR_HIP",
    MK_ERR_STATE: "MK_ERR_STATE",
    MK_ERR_RANGE: "MK_ERR_RANGE",
}


class MeerkatError(RuntimeError):
    """A non-zero status from the C ABI; ``code`` holds the MK_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None
_exiting = False


@atexit.register
def _at_exit():
    global _exiting
    _exiting = True


def finalizer_ok():
    """Whether a ``__del__`` may still call into the library.  Not once the
    interpreter is exiting: the HIP runtime (and a profiler's preloaded
    tool) tear down then, and a handle left open is released with the
    process.  Scripts close their handles explicitly (``close()``)."""
    return not _exiting and not sys.is_finalizing()


vp = ctypes.c_void_p
u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int
f64 = ctypes.c_double
f32 = ctypes.c_float
P = ctypes.POINTER

# name -> argtypes (restype is int unless listed in _RESTYPES)
_PROTOS = {
    "mk_last_error": [],
    "mk_version": [],
    "mk_device_info": [i32, P(i32), P(i32)],
    "mk_pool_create": [u64, i32, P(vp)],
    "mk_pool_destroy": [vp],
    "mk_pool_stats": [vp, P(u64), P(u64)],
    "mk_graph_create": [vp, u32, i32, f64, vp, P(vp)],
    "mk_graph_create_ex": [vp, u32, u32, f64, vp, vp, P(vp)],
    "mk_graph_reverse": [vp, P(vp)],
    "mk_graph_destroy": [vp],
    "mk_graph_insert": [vp, vp, vp, vp, u32, vp],
    "mk_graph_delete": [vp, vp, vp, u32, vp],
    "mk_graph_set_deferred": [vp, i32],
    "mk_graph_sync": [vp],
    "mk_graph_update_weights": [vp, vp, vp, vp, u32, vp],
    "mk_graph_update_slab_pointers": [vp],
    "mk_graph_updated_edges": [vp, vp, vp, vp, u64, P(u64)],
    "mk_graph_search": [vp, vp, vp, u32, vp, vp, vp],
    "mk_graph_last_status": [vp, P(vp), P(u32)],
    "mk_graph_num_vertices": [vp, P(u32)],
    "mk_graph_num_edges": [vp, P(u64)],
    "mk_graph_is_weighted": [vp, P(i32)],
    "mk_graph_degrees": [vp, vp],
    "mk_graph_bucket_offsets": [vp, vp],
    "mk_graph_count_slots": [vp, vp, vp],
    "mk_graph_count_slots_noalloc": [vp, vp, vp],
    "mk_graph_export_csr": [vp, vp, vp, vp],
    "mk_graph_slab_stats": [vp, P(u64), P(u64), P(u32)],
    "mk_bfs_create": [vp, u32, P(vp)],
    "mk_bfs_destroy": [vp],
    "mk_bfs_static": [vp, P(f32)],
    "mk_bfs_incremental": [vp, vp, vp, u32, P(f32)],
    "mk_bfs_decremental": [vp, vp, vp, u32, P(f32)],
    "mk_bfs_read": [vp, vp, vp],
    "mk_bfs_set_source": [vp, u32],
    "mk_bfs_last_stats": [vp, P(u32), P(u64), P(u64)],
    "mk_bfs_set_deferred": [vp, i32],
    "mk_bfs_sync": [vp],
    "mk_bfs_deferred_stats": [vp, P(u64), P(u64)],
    "mk_sssp_create": [vp, u32, P(vp)],
    "mk_sssp_destroy": [vp],
    "mk_sssp_static": [vp, P(f32)],
    "mk_sssp_incremental": [vp, vp, vp, vp, u32, P(f32)],
    "mk_sssp_decremental": [vp, vp, vp, vp, u32, P(f32)],
    "mk_sssp_reweight": [vp, vp, vp, vp, u32, P(f32)],
    "mk_sssp_read": [vp, vp, vp],
    "mk_sssp_set_source": [vp, u32],
    "mk_sssp_last_stats": [vp, P(u32), P(u64), P(u64)],
    "mk_sssp_set_deferred": [vp, i32],
    "mk_sssp_sync": [vp],
    "mk_sssp_deferred_stats": [vp, P(u64), P(u64)],
    "mk_sssp_capture_step": [vp, vp, vp, u32, vp, vp, vp, u32, P(vp)],
    "mk_bfs_capture_step": [vp, vp, vp, u32, vp, vp, u32, P(vp)],
    "mk_step_launch": [vp],
    "mk_step_load": [vp, vp, vp, vp, vp, vp],
    "mk_step_sync": [vp],
    "mk_step_destroy": [vp],
    "mk_pagerank_create": [vp, f32, f32, u32, vp, P(vp)],
    "mk_pagerank_destroy": [vp],
    "mk_pagerank_static": [vp, P(f32)],
    "mk_pagerank_incremental": [vp, vp, vp, u32, P(f32)],
    "mk_pagerank_decremental": [vp, vp, vp, u32, P(f32)],
    "mk_pagerank_read": [vp, vp],
    "mk_pagerank_last_stats": [vp, P(u32), P(f64)],
    "mk_tc_create": [vp, P(vp)],
    "mk_tc_destroy": [vp],
    "mk_tc_static": [vp, P(f32)],
    "mk_tc_incremental": [vp, vp, vp, u32, P(f32)],
    "mk_tc_decremental": [vp, vp, vp, u32, P(f32)],
    "mk_tc_count": [vp, P(u64)],
    "mk_cc_create": [vp, P(vp)],
    "mk_cc_destroy": [vp],
    "mk_cc_static": [vp, P(f32)],
    "mk_cc_incremental": [vp, vp, vp, u32, P(f32)],
    "mk_cc_decremental": [vp, vp, vp, u32, P(f32)],
    "mk_cc_labels": [vp, vp],
    "mk_msq_create": [vp, i32, P(vp)],
    "mk_msq_destroy": [vp],
    "mk_msq_run": [vp, vp, u32, P(f32)],
    "mk_msq_read": [vp, u32, vp],
    "mk_msq_last_stats": [vp, P(u32), P(u64)],
    "mk_msq_open": [vp],
    "mk_msq_admit": [vp, vp, vp, u32],
    "mk_msq_launch": [vp, u32],
    "mk_msq_poll": [vp, P(u64), P(u32)],
    "mk_msq_ready": [vp, P(i32)],
    "mk_tune_set": [ctypes.c_char_p, ctypes.c_int64],
    "mk_tune_get": [ctypes.c_char_p, P(ctypes.c_int64)],
    "mk_profile_enable": [i32],
    "mk_profile_reset": [],
    "mk_profile_count": [P(i32)],
    "mk_profile_get": [i32, P(ctypes.c_char_p), P(u64), P(f64), P(f64)],
    "mk_profile_step_count": [P(i32)],
    "mk_profile_step_get": [i32, P(ctypes.c_char_p), P(u32), P(u32), P(u32), P(f64)],
    "mk_gen_rmat": [u32, u32, f64, f64, f64, u64, i32, P(u64)],
    "mk_gen_er": [u32, u64, u64, P(u64)],
    "mk_gen_fetch": [vp, vp, vp, u64],
    "mk_gen_insert_batch": [vp, u64, u32, u32, u64, vp, vp, vp, u64],
    "mk_gen_delete_batch": [vp, u64, u32, u64, vp, vp],
}
_RESTYPES = {"mk_last_error": ctypes.c_char_p}


def header_symbols():
    """Function names declared in include/meerkat.h."""
    with open(HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mk_[a-z0-9_]+)\s*\(", text)))


def lib():
    """The loaded library; raises OSError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() "
                "or make -C meerkat_amd/csrc"
            )
        try:  # torch carries its own copy of the HIP runtime: it must find
            # the device before the library's runtime is loaded (the other
            # order leaves one of the two without a device)
            import torch

            torch.cuda.is_available()
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _PROTOS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def last_error():
    s = lib().mk_last_error()
    return s.decode() if s else ""


def check(rc):
    """Raise MeerkatError for a non-zero status."""
    if rc != MK_OK:
        raise MeerkatError(rc, last_error())
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))
