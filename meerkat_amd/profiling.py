"""Per-kernel timing from the library's HIP-event profiler (mk_profile_*)."""
import ctypes

from . import _lib


def enable(on=True):
    _lib.call("mk_profile_enable", int(on))


def reset():
    _lib.call("mk_profile_reset")


def totals():
    """{kernel: {"launches", "ms", "algo_bytes", "avg_ms", "achieved_GBps"}}."""
    n = ctypes.c_int()
    _lib.call("mk_profile_count", ctypes.byref(n))
    out = {}
    for i in range(n.value):
        name = ctypes.c_char_p()
        launches, ms, nbytes = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
        _lib.call("mk_profile_get", i, ctypes.byref(name), ctypes.byref(launches),
                  ctypes.byref(ms), ctypes.byref(nbytes))
        k = name.value.decode()
        L = launches.value
        out[k] = {
            "launches": L,
            "ms": ms.value,
            "algo_bytes": nbytes.value,
            "avg_ms": ms.value / L if L else 0.0,
            "achieved_GBps": (nbytes.value / (ms.value * 1e-3) / 1e9) if ms.value > 0 else 0.0,
        }
    return out
