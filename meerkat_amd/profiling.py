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


# ------------------------------------------------------------ rocprofv3 regions
_roctx = None


def _roctx_lib():
    """The roctx library, or False when it cannot be loaded (then profiler
    regions are no-ops)."""
    global _roctx
    if _roctx is None:
        _roctx = False
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                L = ctypes.CDLL(name)
                L.roctxProfilerPause.argtypes = [ctypes.c_uint64]
                L.roctxProfilerResume.argtypes = [ctypes.c_uint64]
                _roctx = L
                break
            except (OSError, AttributeError):
                continue
    return _roctx


class region:
    """Context manager marking a region for ``rocprofv3 --selected-regions``:
    collection resumes on entry and pauses on exit (roctxProfilerResume /
    roctxProfilerPause with thread id 0).  Without the profiler these calls
    do nothing."""

    def __enter__(self):
        L = _roctx_lib()
        if L:
            L.roctxProfilerResume(0)
        return self

    def __exit__(self, *exc):
        L = _roctx_lib()
        if L:
            L.roctxProfilerPause(0)
        return False


# ------------------------------------------------------------ rocprofv3 output
def _read_csv(path):
    import csv

    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def kernel_stats(path, top=20):
    """Rows of a rocprofv3 ``*_kernel_stats.csv`` (``--stats``), busiest
    first: (name, calls, total_ms, avg_us, min_us, max_us, percent)."""
    rows = []
    for r in _read_csv(path):
        rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6,
                     float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                     float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    rows.sort(key=lambda x: -x[2])
    return rows[:top]


def pmc_per_kernel(path, counter):
    """Mean per-dispatch value of `counter` for each kernel name in a
    rocprofv3 ``*_counter_collection.csv``: {name: (dispatches, mean)}."""
    acc = {}
    for r in _read_csv(path):
        if r.get("Counter_Name") != counter:
            continue
        k = r["Kernel_Name"]
        n, s = acc.get(k, (0, 0.0))
        acc[k] = (n + 1, s + float(r["Counter_Value"]))
    return {k: (n, s / n) for k, (n, s) in acc.items()}


def _main(argv):
    import argparse

    ap = argparse.ArgumentParser(prog="python -m meerkat_amd.profiling",
                                 description="Summaries of rocprofv3 CSV output")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats", help="kernel time table from *_kernel_stats.csv")
    a.add_argument("csv")
    a.add_argument("--top", type=int, default=20)
    b = sub.add_parser("pmc", help="mean per-dispatch counter value per kernel")
    b.add_argument("csv")
    b.add_argument("--counter", required=True)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(f"{'calls':>6} {'total_ms':>10} {'avg_us':>10} {'min_us':>9} {'max_us':>10} "
              f"{'%':>6}  kernel")
        for name, calls, tot, avg, mn, mx, pct in kernel_stats(args.csv, args.top):
            print(f"{calls:6d} {tot:10.3f} {avg:10.2f} {mn:9.2f} {mx:10.2f} {pct:6.2f}  {name}")
    else:
        print(f"{'dispatches':>10} {'mean ' + args.counter:>20}  kernel")
        for name, (n, mean) in sorted(pmc_per_kernel(args.csv, args.counter).items(),
                                      key=lambda kv: -kv[1][1] * kv[1][0]):
            print(f"{n:10d} {mean:20.1f}  {name}")


if __name__ == "__main__":
    import sys

    _main(sys.argv[1:])
