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


# ------------------------------------------------------------ profile markers
MARKER_KERNEL = "FillFunctor<short>"


def marker(device=0):
    """Launch a one-element int16 fill kernel and wait for it: a dispatch no
    other code here launches, so pairs of markers delimit a region of a
    rocprofv3 trace or counter collection (see ``region_rows``)."""
    import torch

    torch.zeros(1, dtype=torch.int16, device=device)
    torch.cuda.synchronize()


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


def region_rows(rows, region, marker=MARKER_KERNEL):
    """Rows of a rocprofv3 per-dispatch CSV (kernel trace or counter
    collection) strictly between the (2*region)-th and (2*region+1)-th
    marker dispatches, in dispatch order.  region None: all rows."""
    if region is None:
        return rows
    rows = sorted(rows, key=lambda r: int(r["Dispatch_Id"]))
    ids = []
    for r in rows:
        if marker in r["Kernel_Name"] and (not ids or ids[-1] != r["Dispatch_Id"]):
            ids.append(r["Dispatch_Id"])
    if len(ids) < 2 * region + 2:
        raise ValueError(f"only {len(ids)} marker dispatches; region {region} needs "
                         f"{2 * region + 2}")
    lo, hi = int(ids[2 * region]), int(ids[2 * region + 1])
    return [r for r in rows if lo < int(r["Dispatch_Id"]) < hi]


def trace_stats(path, region=None):
    """Per-kernel (dispatches, total_ms, avg_us, min_us, max_us) from a
    rocprofv3 ``*_kernel_trace.csv``, restricted to a marker region."""
    acc = {}
    for r in region_rows(_read_csv(path), region):
        dt = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        n, tot, mn, mx = acc.get(r["Kernel_Name"], (0, 0.0, float("inf"), 0.0))
        acc[r["Kernel_Name"]] = (n + 1, tot + dt, min(mn, dt), max(mx, dt))
    out = [(k, n, tot / 1e3, tot / n, mn, mx) for k, (n, tot, mn, mx) in acc.items()]
    out.sort(key=lambda x: -x[2])
    return out


def pmc_per_kernel(path, counter, region=None):
    """Mean per-dispatch value of `counter` for each kernel name in a
    rocprofv3 ``*_counter_collection.csv`` (optionally restricted to a marker
    region): {name: (dispatches, mean)}."""
    acc = {}
    for r in region_rows(_read_csv(path), region):
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
    t = sub.add_parser("trace", help="kernel time table from *_kernel_trace.csv")
    t.add_argument("csv")
    t.add_argument("--region", type=int, default=None,
                   help="only dispatches between marker pair REGION (0 = first)")
    t.add_argument("--top", type=int, default=20)
    b = sub.add_parser("pmc", help="mean per-dispatch counter value per kernel")
    b.add_argument("csv")
    b.add_argument("--counter", required=True)
    b.add_argument("--region", type=int, default=None)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(f"{'calls':>6} {'total_ms':>10} {'avg_us':>10} {'min_us':>9} {'max_us':>10} "
              f"{'%':>6}  kernel")
        for name, calls, tot, avg, mn, mx, pct in kernel_stats(args.csv, args.top):
            print(f"{calls:6d} {tot:10.3f} {avg:10.2f} {mn:9.2f} {mx:10.2f} {pct:6.2f}  {name}")
    elif args.cmd == "trace":
        print(f"{'calls':>6} {'total_ms':>10} {'avg_us':>10} {'min_us':>9} {'max_us':>10}  kernel")
        for name, calls, tot, avg, mn, mx in trace_stats(args.csv, args.region)[:args.top]:
            print(f"{calls:6d} {tot:10.3f} {avg:10.2f} {mn:9.2f} {mx:10.2f}  {name}")
    else:
        print(f"{'dispatches':>10} {'mean ' + args.counter:>20}  kernel")
        for name, (n, mean) in sorted(pmc_per_kernel(args.csv, args.counter,
                                                     args.region).items(),
                                      key=lambda kv: -kv[1][1] * kv[1][0]):
            print(f"{n:10d} {mean:20.1f}  {name}")


if __name__ == "__main__":
    import sys

    _main(sys.argv[1:])
