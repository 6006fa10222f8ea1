"""Per-kernel timing from the library's HIP-event profiler (mk_profile_*)."""
import ctypes

from . import _lib


def enable(on=True, loops_only=False):
    """Kernel timing on/off; loops_only times only the persistent BFS/SSSP
    frontier loops (by their own device clock, nearly free), not every
    kernel (a pair of dispatch events each)."""
    _lib.call("mk_profile_enable", (2 if loops_only else 1) if on else 0)


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


def steps():
    """Frontier steps of the persistent BFS/SSSP loops recorded while
    profiling: [{"kernel", "step", "frontier", "hub_items", "us"}], in
    order (step 0 starts each loop launch)."""
    n = ctypes.c_int()
    _lib.call("mk_profile_step_count", ctypes.byref(n))
    out = []
    for i in range(n.value):
        name = ctypes.c_char_p()
        step, small, hubs, us = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        _lib.call("mk_profile_step_get", i, ctypes.byref(name), ctypes.byref(step),
                  ctypes.byref(small), ctypes.byref(hubs), ctypes.byref(us))
        out.append({"kernel": name.value.decode(), "step": step.value, "frontier": small.value,
                    "hub_items": hubs.value, "us": us.value})
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


# Names the library's HIP-event profiler uses for the kernels a rocprofv3
# trace names by their full signature (the bench reports the former).
SCOPE_OF = [
    ("k_expand<true, true, 0>", "k_expand_relax"),
    ("k_expand<true, true, 1>", "k_expand_invalidate"),
    ("k_expand_loop<true, true, 0>", "k_expand_relax"),
    ("k_expand_loop<true, true, 1>", "k_expand_invalidate"),
    ("k_insert<", "k_insert"),
    ("k_delete<", "k_delete"),
    ("k_seed_batch<", "k_seed_batch"),
    ("k_invalidate_batch(", "k_invalidate_batch"),
    ("k_reseed_pull<", "k_reseed_pull"),
    ("k_pull_hubs<", "k_pull_hubs"),
]


def scope_name(kernel):
    for pat, name in SCOPE_OF:
        if pat in kernel:
            return name
    return None


def traffic(fetch_csv, write_csv, region=0, calib_region=1, calib_bytes=1 << 30):
    """HBM bytes per dispatch of each kernel in marker region `region`, from
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) collected in two runs, each
    scaled by the factor that makes the calibration copy in `calib_region`
    (reading and writing `calib_bytes`) read exactly.  Returns ({kernel:
    (fetch_bytes, write_bytes, dispatches)}, fetch_factor, write_factor)."""
    f = pmc_per_kernel(fetch_csv, "FETCH_SIZE", region)
    w = pmc_per_kernel(write_csv, "WRITE_SIZE", region)
    fc = list(pmc_per_kernel(fetch_csv, "FETCH_SIZE", calib_region).values())
    wc = list(pmc_per_kernel(write_csv, "WRITE_SIZE", calib_region).values())
    if len(fc) != 1 or len(wc) != 1:
        raise ValueError("the calibration region must hold exactly one kernel")
    ff = calib_bytes / (fc[0][1] * 1024.0)
    wf = calib_bytes / (wc[0][1] * 1024.0)
    out = {}
    for k in set(f) | set(w):
        fb = f.get(k, (0, 0.0))[1] * 1024.0 * ff
        wb = w.get(k, (0, 0.0))[1] * 1024.0 * wf
        out[k] = (fb, wb, max(f.get(k, (0, 0))[0], w.get(k, (0, 0))[0]))
    return out, ff, wf


def report(trace_csv, fetch_csv, write_csv, commands, top=12, program="bench.py"):
    """Text summary of the timed region (marker pair 0): kernel times from the
    kernel trace and HBM traffic per dispatch from the counter passes."""
    lines = [f"# rocprofv3 summary: timed region of {program} (between marker pair 0)", ""]
    lines += ["Commands:"] + [f"  {c}" for c in commands] + [""]
    lines.append("Kernel trace, timed region:")
    lines.append(f"{'calls':>6} {'total_ms':>10} {'avg_us':>9} {'min_us':>8} {'max_us':>9}  kernel")
    for name, n, tot, avg, mn, mx in trace_stats(trace_csv, 0)[:top]:
        lines.append(f"{n:6d} {tot:10.3f} {avg:9.2f} {mn:8.2f} {mx:9.2f}  {name[:110]}")
    tr, ff, wf = traffic(fetch_csv, write_csv)
    lines += ["", f"Counter scale from the 1 GiB calibration copy (marker pair 1): "
              f"FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {wf:.3f}", "",
              "HBM traffic per dispatch, timed region (scaled):",
              f"{'dispatches':>10} {'fetch_MB':>10} {'write_MB':>10} {'total_MB':>10}  kernel"]
    machine = []
    for k, (fb, wb, n) in sorted(tr.items(), key=lambda kv: -(kv[1][0] + kv[1][1]) * kv[1][2]):
        if fb + wb < 1e5:
            continue
        lines.append(f"{n:10d} {fb / 1e6:10.3f} {wb / 1e6:10.3f} {(fb + wb) / 1e6:10.3f}  {k[:100]}")
        sc = scope_name(k)
        if sc:
            machine.append(f"{sc} {int(round(fb + wb))}")
    return "\n".join(lines) + "\n", "\n".join(machine) + "\n"


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
    rp = sub.add_parser("report", help="timed-region summary and traffic table")
    rp.add_argument("--trace", required=True)
    rp.add_argument("--fetch", required=True)
    rp.add_argument("--write", required=True)
    rp.add_argument("--command", action="append", default=[])
    rp.add_argument("--out", required=True, help="summary text file")
    rp.add_argument("--traffic-out", required=True, help="'<kernel> <bytes>' lines")
    rp.add_argument("--program", default="bench.py", help="what the trace is of (title)")
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
    elif args.cmd == "report":
        text, machine = report(args.trace, args.fetch, args.write, args.command,
                               program=args.program)
        with open(args.out, "w") as f:
            f.write(text)
        with open(args.traffic_out, "w") as f:
            f.write(machine)
        print(text)
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
