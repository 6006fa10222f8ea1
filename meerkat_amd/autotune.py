"""Launch-shape knobs of the library and a measuring autotuner.

Knobs are read when an algorithm handle is created, so a change applies to
handles created afterwards:

- ``waves_per_cu`` (1..64, default 16): waves the grid-stride frontier,
  PageRank, CC and triangle kernels launch per compute unit.
- ``ms_unroll`` (4 or 8, default 8): neighbour distance rows each wave keeps
  in flight in multi-source SSSP (more rows hide more HBM latency at the
  cost of registers).
- ``pull_enter`` / ``pull_exit`` (defaults 14 / 24): the push/pull
  thresholds of static BFS over a graph with the in-edge mirror; a step runs
  bottom-up once the frontier holds more than 1/pull_enter of the buckets
  (0: never) and until it holds fewer than 1/pull_exit of the vertices.

The reference fixes its launch shapes at compile time (e.g.
include/Algorithms/impl/BFSImpl.cuh:620-625); these knobs exist because the
best shape on MI355X depends on the graph's degree skew.
"""
import numpy as np

from . import _lib

KNOBS = {"waves_per_cu": (1, 64), "ms_unroll": (4, 8), "orphan_repair": (0, 1),
         "persistent_steps": (0, 1), "tail_enter": (0, 1024), "coop_launch": (0, 1),
         "mut_waves_per_cu": (1, 256), "barrier_timeout_ms": (0, 600000),
         "pull_enter": (0, 4096), "pull_exit": (1, 4096)}


def set_knob(name, value):
    """Set one knob; raises MeerkatError for an unknown name or bad value."""
    _lib.call("mk_tune_set", name.encode(), int(value))


def get_knob(name):
    v = _lib.ctypes.c_int64()
    _lib.call("mk_tune_get", name.encode(), _lib.ctypes.byref(v))
    return int(v.value)


def knobs():
    return {k: get_knob(k) for k in KNOBS}


def _time(make, run, repeats):
    h = make()
    try:
        run(h)  # first run pays one-time setup (frontier buffers)
        return min(run(h) for _ in range(repeats))
    finally:
        h.close()


def tune(graph, sources=None, waves=(8, 16, 32), unrolls=(4, 8), repeats=2, apply=True):
    """Measure static BFS / SSSP and 64-source SSSP on `graph` for each
    setting and (with `apply`) keep the fastest.  Returns
    ``{"waves_per_cu": best, "ms_unroll": best, "timings": {...}}`` with
    milliseconds per setting.  The knobs are restored when `apply` is
    false."""
    from .algorithms import BFS, SSSP, MultiSource

    before = knobs()
    if sources is None:
        sources = np.random.default_rng(1).integers(0, graph.n, 64).astype(np.uint32)
    timings = {"waves_per_cu": {}, "ms_unroll": {}}
    chosen = dict(before)
    try:
        for w in waves:
            set_knob("waves_per_cu", w)
            t = _time(lambda: BFS(graph, int(sources[0])), lambda h: h.Static(), repeats)
            if graph.weighted:
                t += _time(lambda: SSSP(graph, int(sources[0])), lambda h: h.Static(), repeats)
            timings["waves_per_cu"][w] = t
        best_w = min(timings["waves_per_cu"], key=timings["waves_per_cu"].get)
        set_knob("waves_per_cu", best_w)
        best_u = before["ms_unroll"]
        if graph.weighted:
            for u in unrolls:
                set_knob("ms_unroll", u)
                timings["ms_unroll"][u] = _time(lambda: MultiSource(graph, sssp=True),
                                                lambda h: h.run(sources), repeats)
            best_u = min(timings["ms_unroll"], key=timings["ms_unroll"].get)
        if apply:
            chosen = {"waves_per_cu": best_w, "ms_unroll": best_u}
    finally:
        for k, v in chosen.items():
            set_knob(k, v)
    return {**knobs(), "timings": timings}
