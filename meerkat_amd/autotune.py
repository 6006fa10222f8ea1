"""Launch-shape knobs of the library and a measuring autotuner.

Knobs are read when an algorithm handle is created, so a change applies to
handles created afterwards:

- ``waves_per_cu`` (1..64, default 16): waves the grid-stride frontier,
  CC and triangle kernels launch per compute unit.
- ``ms_unroll`` (4 or 8, default 8): neighbour distance rows each wave keeps
  in flight in multi-source SSSP (more rows hide more HBM latency at the
  cost of registers).
- ``ms_waves_per_cu`` (1..64, default 24): waves per compute unit of the
  multi-source BFS / SSSP steps (random row reads: more waves keep more in
  flight, up to what a CU holds).
- ``pr_waves_per_cu`` (1..64, default 32): waves per compute unit of the
  PageRank iteration, also capped at what a CU holds.
- ``pull_enter`` / ``pull_exit`` (defaults 14 / 24): the push/pull
  thresholds of static BFS over a graph with the in-edge mirror; a step runs
  bottom-up once the frontier holds more than 1/pull_enter of the buckets
  (0: never) and until it holds fewer than 1/pull_exit of the vertices.
- ``sssp_delta`` (default 0): static SSSP runs near-far with this bucket
  width (vertices at or past the near bound wait in a far list until the
  near frontier drains), cutting the re-relaxations of label-correcting;
  ``sssp_far_cap`` bounds each far list (0: one entry per vertex).

The reference fixes its launch shapes at compile time (e.g.
include/Algorithms/impl/BFSImpl.cuh:620-625); these knobs exist because the
best shape on MI355X depends on the graph's degree skew.
"""
import numpy as np

from . import _lib

KNOBS = {"waves_per_cu": (1, 64), "ms_unroll": (4, 8), "ms_waves_per_cu": (1, 64),
         "pr_waves_per_cu": (1, 64), "orphan_repair": (0, 1),
         "persistent_steps": (0, 1), "tail_enter": (0, 1024), "coop_launch": (0, 1),
         "mut_waves_per_cu": (1, 256), "barrier_timeout_ms": (0, 600000),
         "pull_enter": (0, 4096), "pull_exit": (1, 4096), "sssp_delta": (0, 1 << 24),
         "sssp_far_cap": (0, 0xFFFFFFF0)}


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


def tune(graph, sources=None, waves=(8, 16, 32), unrolls=(4, 8),
         pulls=((14, 24), (32, 24), (8, 48)), deltas=(0, 32, 96), tails=(64, 128, 256),
         ms_waves=(16, 24), repeats=2, apply=True):
    """Measure on `graph` and (with `apply`) keep the fastest setting of
    each knob in turn, the ones already chosen held:

    - ``waves_per_cu`` over `waves`: static BFS + static SSSP;
    - the push/pull thresholds ``(pull_enter, pull_exit)`` over `pulls`:
      static BFS (graphs with the in-edge mirror only);
    - ``sssp_delta`` over `deltas`: static SSSP (weighted graphs);
    - ``tail_enter`` over `tails`: static BFS + SSSP (the frontier's small
      steps);
    - ``ms_unroll`` over `unrolls`, then ``ms_waves_per_cu`` over
      `ms_waves`: a 64-source SSSP (weighted graphs).

    Returns the knobs after the call with ``"timings"``: milliseconds per
    setting of each knob measured.  The knobs are restored when `apply` is
    false.  Static calls run from ``sources[0]``; pass sources with
    out-edges (a source without any is a trivial traversal)."""
    from .algorithms import BFS, SSSP, MultiSource

    before = knobs()
    if sources is None:
        sources = np.random.default_rng(1).integers(0, graph.n, 64).astype(np.uint32)
    src = int(sources[0])

    def bfs():
        return _time(lambda: BFS(graph, src), lambda h: h.Static(), repeats)

    def sssp():
        return _time(lambda: SSSP(graph, src), lambda h: h.Static(), repeats)

    def both():
        return bfs() + (sssp() if graph.weighted else 0.0)

    def msssp():
        return _time(lambda: MultiSource(graph, sssp=True), lambda h: h.run(sources), repeats)

    reverse = bool(getattr(graph, "reverse", False))
    # (knob or knob pair, candidates, objective), in order
    plan = [(("waves_per_cu",), [(w,) for w in waves], both)]
    if reverse:
        plan.append((("pull_enter", "pull_exit"), [tuple(p) for p in pulls], bfs))
    if graph.weighted:
        plan.append((("sssp_delta",), [(x,) for x in deltas], sssp))
    plan.append((("tail_enter",), [(t,) for t in tails], both))
    if graph.weighted:
        plan.append((("ms_unroll",), [(u,) for u in unrolls], msssp))
        plan.append((("ms_waves_per_cu",), [(w,) for w in ms_waves], msssp))
    timings = {}
    try:
        for names, cands, objective in plan:
            key = "/".join(names)
            timings[key] = {}
            for c in cands:
                for k, v in zip(names, c):
                    set_knob(k, v)
                timings[key][c if len(c) > 1 else c[0]] = objective()
            best = min(timings[key], key=timings[key].get)
            best = best if isinstance(best, tuple) else (best,)
            for k, v in zip(names, best):
                set_knob(k, v)
    finally:
        if not apply:
            for k, v in before.items():
                set_knob(k, v)
    return {**knobs(), "timings": timings}


def tune_load_factor(n, src, dst, w=None, factors=(0.5, 0.7, 0.9), reverse=False, source=0,
                     pool_bytes=1 << 30, repeats=2):
    """The load factor of a graph is fixed when it is built (the reference's
    DynamicGraph argument, include/Graph/SlabGraph.cuh:382-384): build the
    edge list (src, dst[, w]) at each of `factors` and time the build
    (insert rate) and static BFS (+ SSSP when weighted) from `source`.
    Returns ``{"load_factor": fastest traversal, "timings": {lf: {...}}}``;
    the build rate is reported beside it, since a smaller load factor
    spends more head slabs (memory) for shorter chains."""
    import time

    import torch

    from . import gen
    from .algorithms import BFS, SSSP
    from .graph import DynamicGraph, SlabPool, device_u32

    weighted = w is not None
    S, D = device_u32(src), device_u32(dst)
    Wd = device_u32(w) if weighted else None
    hints, in_hints = gen.degree_hints(n, src), gen.degree_hints(n, dst)
    timings = {}
    for lf in factors:
        pool = SlabPool(pool_bytes)
        g = DynamicGraph(n, pool, lf, hints, weighted=weighted, reverse=reverse,
                         in_degree_hints=in_hints)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.InsertEdges(S, D, Wd)
        torch.cuda.synchronize()
        build = time.perf_counter() - t0
        t = _time(lambda: BFS(g, source), lambda h: h.Static(), repeats)
        if weighted:
            t += _time(lambda: SSSP(g, source), lambda h: h.Static(), repeats)
        timings[lf] = {"traversal_ms": t, "build_edges_per_s": len(src) / build,
                       "slabs": g.slab_stats()}
        g.close()
        pool.close()
    best = min(timings, key=lambda lf: timings[lf]["traversal_ms"])
    return {"load_factor": best, "timings": timings}
