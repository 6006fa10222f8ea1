"""The hot path's other rows at their full sizes, one per invocation, for
kernel traces and counter passes (rocprofv3) beside bench.py's update step.

    python -m meerkat_amd.workloads pagerank   # ER 4,194,304 x 33,554,432: static PageRank,
                                               # then 3 x (1,000,000 inserts + incremental)
    python -m meerkat_amd.workloads tc_cc      # RMAT-22 undirected: static TC and CC, then
                                               # 3 x (500,000 undirected inserts + incremental)
    python -m meerkat_amd.workloads static     # RMAT-24 weighted with the in-edge mirror:
                                               # static BFS and SSSP from 8 sources
    python -m meerkat_amd.workloads multi      # RMAT-24: 64-source BFS and SSSP, twice each
    python -m meerkat_amd.workloads weights    # RMAT-24 weighted, mirror, update tracking:
                                               # 3 x (1,000,000 weight changes + SSSP
                                               # reweight), then 3 x (1,000,000 inserts,
                                               # the graph's updated edges, incremental SSSP)

Launch knobs (mk_tune_set) may follow as name=value, e.g. sssp_delta=64.

Inputs are built and staged in HBM before the measured region, which is
delimited by marker pair 0 (profiling.marker); a 1 GiB copy between marker
pair 1 follows, the calibration of rocprofv3's FETCH_SIZE / WRITE_SIZE
counters.  Prints one JSON line: the calls' own times, and the library's
algorithmic bytes and achieved GB/s per kernel over the region (its kernel
timing, profiling level 1).  Results are checked by the tests, not here.
"""
import json
import sys
import time

import numpy as np


def _sync():
    import torch

    torch.cuda.synchronize()


def _region(fn):
    """Run fn() between marker pair 0 with every kernel timed; returns
    (fn's result, wall seconds, per-kernel totals)."""
    import torch

    from . import profiling

    dev = torch.device("cuda:0")
    _sync()
    profiling.reset()
    profiling.enable(True)
    profiling.marker(dev)
    t0 = time.perf_counter()
    out = fn()
    _sync()
    el = time.perf_counter() - t0
    profiling.marker(dev)
    profiling.enable(False)
    tot = profiling.totals()
    # calibration copy (marker pair 1)
    src = torch.ones((1 << 30) // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    profiling.marker(dev)
    torch.add(src, 0.0, out=dst)
    profiling.marker(dev)
    del src, dst
    kern = {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                "algo_MB": round(v["algo_bytes"] / 1e6, 3),
                "GBps": round(v["achieved_GBps"], 1)} for k, v in tot.items()}
    return out, el, kern


def pagerank():
    from . import gen
    from .algorithms import PageRank
    from .graph import DynamicGraph, SlabPool, device_u32

    n, s, d, _ = gen.erdos_renyi(4_194_304, 33_554_432, seed=5)
    pool = SlabPool(1 << 32)
    rev = DynamicGraph(n, pool, 0.7, gen.degree_hints(n, d))
    rev.InsertEdges(d, s)
    pr = PageRank(rev, 0.85, 1e-7, out_degree=gen.degree_hints(n, s))
    pr.Static()  # loads the kernels
    keys = gen.keys_of(s, d)
    batches = []
    for i in range(3):
        bs, bd, _ = gen.insert_batch(keys, n, 1_000_000, seed=6 * 1000 + i)
        batches.append((device_u32(bs), device_u32(bd)))
        keys = np.union1d(keys, gen.keys_of(bs, bd))

    def work():
        rec = {"static_ms": pr.Static(), "static_iterations": pr.last_stats()["iterations"],
               "incremental_ms": [], "iterations": []}
        for bs, bd in batches:
            rev.InsertEdges(bd, bs)
            rec["incremental_ms"].append(pr.Incremental(bs, bd))
            rec["iterations"].append(pr.last_stats()["iterations"])
        return rec

    rec, el, kern = _region(work)
    pr.close()
    rev.close()
    pool.close()
    return {"workload": "er4m_pagerank_inserts", "vertices": n, "edges": len(s), **rec,
            "region_s": round(el, 4), "kernels": kern}


def tc_cc():
    from . import gen
    from .algorithms import CC, TC
    from .graph import DynamicGraph, SlabPool, device_u32

    n, s, d, _ = gen.rmat(22, 16, seed=7, undirected=True)
    pool = SlabPool(1 << 33)
    g = DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    tc, cc = TC(g), CC(g)
    cc.Static()  # loads the kernels
    keys = gen.keys_of(s, d)
    batches = []
    for i in range(3):
        bs, bd, _ = gen.insert_batch(keys, n, 500_000, seed=8 * 1000 + i)
        k = np.unique(np.r_[gen.keys_of(bs, bd), gen.keys_of(bd, bs)])
        S, D = gen.split_keys(k)
        batches.append((device_u32(S), device_u32(D)))
        keys = np.union1d(keys, k)

    def work():
        rec = {"static_tc_ms": tc.Static(), "static_cc_ms": cc.Static(),
               "incremental_tc_ms": [], "incremental_cc_ms": []}
        for S, D in batches:
            g.InsertEdges(S, D)
            rec["incremental_tc_ms"].append(tc.Incremental(S, D))
            rec["incremental_cc_ms"].append(cc.Incremental(S, D))
        rec["triangles"] = int(tc.GetCount())
        return rec

    rec, el, kern = _region(work)
    cc.close()
    tc.close()
    g.close()
    pool.close()
    return {"workload": "rmat22_undirected_tc_cc_inserts", "vertices": n,
            "directed_entries": len(s), **rec, "region_s": round(el, 4), "kernels": kern}


def _rmat24_mirrored():
    from . import gen
    from .graph import DynamicGraph, SlabPool

    n, s, d, w = gen.rmat(24, 16, seed=9, weighted=True)
    pool = SlabPool(16 << 30)
    g = DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                     in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    rng = np.random.default_rng(24)
    deg = np.bincount(s, minlength=n)
    # sources with out-edges (a query from a vertex without any is trivial)
    sources = rng.choice(np.nonzero(deg)[0], size=64, replace=False).astype(np.uint32)
    return n, len(s), pool, g, sources


def static():
    from .algorithms import BFS, SSSP

    n, m, pool, g, sources = _rmat24_mirrored()
    b, sp = BFS(g, int(sources[0])), SSSP(g, int(sources[0]))
    b.Static()  # load the kernels
    sp.Static()

    def work():
        rec = {"bfs_ms": [], "sssp_ms": [], "bfs_slabs": [], "sssp_edges": []}
        for src in sources[:8]:
            b.set_source(int(src))
            rec["bfs_ms"].append(b.Static())
            rec["bfs_slabs"].append(b.last_stats()["slabs_read"])
            sp.set_source(int(src))
            rec["sssp_ms"].append(sp.Static())
            rec["sssp_edges"].append(sp.last_stats()["edges_read"])
        return rec

    rec, el, kern = _region(work)
    sp.close()
    b.close()
    g.close()
    pool.close()
    return {"workload": "rmat24_static_bfs_sssp", "vertices": n, "edges": m,
            "sources": [int(x) for x in sources[:8]], **rec, "region_s": round(el, 4),
            "kernels": kern}


def multi():
    from .algorithms import MultiSource

    n, m, pool, g, sources = _rmat24_mirrored()
    mb, ms = MultiSource(g, sssp=False), MultiSource(g, sssp=True)
    mb.run(sources[:2])  # load the kernels
    ms.run(sources[:2])

    def work():
        rec = {"msbfs_ms": [], "mssp_ms": [], "msbfs_edges": [], "mssp_edges": []}
        for _ in range(2):
            rec["msbfs_ms"].append(mb.run(sources))
            rec["msbfs_edges"].append(mb.last_stats()["edges_read"])
            rec["mssp_ms"].append(ms.run(sources))
            rec["mssp_edges"].append(ms.last_stats()["edges_read"])
        return rec

    rec, el, kern = _region(work)
    ms.close()
    mb.close()
    g.close()
    pool.close()
    return {"workload": "rmat24_multi_source_64", "vertices": n, "edges": m, **rec,
            "region_s": round(el, 4), "kernels": kern}


def weights():
    from . import gen
    from .algorithms import SSSP
    from .graph import DynamicGraph, SlabPool, device_u32

    n, s, d, w = gen.rmat(24, 16, seed=9, weighted=True)
    pool = SlabPool(16 << 30)
    g = DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                     in_degree_hints=gen.degree_hints(n, d), track_updates=True)
    g.InsertEdges(s, d, w)
    rng = np.random.default_rng(25)
    deg = np.bincount(s, minlength=n)
    source = int(rng.choice(np.nonzero(deg)[0]))
    sp = SSSP(g, source)
    sp.Static()
    # weight batches: distinct present edges, new weights in [1, 100]
    wb = []
    for i in range(3):
        idx = rng.choice(len(s), size=1_000_000, replace=False)
        wb.append((device_u32(s[idx]), device_u32(d[idx]),
                   device_u32(rng.integers(1, 101, len(idx)).astype(np.uint32))))
    keys = gen.keys_of(s, d)
    ib = []
    for i in range(3):
        bs, bd, bw = gen.insert_batch(keys, n, 1_000_000, seed=26 * 1000 + i, weighted=True,
                                      wseed=27 + i)
        ib.append((device_u32(bs), device_u32(bd), device_u32(bw)))
        keys = np.union1d(keys, gen.keys_of(bs, bd))
    g.UpdateSlabPointers()
    _sync()

    def wall(fn):
        _sync()
        t0 = time.perf_counter()
        r = fn()
        _sync()
        return r, (time.perf_counter() - t0) * 1e3

    def work():
        rec = {"update_weights_ms": [], "reweight_ms": [], "insert_ms": [],
               "updated_edges_ms": [], "updated_edges": [], "incremental_ms": [],
               "update_slab_pointers_ms": []}
        for bs, bd, bw in wb:
            rec["update_weights_ms"].append(wall(lambda: g.UpdateEdges(bs, bd, bw))[1])
            rec["reweight_ms"].append(sp.Reweight(bs, bd, bw))
        for bs, bd, bw in ib:
            rec["insert_ms"].append(wall(lambda: g.InsertEdges(bs, bd, bw))[1])
            (us, ud, uw), t = wall(g.UpdatedEdges)
            rec["updated_edges_ms"].append(t)
            rec["updated_edges"].append(int(us.numel()))
            rec["incremental_ms"].append(sp.Incremental(us, ud, uw))
            rec["update_slab_pointers_ms"].append(wall(g.UpdateSlabPointers)[1])
            del us, ud, uw
        return rec

    rec, el, kern = _region(work)
    sp.close()
    g.close()
    pool.close()
    return {"workload": "rmat24_weight_updates_and_updated_edges", "vertices": n,
            "edges": len(s), "source": source, **rec, "region_s": round(el, 4),
            "kernels": kern}


WORKLOADS = {"pagerank": pagerank, "tc_cc": tc_cc, "static": static, "multi": multi,
             "weights": weights}


def main(argv):
    knobs = [a.split("=", 1) for a in argv[1:]]
    if not argv or argv[0] not in WORKLOADS or any(len(k) != 2 for k in knobs):
        print(__doc__, file=sys.stderr)
        return 2
    import torch

    from . import autotune

    assert torch.cuda.is_available(), "the workloads need a GPU"
    for name, value in knobs:
        autotune.set_knob(name, int(value))
    out = WORKLOADS[argv[0]]()
    out["knobs"] = autotune.knobs()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
