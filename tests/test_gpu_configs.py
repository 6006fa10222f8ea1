"""The four parity workloads at their full sizes, on the GPU, checked against
the oracle's from-scratch recompute on the updated graph:

- C1: RMAT scale 16 (seed 1), static BFS from 0, then 10 batches of 10,000
  new edges (seed 2), incremental BFS after each: levels bit-exact, parents
  valid, after every batch.
- C2: RMAT scale 20, weights uniform [1, 100] (seed 3), 20 batches of
  50,000 deletions of existing edges then 50,000 new-edge insertions
  (seed 4); decremental then incremental SSSP from 0: distances bit-exact
  and parents valid after every batch.
- C3: directed Erdos-Renyi, 4,194,304 vertices, 33,554,432 edges (seed 5),
  10 batches of 1,000,000 new edges (seed 6); incremental PageRank (d 0.85,
  stop at L1 delta 1e-7): L1 distance to the FP64 PageRank <= 1e-6 after
  every batch.
- C4: undirected RMAT scale 22 (seed 7), 10 batches of 500,000 new
  undirected edges (both directions, seed 8); incremental triangle count and
  connected components: component labels and the triangle count bit-exact
  at the static count and after every batch.

Per-batch GPU times are written as one JSON line per workload to
$MK_RESULTS_DIR/configs.jsonl when that variable names a directory."""
import json
import os
import time

import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import gen

pytestmark = pytest.mark.gpu


def record(name, data):
    out = os.environ.get("MK_RESULTS_DIR")
    print(name, json.dumps(data), flush=True)
    if out and os.path.isdir(out):
        with open(os.path.join(out, "configs.jsonl"), "a") as f:
            f.write(json.dumps({"workload": name, **data}) + "\n")


def sync():
    import torch

    torch.cuda.synchronize()


def test_c1_rmat16_bfs_inserts(gpu):
    n, s, d, _ = gen.rmat(16, 16, seed=1)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    b = mk.BFS(g, 0)
    b.Static()  # the first call also pays one-time kernel loading
    static_ms = b.Static()
    keys = gen.keys_of(s, d)
    assert np.array_equal(b.Distances(), oracle.bfs(oracle.csr(n, s, d), 0))
    upd, incr = [], []
    for i in range(10):
        bs, bd, _ = gen.insert_batch(keys, n, 10_000, seed=2 * 1000 + i)
        sd, dd = mk.device_u32(bs), mk.device_u32(bd)
        sync()
        t = time.perf_counter()
        g.InsertEdges(sd, dd)
        upd.append(time.perf_counter() - t)
        incr.append(b.Incremental(sd, dd))
        keys = np.union1d(keys, gen.keys_of(bs, bd))
        s2, d2 = gen.split_keys(keys)
        ref = oracle.csr(n, s2, d2)
        dist, par = b.read()
        assert np.array_equal(dist, oracle.bfs(ref, 0)), i
        assert oracle.check_parents_fast(ref, dist, par, 0) == 0, i
    record("C1", {"vertices": n, "edges": int(len(s)), "static_bfs_ms": static_ms,
                  "incremental_bfs_ms": incr, "insert_edges_per_s": 10_000 / np.mean(upd)})


def test_c2_rmat20_sssp_mixed(gpu):
    n, s, d, w = gen.rmat(20, 16, seed=3, weighted=True)
    pool = mk.SlabPool(1 << 31)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()  # the first call also pays one-time kernel loading
    static_ms = sp.Static()
    keys, wt = gen.keys_of(s, d), w.copy()
    upd, decr, incr = [], [], []
    for i in range(20):
        ds, dd = gen.delete_batch(keys, 50_000, seed=4 * 1000 + i)
        bs, bd, bw = gen.insert_batch(keys, n, 50_000, seed=4 * 1000 + 500 + i, weighted=True,
                                      wseed=4 * 1000 + 700 + i)
        dsd, ddd = mk.device_u32(ds), mk.device_u32(dd)
        bsd, bdd, bwd = mk.device_u32(bs), mk.device_u32(bd), mk.device_u32(bw)
        sync()
        t = time.perf_counter()
        g.DeleteEdges(dsd, ddd)
        t1 = time.perf_counter()
        decr.append(sp.Decremental(dsd, ddd))
        t2 = time.perf_counter()
        g.InsertEdges(bsd, bdd, bwd)
        t3 = time.perf_counter()
        incr.append(sp.Incremental(bsd, bdd, bwd))
        upd.append((t1 - t) + (t3 - t2))
        keep = ~np.isin(keys, gen.keys_of(ds, dd))
        keys, wt = keys[keep], wt[keep]
        nk = gen.keys_of(bs, bd)
        keys, wt = np.r_[keys, nk], np.r_[wt, bw]
        order = np.argsort(keys)
        keys, wt = keys[order], wt[order]
        s2, d2 = gen.split_keys(keys)
        ref = oracle.csr(n, s2, d2, wt)
        dist, par = sp.read()
        assert np.array_equal(dist, oracle.dijkstra(ref, 0)), i
        assert oracle.check_parents_fast(ref, dist, par, 0, weights=True) == 0, i
    record("C2", {"vertices": n, "edges": int(len(s)), "static_sssp_ms": static_ms,
                  "decremental_sssp_ms": decr, "incremental_sssp_ms": incr,
                  "update_edges_per_s": 100_000 / np.mean(upd)})


def test_c3_er4m_pagerank_inserts(gpu):
    n, s, d, _ = gen.erdos_renyi(4_194_304, 33_554_432, seed=5)
    pool = mk.SlabPool(1 << 32)
    # reference style: a graph of reversed edges with host out-degrees
    rev = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, d))
    rev.InsertEdges(d, s)
    pr = mk.PageRank(rev, 0.85, 1e-7, out_degree=gen.degree_hints(n, s))
    static_ms = pr.Static()
    keys = gen.keys_of(s, d)
    upd, incr, iters, errs = [], [], [], []
    for i in range(10):
        bs, bd, _ = gen.insert_batch(keys, n, 1_000_000, seed=6 * 1000 + i)
        bsd, bdd = mk.device_u32(bs), mk.device_u32(bd)
        sync()
        t = time.perf_counter()
        rev.InsertEdges(bdd, bsd)
        upd.append(time.perf_counter() - t)
        incr.append(pr.Incremental(bsd, bdd))
        iters.append(pr.last_stats()["iterations"])
        keys = np.union1d(keys, gen.keys_of(bs, bd))
        s2, d2 = gen.split_keys(keys)
        g = oracle.csr(n, s2, d2)
        ref, _, delta = oracle.pagerank(oracle.transpose(g), g.degrees(), 0.85, 1e-14, 5000)
        assert delta <= 1e-14
        err = float(np.abs(pr.Values().astype(np.float64) - ref).sum())
        errs.append(err)
        assert err <= 1e-6, (i, err)
    record("C3", {"vertices": n, "edges": int(len(s)), "static_pagerank_ms": static_ms,
                  "incremental_pagerank_ms": incr, "iterations": iters, "l1_error": errs,
                  "insert_edges_per_s": 1_000_000 / np.mean(upd)})


def test_c4_rmat22_tc_cc_inserts(gpu):
    n, s, d, _ = gen.rmat(22, 16, seed=7, undirected=True)
    pool = mk.SlabPool(1 << 33)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    tc, cc = mk.TC(g), mk.CC(g)
    tc_static_ms = tc.Static()
    cc_static_ms = cc.Static()
    keys = gen.keys_of(s, d)
    ref = oracle.csr(n, s, d)
    assert tc.GetCount() == oracle.tc(ref)
    assert np.array_equal(cc.Labels(), oracle.cc(ref)[0])
    upd, sizes, tci, cci = [], [], [], []
    for i in range(10):
        bs, bd, _ = gen.insert_batch(keys, n, 500_000, seed=8 * 1000 + i)
        # both directions of each new undirected edge, once
        k = np.unique(np.r_[gen.keys_of(bs, bd), gen.keys_of(bd, bs)])
        S, D = gen.split_keys(k)
        Sd, Dd = mk.device_u32(S), mk.device_u32(D)
        sync()
        t = time.perf_counter()
        g.InsertEdges(Sd, Dd)
        upd.append(time.perf_counter() - t)
        sizes.append(len(k))
        tci.append(tc.Incremental(Sd, Dd))
        cci.append(cc.Incremental(Sd, Dd))
        keys = np.union1d(keys, k)
        s2, d2 = gen.split_keys(keys)
        ref = oracle.csr(n, s2, d2)
        assert np.array_equal(cc.Labels(), oracle.cc(ref)[0]), i
        assert tc.GetCount() == oracle.tc(ref), i
        print(f"C4 batch {i}: {tc.GetCount()} triangles, labels exact", flush=True)
    record("C4", {"vertices": n, "directed_entries": int(len(s)), "static_tc_ms": tc_static_ms,
                  "static_cc_ms": cc_static_ms, "incremental_tc_ms": tci,
                  "incremental_cc_ms": cci, "triangles": int(tc.GetCount()),
                  "insert_entries_per_s": float(np.sum(sizes) / np.sum(upd))})
