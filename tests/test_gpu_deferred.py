"""Deferred mutation checks (mk_graph_set_deferred): a mutation call returns
once queued; the next call on the graph, or the BFS/SSSP call queued after
it, settles its outcome.  Results bit-exact against the oracle; failures
surface once, with the graph restored as the synchronous call leaves it."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import _lib, gen

pytestmark = pytest.mark.gpu


def check(algo, n, s, d, w=None, source=0):
    ref = oracle.csr(n, s, d, w)
    dist, par = algo.read()
    exp = oracle.dijkstra(ref, source) if w is not None else oracle.bfs(ref, source)
    assert np.array_equal(dist, exp), int((dist != exp).sum())
    assert oracle.check_parents_fast(ref, dist, par, source, weights=w is not None) == 0


@pytest.mark.parametrize("mirror", [False, True])
def test_deferred_batches_keep_sssp_exact(gpu, mirror):
    n, s, d, w = gen.rmat(14, 8, seed=61, weighted=True)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=mirror,
                        in_degree_hints=gen.degree_hints(n, d) if mirror else None)
    g.InsertEdges(s, d, w)
    g.set_deferred(True)
    sp = mk.SSSP(g, 0)
    sp.Static()
    keys, wt = gen.keys_of(s, d), w.copy()
    for i in range(5):
        ds, dd = gen.delete_batch(keys, 2000, seed=700 + i)
        g.DeleteEdges(ds, dd)
        sp.Decremental(ds, dd, None)
        keep = ~np.isin(keys, gen.keys_of(ds, dd))
        keys, wt = keys[keep], wt[keep]
        bs, bd, bw = gen.insert_batch(keys, n, 2000, seed=800 + i, weighted=True, wseed=900 + i)
        g.InsertEdges(bs, bd, bw)
        sp.Incremental(bs, bd, bw)
        keys, wt = np.r_[keys, gen.keys_of(bs, bd)], np.r_[wt, bw]
        order = np.argsort(keys)
        keys, wt = keys[order], wt[order]
        s2, d2 = gen.split_keys(keys)
        check(sp, n, s2, d2, wt)
    assert g.num_edges() == len(keys)


def test_deferred_range_error_surfaces_in_the_traversal(gpu):
    n, s, d, _ = gen.rmat(12, 8, seed=62)
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    g.set_deferred(True)
    b = mk.BFS(g, 0)
    b.Static()
    before = g.export_csr()
    bad_s, bad_d = np.array([0, 1, n + 3], np.uint32), np.array([5, 6, 7], np.uint32)
    g.InsertEdges(bad_s, bad_d)  # queued: returns at once
    with pytest.raises(mk.MeerkatError) as ei:
        b.Incremental(bad_s, bad_d)
    assert ei.value.code == _lib.MK_ERR_RANGE
    with pytest.raises(mk.MeerkatError) as ei:  # results invalid until Static
        b.Incremental(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert ei.value.code == _lib.MK_ERR_STATE
    after = g.export_csr()  # nothing was applied
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    b.Static()
    check(b, n, s, d)


def test_deferred_pool_failure_surfaces_once_and_rolls_back(gpu):
    n = 200000
    pool = mk.SlabPool(2048 * 256)
    g = mk.DynamicGraph(n, pool, 0.7, None, reverse=True)
    g.InsertEdges(np.zeros(500, np.uint32), np.arange(1, 501, dtype=np.uint32))
    g.set_deferred(True)
    before_f, before_r = g.export_csr(), g.mirror().export_csr()
    g.InsertEdges(np.zeros(150000, np.uint32), np.arange(1, 150001, dtype=np.uint32))
    # the next mutation reports the queued batch's failure and does not run
    with pytest.raises(mk.MeerkatError) as ei:
        g.InsertEdges(np.array([9], np.uint32), np.array([10], np.uint32))
    assert ei.value.code == _lib.MK_ERR_POOL_FULL
    g.sync()  # settled: nothing pending any more
    after_f, after_r = g.export_csr(), g.mirror().export_csr()
    assert np.array_equal(before_f[1], after_f[1]) and np.array_equal(before_r[1], after_r[1])
    assert g.num_edges() == 500
    g.InsertEdges(np.array([9], np.uint32), np.array([10], np.uint32))
    g.sync()
    assert g.num_edges() == 501


# ---------------------------------------------------------------- deferred
# BFS/SSSP calls (set_deferred on the handle): Incremental / Decremental
# return once queued; the handle's next call, read or sync settles them.

@pytest.mark.parametrize("mirror", [False, True])
def test_deferred_calls_run_ahead_and_stay_exact(gpu, mirror):
    n, s, d, w = gen.rmat(14, 8, seed=63, weighted=True)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=mirror,
                        in_degree_hints=gen.degree_hints(n, d) if mirror else None)
    g.InsertEdges(s, d, w)
    g.set_deferred(True)
    sp, bf = mk.SSSP(g, 0), mk.BFS(g, 0)
    sp.Static()
    bf.Static()
    sp.set_deferred(True)
    bf.set_deferred(True)
    keys, wt = gen.keys_of(s, d), w.copy()
    for i in range(6):
        ds, dd = gen.delete_batch(keys, 1500, seed=1700 + i)
        g.DeleteEdges(mk.device_u32(ds), mk.device_u32(dd))
        assert sp.Decremental(ds, dd, None) == 0.0  # queued, not measured
        assert bf.Decremental(ds, dd) == 0.0
        keep = ~np.isin(keys, gen.keys_of(ds, dd))
        keys, wt = keys[keep], wt[keep]
        bs, bd, bw = gen.insert_batch(keys, n, 1500, seed=1800 + i, weighted=True,
                                      wseed=1900 + i)
        g.InsertEdges(bs, bd, bw)
        sp.Incremental(bs, bd, bw)
        bf.Incremental(bs, bd)
        keys, wt = np.r_[keys, gen.keys_of(bs, bd)], np.r_[wt, bw]
        order = np.argsort(keys)
        keys, wt = keys[order], wt[order]
        if i % 3 == 2:  # a few rounds queued back to back, then a read settles
            s2, d2 = gen.split_keys(keys)
            check(sp, n, s2, d2, wt)
            check(bf, n, s2, d2)
            assert sp.last_stats()["edges_read"] > 0
    sp.sync()
    sp.set_deferred(False)
    assert sp.Incremental(np.zeros(0, np.uint32), np.zeros(0, np.uint32)) >= 0.0


@pytest.mark.parametrize("weighted", [False, True])
def test_deferred_call_past_one_launch_is_recomputed(gpu, weighted):
    # a 1000-vertex path: after the reconnecting insert the relax runs ~480
    # steps, more than one loop launch reserves; a deferred call cannot be
    # continued in place, so settling it recomputes from scratch
    n = 1000
    s = np.arange(n - 1, dtype=np.uint32)
    d = s + 1
    w = (np.arange(n - 1, dtype=np.uint32) % 7 + 1) if weighted else None
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(n, pool, 0.7, None, weighted=weighted)
    g.InsertEdges(s, d, w)
    algo = mk.SSSP(g, 0) if weighted else mk.BFS(g, 0)
    algo.Static()
    algo.set_deferred(True)
    cs, cd = np.array([500], np.uint32), np.array([501], np.uint32)
    g.DeleteEdges(cs, cd)
    algo.Decremental(cs, cd) if not weighted else algo.Decremental(cs, cd, None)
    keep = ~((s == 500) & (d == 501))
    s1, d1 = s[keep], d[keep]
    w1 = w[keep] if weighted else None
    bs, bd = np.array([499, 10], np.uint32), np.array([700, 520], np.uint32)
    bw = np.array([3, 5], np.uint32) if weighted else None
    g.InsertEdges(bs, bd, bw)
    if weighted:
        algo.Incremental(bs, bd, bw)
    else:
        algo.Incremental(bs, bd)
    s2, d2 = np.r_[s1, bs], np.r_[d1, bd]
    w2 = np.r_[w1, bw] if weighted else None
    check(algo, n, s2, d2, w2)  # the read settles (and recomputes)


def test_deferred_call_on_a_failed_batch_reports_it(gpu):
    # the mutation a deferred call runs on is settled when the call is
    # queued: its failure is the call's error, and the handle stays invalid
    # until Static
    n, s, d, _ = gen.rmat(12, 8, seed=64)
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    g.set_deferred(True)
    b = mk.BFS(g, 0)
    b.Static()
    b.set_deferred(True)
    bad_s, bad_d = np.array([0, 1, n + 3], np.uint32), np.array([5, 6, 7], np.uint32)
    g.InsertEdges(bad_s, bad_d)
    with pytest.raises(mk.MeerkatError) as ei:
        b.Incremental(bad_s, bad_d)
    assert ei.value.code == _lib.MK_ERR_RANGE
    with pytest.raises(mk.MeerkatError) as ei:
        b.Incremental(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert ei.value.code == _lib.MK_ERR_STATE
    b.Static()
    check(b, n, s, d)


@pytest.mark.parametrize("failure", ["range", "pool"])
def test_failed_batch_is_reported_to_every_handle(gpu, failure):
    # the usual pattern of one graph with an SSSP and a BFS handle, deferred
    # mutations and deferred calls: the first call settles the failed batch
    # and reports it; the second handle, given the same batch, must not run
    # on edges the graph never applied (its seed would relax them unchecked)
    if failure == "range":
        n, s, d, w = gen.rmat(12, 8, seed=65, weighted=True)
        pool = mk.SlabPool(1 << 26)
        bad_s = np.array([0, 1, n + 3], np.uint32)
        bad_d = np.array([5, 6, 7], np.uint32)
        code = _lib.MK_ERR_RANGE
    else:
        n = 200000
        s = np.zeros(500, np.uint32)
        d = np.arange(1, 501, dtype=np.uint32)
        w = np.full(500, 3, np.uint32)
        pool = mk.SlabPool(2048 * 256)
        bad_s = np.r_[np.zeros(100000, np.uint32), np.array([1], np.uint32)]
        bad_d = np.r_[np.arange(1000, 101000, dtype=np.uint32), np.array([0], np.uint32)]
        code = _lib.MK_ERR_POOL_FULL
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    g.set_deferred(True)
    sp, bf = mk.SSSP(g, 0), mk.BFS(g, 0)
    sp.Static()
    bf.Static()
    sp.set_deferred(True)
    bf.set_deferred(True)
    bs, bd = mk.device_u32(bad_s), mk.device_u32(bad_d)
    bw = mk.device_u32(np.ones(len(bad_s), np.uint32))
    g.InsertEdges(bs, bd, bw)
    with pytest.raises(mk.MeerkatError) as ei:
        sp.Incremental(bs, bd, bw)
    assert ei.value.code == code
    with pytest.raises(mk.MeerkatError) as ei:
        bf.Incremental(bs, bd)
    assert ei.value.code == code
    for h in (sp, bf):  # both invalid until Static
        with pytest.raises(mk.MeerkatError) as ei:
            h.Incremental(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
        assert ei.value.code == _lib.MK_ERR_STATE
    sp.Static()
    bf.Static()
    check(sp, n, s, d, w)
    check(bf, n, s, d)
    # the next mutation clears the record: a good batch runs everywhere
    gs, gd = mk.device_u32(np.array([2], np.uint32)), mk.device_u32(np.array([3], np.uint32))
    gw = mk.device_u32(np.array([4], np.uint32))
    g.InsertEdges(gs, gd, gw)
    sp.Incremental(gs, gd, gw)
    bf.Incremental(gs, gd)
    s2, d2, w2 = np.r_[s, [2]].astype(np.uint32), np.r_[d, [3]].astype(np.uint32), \
        np.r_[w, [4]].astype(np.uint32)
    check(sp, n, s2, d2, w2)
    check(bf, n, s2, d2)
