"""Update-aware iteration: UpdateSlabPointers (include/Graph/SlabGraphImpl.cuh:
64-82 -> device/UpdateSlablistPointers.cuh:4-104) and the iteration of the
edges inserted since (the markers of include/Graph/SlabGraph.cuh:377-379),
mk_graph_updated_edges.

The expected set is computed on the host from the batches: edges applied
by inserts since the last UpdateSlabPointers (since creation before the
first), minus those deleted since, with their weights in the graph.  Cases:
duplicates and self-loops in a batch, edges already present, deletion of
new and of old edges, re-insertion of a deleted edge into a reused
(tombstoned) slot, hub chains growing overflow slabs after the marks were
cleared, weight updates (not insertions), and a graph of over a million
buckets (the iterator's waves then scan many 64-bucket chunks each).  Then
the consumer the iteration is for: incremental BFS / SSSP driven by the
graph's own updated edges after several insert batches, bit-exact against
the oracle."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from hostedges import HostEdges
from meerkat_amd import _lib, gen

pytestmark = pytest.mark.gpu


def updated_set(g):
    s, d, w = g.UpdatedEdges()
    s = s.cpu().numpy().view(np.uint32)
    d = d.cpu().numpy().view(np.uint32)
    keys = gen.keys_of(s, d)
    assert len(np.unique(keys)) == len(keys), "an edge reported twice"
    if w is None:
        return set(keys.tolist()), None
    w = w.cpu().numpy().view(np.uint32)
    return set(keys.tolist()), dict(zip(keys.tolist(), w.tolist()))


@pytest.mark.parametrize("weighted", [False, True])
def test_updated_edges_follow_inserts_and_deletes(gpu, weighted):
    n, s, d, w = gen.rmat(12, 16, seed=51, weighted=weighted)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted,
                        track_updates=True)
    g.InsertEdges(s, d, w)
    host = HostEdges(n, s, d, w)
    # before the first UpdateSlabPointers every edge is new
    got, gw = updated_set(g)
    assert got == set(host.keys.tolist())
    g.UpdateSlabPointers()
    assert updated_set(g)[0] == set()

    rng = np.random.default_rng(3)
    a_s, a_d, a_w = gen.insert_batch(host.keys, n, 3000, seed=4, weighted=True)
    # ride-alongs: present edges, duplicates inside the batch, self-loops
    old = rng.choice(len(s), 200, replace=False)
    bs = np.concatenate([a_s, s[old], a_s[:100], np.arange(3, dtype=np.uint32)])
    bd = np.concatenate([a_d, d[old], a_d[:100], np.arange(3, dtype=np.uint32)])
    # (a duplicate carries its first occurrence's weight: which copy of a
    # repeated edge lands first is not fixed)
    bw = (np.concatenate([a_w, np.full(200, 9, np.uint32), a_w[:100], np.full(3, 9, np.uint32)])
          if weighted else None)
    g.InsertEdges(bs, bd, bw)
    host.insert(bs, bd, bw)
    expect = set(gen.keys_of(a_s, a_d).tolist())
    assert updated_set(g)[0] == expect
    # delete some new and some old edges; re-insert one old edge (it lands in
    # a tombstoned slot or elsewhere: new either way)
    kill_new = rng.choice(len(a_s), 500, replace=False)
    kill_old = rng.choice(len(s), 300, replace=False)
    ds = np.concatenate([a_s[kill_new], s[kill_old]])
    dd = np.concatenate([a_d[kill_new], d[kill_old]])
    g.DeleteEdges(ds, dd)
    host.delete(ds, dd)
    expect -= set(gen.keys_of(a_s[kill_new], a_d[kill_new]).tolist())
    back = kill_old[:50]
    rw = (np.arange(50, dtype=np.uint32) % 100 + 1) if weighted else None
    g.InsertEdges(s[back], d[back], rw)
    host.insert(s[back], d[back], rw)
    expect |= set(gen.keys_of(s[back], d[back]).tolist())
    if weighted:  # a weight update is not an insertion
        g.UpdateEdges(s[:20], d[:20], np.full(20, 55, np.uint32))
        host.update_weights(s[:20], d[:20], np.full(20, 55, np.uint32))
    got, gw = updated_set(g)
    assert got == expect
    if weighted:
        ref = host.csr()
        src_e = np.repeat(np.arange(n, dtype=np.uint32), np.diff(ref.row_off).astype(np.int64))
        wmap = dict(zip(gen.keys_of(src_e, ref.col).tolist(), ref.w.tolist()))
        assert all(gw[k] == wmap[k] for k in got)
    g.UpdateSlabPointers()
    assert updated_set(g)[0] == set()
    g.close()
    pool.close()


def test_hub_chain_overflow_after_marks_cleared(gpu):
    # one vertex with thousands of out-edges and a single bucket: its chain
    # grows overflow slabs across mark generations
    n = 1 << 14
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(n, pool, 0.7, None, track_updates=True)
    first = np.arange(1, 2001, dtype=np.uint32)
    g.InsertEdges(np.zeros(2000, np.uint32), first)
    g.UpdateSlabPointers()
    second = np.arange(2001, 6001, dtype=np.uint32)
    g.InsertEdges(np.zeros(4000, np.uint32), second)
    g.DeleteEdges(np.zeros(100, np.uint32), first[:100])
    g.InsertEdges(np.zeros(100, np.uint32), first[:100])  # reuse tombstones: new again
    got, _ = updated_set(g)
    assert got == set(gen.keys_of(np.zeros(4100, np.uint32),
                                  np.concatenate([second, first[:100]])).tolist())
    assert g.slab_stats()["max_chain"] > 60
    g.close()
    pool.close()


def test_untracked_graph_refuses_iteration(gpu):
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(1024, pool, 0.7, None)
    g.UpdateSlabPointers()  # a no-op, as the reference's with nothing to move
    with pytest.raises(mk.MeerkatError) as e:
        g.UpdatedEdges()
    assert e.value.code == _lib.MK_ERR_STATE
    g.close()
    pool.close()


@pytest.mark.parametrize("weighted", [False, True])
def test_incremental_from_the_graphs_updated_edges(gpu, weighted):
    # three insert batches, then one Incremental over the graph's own
    # updated edges (no batch kept by the caller)
    n, s, d, w = gen.rmat(14, 8, seed=52, weighted=weighted)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted,
                        track_updates=True)
    g.InsertEdges(s, d, w)
    host = HostEdges(n, s, d, w)
    algo = mk.SSSP(g, 0) if weighted else mk.BFS(g, 0)
    algo.Static()
    prev_keys = set()
    for rnd in range(2):
        g.UpdateSlabPointers()
        new_keys = set()
        for i in range(3):
            bs, bd, bw = gen.insert_batch(host.keys, n, 4000, seed=60 + 10 * rnd + i,
                                          weighted=weighted)
            g.InsertEdges(bs, bd, bw)
            before = len(host.added)
            host.insert(bs, bd, bw)
            new_keys |= set(gen.keys_of(bs, bd).tolist()) - set(host.keys.tolist())
            assert len(host.added) > before
        us, ud, uw = g.UpdatedEdges()
        got = gen.keys_of(us.cpu().numpy().view(np.uint32), ud.cpu().numpy().view(np.uint32))
        # the first round's edges are old now: only this round's are reported
        assert set(got.tolist()) == new_keys - prev_keys
        prev_keys = prev_keys | new_keys
        if weighted:
            algo.Incremental(us, ud, uw)
        else:
            algo.Incremental(us, ud)
        ref = host.csr()
        dist, par = algo.read()
        exp = oracle.dijkstra(ref, 0) if weighted else oracle.bfs(ref, 0)
        assert np.array_equal(dist, exp), int((dist != exp).sum())
        assert oracle.check_parents_fast(ref, dist, par, 0, weights=weighted) == 0
    algo.close()
    g.close()
    pool.close()


def test_updated_edges_over_many_bucket_chunks(gpu):
    # RMAT-20: over a million buckets, so each wave of the iterator scans
    # several 64-bucket chunks; the iterator's count pass (no output) and its
    # emit pass must agree, and report exactly the applied inserts with
    # their weights, once each
    n, s, d, w = gen.rmat(20, 8, seed=53, weighted=True)
    pool = mk.SlabPool(1 << 30)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True,
                        track_updates=True)
    g.InsertEdges(s, d, w)
    assert g.slab_stats()["head_slabs"] > 64 * 8192
    g.UpdateSlabPointers()
    keys = gen.keys_of(s, d)
    bs, bd, bw = gen.insert_batch(keys, n, 200_000, seed=54, weighted=True)
    g.InsertEdges(bs, bd, bw)
    cnt = _lib.ctypes.c_uint64()
    _lib.call("mk_graph_updated_edges", g.handle, None, None, None, 0, _lib.ctypes.byref(cnt))
    assert cnt.value == len(bs)
    got, wts = updated_set(g)
    exp = gen.keys_of(bs, bd).tolist()
    assert got == set(exp)
    assert all(wts[k] == int(x) for k, x in zip(exp, bw))
    g.close()
    pool.close()
