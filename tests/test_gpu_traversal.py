"""GPU parity of BFS and SSSP (static / incremental / decremental) against the
oracle's from-scratch recompute on the updated graph: distances bit-exact,
parents valid (dist[p] + w(p, v) == dist[v] with (p, v) present).

Workload shapes follow the reference drivers (src/BFS.cu, src/SSSP.cu):
build, Static, then per batch mutate the graph and call Incremental or
Decremental with the same batch.
"""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import gen

pytestmark = pytest.mark.gpu


class EdgeSet:
    """Host copy of the current edge set (sorted keys + weights)."""

    def __init__(self, s, d, w=None):
        self.keys = gen.keys_of(s, d)
        order = np.argsort(self.keys)
        self.keys = self.keys[order]
        self.w = None if w is None else w[order]

    def insert(self, s, d, w=None):
        k = gen.keys_of(s, d)
        self.keys = np.concatenate([self.keys, k])
        if self.w is not None:
            self.w = np.concatenate([self.w, w])
        order = np.argsort(self.keys, kind="stable")
        self.keys = self.keys[order]
        if self.w is not None:
            self.w = self.w[order]

    def delete(self, s, d):
        keep = ~np.isin(self.keys, gen.keys_of(s, d))
        self.keys = self.keys[keep]
        if self.w is not None:
            self.w = self.w[keep]

    def csr(self, n):
        s, d = gen.split_keys(self.keys)
        return oracle.csr(n, s, d, self.w)


def check_bfs(algo, es, n, source):
    ref = es.csr(n)
    dist, par = algo.read()
    exp = oracle.bfs(ref, source)
    assert np.array_equal(dist, exp), int((dist != exp).sum())
    assert oracle.check_parents_fast(ref, dist, par, source) == 0


def check_sssp(algo, es, n, source):
    ref = es.csr(n)
    dist, par = algo.read()
    exp = oracle.dijkstra(ref, source)
    assert np.array_equal(dist, exp), int((dist != exp).sum())
    assert oracle.check_parents_fast(ref, dist, par, source, weights=True) == 0


@pytest.mark.parametrize("hints", [True, False])
def test_bfs_static_matches_oracle(gpu, hints):
    n, s, d, _ = gen.rmat(14, 16, seed=1)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s) if hints else None)
    g.InsertEdges(s, d)
    b = mk.BFS(g, 0)
    b.Static()
    check_bfs(b, EdgeSet(s, d), n, 0)
    # a source in the middle of the graph, and one with no out-edges
    isolated = int(np.nonzero(gen.degree_hints(n, s) == 0)[0][0])
    for src in (123, isolated):
        b2 = mk.BFS(g, src)
        b2.Static()
        check_bfs(b2, EdgeSet(s, d), n, src)


def test_bfs_incremental_config1_shape(gpu):
    # RMAT-16, edge factor 16, 10 batches of 10,000 new edges (config 1)
    n, s, d, _ = gen.rmat(16, 16, seed=1)
    pool = mk.SlabPool(1 << 29)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    es = EdgeSet(s, d)
    b = mk.BFS(g, 0)
    b.Static()
    check_bfs(b, es, n, 0)
    for i in range(10):
        bs, bd, _ = gen.insert_batch(es.keys, n, 10000, seed=100 + i)
        g.InsertEdges(bs, bd)
        es.insert(bs, bd)
        b.Incremental(bs, bd)
        check_bfs(b, es, n, 0)


@pytest.mark.parametrize("reverse", [False, True])
def test_bfs_decremental_and_mixed(gpu, reverse):
    n, s, d, _ = gen.rmat(14, 8, seed=2)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), reverse=reverse,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d)
    es = EdgeSet(s, d)
    b = mk.BFS(g, 0)
    b.Static()
    for i in range(6):
        ds, dd = gen.delete_batch(es.keys, 4000, seed=200 + i)
        g.DeleteEdges(ds, dd)
        es.delete(ds, dd)
        b.Decremental(ds, dd)
        check_bfs(b, es, n, 0)
        bs, bd, _ = gen.insert_batch(es.keys, n, 4000, seed=300 + i)
        g.InsertEdges(bs, bd)
        es.insert(bs, bd)
        b.Incremental(bs, bd)
        check_bfs(b, es, n, 0)


@pytest.mark.parametrize("reverse", [False, True])
def test_bfs_decremental_cuts_off_subtrees(gpu, reverse):
    # a path 0-1-2-...-99 plus a shortcut 0->50: deleting 0->50 pushes the
    # tail back to path levels; deleting 10->11 disconnects 11..99
    n = 100
    s = np.r_[np.arange(99), 0].astype(np.uint32)
    d = np.r_[np.arange(1, 100), 50].astype(np.uint32)
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(n, pool, 0.7, None, reverse=reverse)
    g.InsertEdges(s, d)
    es = EdgeSet(s, d)
    b = mk.BFS(g, 0)
    b.Static()
    check_bfs(b, es, n, 0)
    for cut in ([0], [10]):
        cs = np.array(cut, np.uint32)
        cd = np.array([50 if cut == [0] else 11], np.uint32)
        g.DeleteEdges(cs, cd)
        es.delete(cs, cd)
        b.Decremental(cs, cd)
        check_bfs(b, es, n, 0)
    dist, _ = b.read()
    assert (dist[11:] == mk.INF).all() and dist[10] == 10


def test_incremental_empty_and_selfloop_batches(gpu):
    n, s, d, _ = gen.rmat(10, 8, seed=4)
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    b = mk.BFS(g, 0)
    b.Static()
    e = np.zeros(0, np.uint32)
    b.Incremental(e, e)
    b.Decremental(e, e)
    loops = np.arange(5, dtype=np.uint32)
    g.InsertEdges(loops, loops)
    b.Incremental(loops, loops)
    check_bfs(b, EdgeSet(s, d), n, 0)


@pytest.mark.parametrize("reverse", [False, True])
def test_sssp_static_incremental_decremental_config2_shape(gpu, reverse):
    # config 2 shape at scale 15: mixed batches applied as delete ->
    # Decremental, then insert -> Incremental
    n, s, d, w = gen.rmat(15, 16, seed=3, weighted=True)
    pool = mk.SlabPool(1 << 29)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=reverse,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    es = EdgeSet(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    check_sssp(sp, es, n, 0)
    for i in range(5):
        ds, dd = gen.delete_batch(es.keys, 5000, seed=400 + i)
        g.DeleteEdges(ds, dd)
        es.delete(ds, dd)
        sp.Decremental(ds, dd)
        check_sssp(sp, es, n, 0)
        bs, bd, bw = gen.insert_batch(es.keys, n, 5000, seed=500 + i, weighted=True, wseed=9)
        g.InsertEdges(bs, bd, bw)
        es.insert(bs, bd, bw)
        sp.Incremental(bs, bd, bw)
        check_sssp(sp, es, n, 0)


def test_bfs_on_weighted_graph_ignores_weights(gpu):
    n, s, d, w = gen.rmat(12, 16, seed=5, weighted=True)
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    b = mk.BFS(g, 0)
    b.Static()
    check_bfs(b, EdgeSet(s, d), n, 0)


def test_sssp_rejects_unweighted_graph(gpu):
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(10, pool)
    with pytest.raises(mk.MeerkatError):
        mk.SSSP(g, 0)


def test_sssp_incremental_uses_graph_weight_for_existing_edges(gpu):
    # the batch repeats existing edges with smaller weights: the graph keeps
    # the original weights, so the relaxation must use those (status path),
    # and a copy of the batch (different pointers) must give the same result
    # through the search path
    n, s, d, w = gen.rmat(12, 16, seed=8, weighted=True)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    es = EdgeSet(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    bs, bd, bw = gen.insert_batch(es.keys, n, 3000, seed=77, weighted=True)
    rs, rd = s[:2000], d[:2000]
    rw = np.ones(2000, np.uint32)  # smaller than the stored weights
    S = np.r_[bs, rs]
    D = np.r_[bd, rd]
    Wt = np.r_[bw, rw]
    s_dev, d_dev, w_dev = g.InsertEdges(S, D, Wt)
    es.insert(bs, bd, bw)
    sp.Incremental(s_dev, d_dev, w_dev)  # same device batch: status path
    check_sssp(sp, es, n, 0)
    sp2 = mk.SSSP(g, 0)
    sp2.Static()
    bs2, bd2, bw2 = gen.insert_batch(es.keys, n, 2000, seed=78, weighted=True)
    g.InsertEdges(bs2, bd2, bw2)
    es.insert(bs2, bd2, bw2)
    sp.Incremental(bs2, bd2, bw2)  # fresh device copies: search path
    sp2.Static()
    check_sssp(sp, es, n, 0)
    assert np.array_equal(sp.Distances(), sp2.Distances())
