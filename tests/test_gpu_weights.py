"""Weight updates: DynamicGraph::UpdateEdges (include/Graph/SlabGraphImpl.cuh:
84-103, compiled out in the reference; device/UpdateEdgeKernel.cuh:7-37)
and SSSP kept exact across them (mk_sssp_reweight).

Graph level: present edges take the new weight in the graph and in its
in-edge mirror, absent edges are not inserted, self-loops are skipped, a
batch with an id >= n leaves the graph unchanged, a set graph refuses.
SSSP level: after each batch of mixed weight increases and decreases
(including tree edges, so subtrees are invalidated) the distances equal the
oracle's Dijkstra on the reweighted edge set, parents valid; with and
without the mirror, with one launch per step, deferred, and at the RMAT-20
size of the weighted SSSP workload."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from hostedges import HostEdges
from meerkat_amd import _lib, autotune, gen

pytestmark = pytest.mark.gpu


def build(scale, seed, reverse, weighted=True, ef=16):
    n, s, d, w = gen.rmat(scale, ef, seed=seed, weighted=weighted)
    pool = mk.SlabPool(1 << 30)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted, reverse=reverse,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    return n, s, d, w, pool, g


def check(sp, host, source=0):
    ref = host.csr()
    dist, par = sp.read()
    exp = oracle.dijkstra(ref, source)
    assert np.array_equal(dist, exp), int((dist != exp).sum())
    assert oracle.check_parents_fast(ref, dist, par, source, weights=True) == 0


def weight_batch(rng, n, s, d, sp, count, tree_frac=0.5):
    """`count` distinct present edges, half of them tree edges of the current
    SSSP tree, each with a new weight in [1, 100] (higher or lower)."""
    dist, par = sp.read()
    reached = np.nonzero((dist != 0xFFFFFFFF) & (par != 0xFFFFFFFF) &
                         (np.arange(n) != sp.source))[0]
    nt = min(int(count * tree_frac), len(reached))
    tv = rng.choice(reached, size=nt, replace=False).astype(np.uint32)
    tu = par[tv]
    keys = set(gen.keys_of(tu, tv).tolist())
    idx = rng.choice(len(s), size=4 * count, replace=False)
    rs, rd = [], []
    for a, b in zip(s[idx].tolist(), d[idx].tolist()):
        k = (a << 32) | b
        if k not in keys:
            keys.add(k)
            rs.append(a)
            rd.append(b)
        if len(rs) >= count - nt:
            break
    bs = np.concatenate([tu, np.array(rs, np.uint32)])
    bd = np.concatenate([tv, np.array(rd, np.uint32)])
    bw = rng.integers(1, 101, size=len(bs)).astype(np.uint32)
    return bs, bd, bw


@pytest.mark.parametrize("reverse", [False, True])
def test_update_edges_sets_present_weights_only(gpu, reverse):
    n, s, d, w, pool, g = build(12, 41, reverse)
    host = HostEdges(n, s, d, w)
    rng = np.random.default_rng(1)
    idx = rng.choice(len(s), size=3000, replace=False)
    us, ud = s[idx], d[idx]
    uw = rng.integers(1, 101, size=len(us)).astype(np.uint32)
    # absent edges and self-loops ride along
    absent_s, absent_d, _ = gen.insert_batch(host.keys, n, 500, seed=2)
    loops = np.arange(5, dtype=np.uint32)
    bs = np.concatenate([us, absent_s, loops])
    bd = np.concatenate([ud, absent_d, loops])
    bw = np.concatenate([uw, np.full(len(absent_s) + 5, 77, np.uint32)])
    m0 = g.num_edges()
    g.UpdateEdges(bs, bd, bw)
    st = g.last_status()
    assert (st[:len(us)] == _lib.EDGE_APPLIED).all()
    assert (st[len(us):len(us) + len(absent_s)] == _lib.EDGE_UNCHANGED).all()
    assert (st[len(us) + len(absent_s):] == _lib.EDGE_SELF_LOOP).all()
    assert g.num_edges() == m0
    host.update_weights(bs, bd, bw)
    ref = host.csr()
    row, col, ww = g.export_csr()
    assert np.array_equal(row, ref.row_off) and np.array_equal(col, ref.col)
    assert np.array_equal(ww, ref.w)
    found, fw = g.SearchEdges(us, ud)
    assert found.all() and np.array_equal(fw, uw)
    assert not g.SearchEdges(absent_s, absent_d)[0].any()
    if reverse:
        # the mirror holds (dst -> src, w) with the new weights
        src_e, dst_e = _pairs(ref)
        rt = oracle.csr(n, dst_e, src_e, ref.w)
        mrow, mcol, mw = g.mirror().export_csr()
        assert np.array_equal(mrow, rt.row_off) and np.array_equal(mcol, rt.col)
        assert np.array_equal(mw, rt.w)
    g.close()
    pool.close()


def _pairs(ref):
    """(src, dst) of every edge of an oracle CSR, row by row."""
    src = np.repeat(np.arange(ref.n, dtype=np.uint32),
                    np.diff(ref.row_off).astype(np.int64))
    return src, ref.col


def test_update_edges_errors_leave_graph_unchanged(gpu):
    n, s, d, w, pool, g = build(10, 42, True)
    before = g.export_csr()
    bs = np.array([s[0], n + 3], np.uint32)
    bd = np.array([d[0], 1], np.uint32)
    with pytest.raises(mk.MeerkatError) as e:
        g.UpdateEdges(bs, bd, np.array([55, 66], np.uint32))
    assert e.value.code == _lib.MK_ERR_RANGE
    after = g.export_csr()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
    g.close()
    pool.close()
    n, s, d, _, pool, g = build(10, 42, False, weighted=False)
    with pytest.raises(mk.MeerkatError) as e:
        g.UpdateEdges(s[:4], d[:4], np.ones(4, np.uint32))
    assert e.value.code == _lib.MK_ERR_BAD_ARG
    g.close()
    pool.close()


@pytest.mark.parametrize("reverse,persistent", [(True, 1), (False, 1), (True, 0)])
def test_sssp_reweight_matches_dijkstra(gpu, reverse, persistent):
    old = autotune.get_knob("persistent_steps")
    autotune.set_knob("persistent_steps", persistent)
    try:
        n, s, d, w, pool, g = build(14, 43, reverse)
        host = HostEdges(n, s, d, w)
        sp = mk.SSSP(g, 0)
        sp.Static()
        check(sp, host)
        rng = np.random.default_rng(7)
        for i in range(6):
            bs, bd, bw = weight_batch(rng, n, s, d, sp, 3000)
            if i == 1:
                bw = np.maximum(bw, 90).astype(np.uint32)  # increases only
            elif i == 2:
                bw = np.ones_like(bw)  # decreases only
            g.UpdateEdges(bs, bd, bw)
            host.update_weights(bs, bd, bw)
            sp.Reweight(bs, bd, bw)
            check(sp, host)
        # a deletion and an insertion after the reweights stay exact too
        ds, dd = gen.delete_batch(gen.keys_of(*_pairs(host.csr())), 2000, seed=9)
        g.DeleteEdges(ds, dd)
        host.delete(ds, dd)
        sp.Decremental(ds, dd)
        check(sp, host)
        sp.close()
        g.close()
        pool.close()
    finally:
        autotune.set_knob("persistent_steps", old)


def test_sssp_reweight_deferred_and_bfs_unaffected(gpu):
    n, s, d, w, pool, g = build(13, 44, True)
    host = HostEdges(n, s, d, w)
    sp = mk.SSSP(g, 0)
    bf = mk.BFS(g, 0)
    sp.Static()
    bf.Static()
    levels = bf.Distances()
    g.set_deferred(True)
    sp.set_deferred(True)
    rng = np.random.default_rng(8)
    keep = []
    for i in range(4):
        bs, bd, bw = weight_batch(rng, n, s, d, sp, 1500)
        dbs, dbd, dbw = mk.device_u32(bs), mk.device_u32(bd), mk.device_u32(bw)
        keep.append((dbs, dbd, dbw))
        g.UpdateEdges(dbs, dbd, dbw)
        sp.Reweight(dbs, dbd, dbw)
        host.update_weights(bs, bd, bw)
        check(sp, host)  # the read settles the deferred call
    assert sp.deferred_stats()["settled"] == 4
    g.set_deferred(False)
    # BFS levels do not depend on weights
    bf.Static()
    assert np.array_equal(bf.Distances(), levels)
    sp.close()
    bf.close()
    g.close()
    pool.close()


def test_sssp_reweight_rmat20(gpu):
    # RMAT-20 weighted (the SSSP workload's graph), 5 batches of 100,000
    # weight changes, checked after each
    n, s, d, w, pool, g = build(20, 3, True)
    host = HostEdges(n, s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    rng = np.random.default_rng(11)
    for i in range(5):
        bs, bd, bw = weight_batch(rng, n, s, d, sp, 100_000)
        g.UpdateEdges(bs, bd, bw)
        host.update_weights(bs, bd, bw)
        sp.Reweight(bs, bd, bw)
        ref = host.csr()
        dist, par = sp.read()
        exp = oracle.sssp_par(ref, 0, oracle.num_threads())
        assert np.array_equal(dist, exp), (i, int((dist != exp).sum()))
        assert oracle.check_parents_fast(ref, dist, par, 0, weights=True) == 0
    sp.close()
    g.close()
    pool.close()
