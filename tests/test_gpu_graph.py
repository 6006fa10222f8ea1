"""GPU parity of the slab pool and slab-hash adjacency against the oracle.

Restates the reference's test/GraphTest.cu cases (DegreeCountTest :647-714,
ElementTest :716-871, PartialDelete :873-956, FullDelete :958-1007,
WGraphTest :1111-1215) on seeded generated graphs, plus the semantics the
reference leaves to its absent SlabHash submodule: duplicate suppression
under concurrency, tombstone reuse, pool exhaustion and range errors.
"""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import _lib, gen

pytestmark = pytest.mark.gpu


def build(n, s, d, w=None, hints=True, weighted=False, pool_bytes=1 << 28, lf=0.7):
    pool = mk.SlabPool(pool_bytes)
    h = gen.degree_hints(n, s) if hints else None
    g = mk.DynamicGraph(n, pool, lf, h, weighted=weighted)
    if len(s):
        g.InsertEdges(s, d, w)
    return pool, g


def assert_same_adjacency(g, ref):
    row, col, w = g.export_csr()
    assert np.array_equal(row, ref.row_off)
    assert np.array_equal(col, ref.col)
    if ref.w is not None:
        assert np.array_equal(w, ref.w)


@pytest.fixture(scope="module")
def rmat12(gpu):
    n, s, d, _ = gen.rmat(12, 16, seed=1)
    return n, s, d


def test_degree_and_elements_with_duplicates_and_self_loops(gpu):
    # raw numpy-drawn edges with duplicates and self-loops (golden fixture)
    from conftest import load_golden

    z = load_golden("dense_directed")
    n = int(z["n"])
    ref = oracle.csr(n, z["raw_src"], z["raw_dst"])
    pool, g = build(n, z["raw_src"], z["raw_dst"], hints=False)
    live, tomb = g.count_slots()
    assert np.array_equal(live, ref.degrees())  # DegreeCountTest
    assert np.array_equal(g.degrees(), ref.degrees())
    assert tomb.sum() == 0
    assert_same_adjacency(g, ref)  # ElementTest
    st = g.last_status()
    self_loops = z["raw_src"] == z["raw_dst"]
    assert (st[self_loops] == _lib.EDGE_SELF_LOOP).all()
    assert (st == _lib.EDGE_APPLIED).sum() == ref.m


@pytest.mark.parametrize("hints", [True, False])
def test_partial_then_full_delete(rmat12, hints):
    n, s, d = rmat12
    ref = oracle.csr(n, s, d)
    pool, g = build(n, s, d, hints=hints)
    assert_same_adjacency(g, ref)
    rng = np.random.default_rng(45)
    mark = rng.random(len(s)) < 0.45  # GraphTest.cu:520-529 marks ~45%
    g.DeleteEdges(s[mark], d[mark])
    assert (g.last_status() == _lib.EDGE_APPLIED).all()
    rest = oracle.csr(n, s[~mark], d[~mark])
    assert_same_adjacency(g, rest)  # PartialDelete
    assert np.array_equal(g.degrees(), rest.degrees())
    g.DeleteEdges(s[~mark], d[~mark])
    live, tomb = g.count_slots()  # FullDelete: tombstones == original degree
    assert live.sum() == 0
    assert np.array_equal(tomb, ref.degrees())
    assert g.num_edges() == 0
    # deleting again changes nothing
    g.DeleteEdges(s[:100], d[:100])
    assert (g.last_status() == _lib.EDGE_UNCHANGED).all()


def test_tombstones_are_reused(rmat12):
    n, s, d = rmat12
    pool, g = build(n, s, d, hints=False)
    before = g.slab_stats()
    used_before, _ = pool.stats()
    g.DeleteEdges(s, d)
    g.InsertEdges(s, d)
    assert g.slab_stats()["overflow_slabs"] == before["overflow_slabs"]
    assert pool.stats()[0] == used_before
    assert_same_adjacency(g, oracle.csr(n, s, d))


def test_concurrent_duplicates_inserted_once(gpu):
    n, s, d, _ = gen.rmat(11, 16, seed=9)
    reps = 8
    rng = np.random.default_rng(1)
    perm = rng.permutation(len(s) * reps)
    S = np.tile(s, reps)[perm]
    D = np.tile(d, reps)[perm]
    for hints in (True, False):
        pool, g = build(n, S, D, hints=hints)
        assert (g.last_status() == _lib.EDGE_APPLIED).sum() == len(s)
        assert np.array_equal(g.degrees(), oracle.csr(n, s, d).degrees())
        assert_same_adjacency(g, oracle.csr(n, s, d))


@pytest.mark.parametrize("hints", [True, False])
def test_repeated_deletes_and_reinserts_on_a_hub(gpu, hints):
    # every key of one batch repeated, in random order, on a hub whose keys
    # share few buckets (without hints: one chain of ~50 slabs): a repeated
    # delete applies once; re-inserting deleted keys, each several times,
    # applies each once into the freed slots, beside new keys
    n, h = 5000, 3000
    rng = np.random.default_rng(12)
    s = np.r_[np.zeros(h, np.uint32), rng.integers(1, n, 20000).astype(np.uint32)]
    d = np.r_[np.arange(1, h + 1, dtype=np.uint32), rng.integers(0, n, 20000).astype(np.uint32)]
    keep = s != d
    s, d = s[keep], d[keep]
    pool, g = build(n, s, d, hints=hints)
    ref = oracle.csr(n, s, d)
    assert_same_adjacency(g, ref)
    gone = np.arange(1, h + 1, 2, dtype=np.uint32)  # every other hub key
    perm = rng.permutation(len(gone) * 6)
    g.DeleteEdges(np.zeros(len(perm), np.uint32), np.tile(gone, 6)[perm])
    st = g.last_status()
    assert (st == _lib.EDGE_APPLIED).sum() == len(gone)
    assert ((st == _lib.EDGE_APPLIED) | (st == _lib.EDGE_UNCHANGED)).all()
    live = np.ones(len(s), bool)
    live[:h][np.isin(d[:h], gone)] = False
    assert_same_adjacency(g, oracle.csr(n, s[live], d[live]))
    used = pool.stats()[0]
    back = np.r_[np.tile(gone, 6), np.tile(np.arange(h + 1, h + 401, dtype=np.uint32), 4)]
    perm = rng.permutation(len(back))
    g.InsertEdges(np.zeros(len(back), np.uint32), back[perm])
    assert (g.last_status() == _lib.EDGE_APPLIED).sum() == len(gone) + 400
    s2 = np.r_[s, np.zeros(400, np.uint32)]
    d2 = np.r_[d, np.arange(h + 1, h + 401, dtype=np.uint32)]
    assert_same_adjacency(g, oracle.csr(n, s2, d2))
    assert np.array_equal(g.degrees(), oracle.csr(n, s2, d2).degrees())
    # the freed slots took the re-inserted keys: 400 new keys need at most
    # ceil(400 / 63) more slabs in the hub's chains
    assert pool.stats()[0] - used <= -(-400 // 63)
    g.close()
    pool.close()


def test_weighted_pairs_and_keep_first_weight(gpu):
    n, s, d, w = gen.rmat(11, 16, seed=3, weighted=True)
    ref = oracle.csr(n, s, d, w)
    pool, g = build(n, s, d, w, weighted=True)
    assert_same_adjacency(g, ref)  # WGraphTest.ElementTest
    live, _ = g.count_slots()
    assert np.array_equal(live, ref.degrees())  # WGraphTest.DegreeCountTest
    # re-inserting existing edges with other weights changes nothing
    g.InsertEdges(s[:500], d[:500], (w[:500] % 7) + 200)
    assert (g.last_status() == _lib.EDGE_UNCHANGED).all()
    assert_same_adjacency(g, ref)
    found, ww = g.SearchEdges(s[:1000], d[:1000])
    assert found.all() and np.array_equal(ww, w[:1000])
    missing = gen.insert_batch(gen.keys_of(s, d), n, 300, seed=4)
    found, ww = g.SearchEdges(missing[0], missing[1])
    assert not found.any() and (ww == 0xFFFFFFFF).all()


def test_long_chains_hub_without_hints(gpu):
    n = 20000
    hub = np.zeros(15000, np.uint32)
    nbr = np.arange(1, 15001, dtype=np.uint32)
    pool, g = build(n, hub, nbr, hints=False)
    st = g.slab_stats()
    assert st["max_chain"] == int(np.ceil(15000 / 63))
    assert pool.stats()[0] == st["overflow_slabs"]
    assert_same_adjacency(g, oracle.csr(n, hub, nbr))
    g.close()
    assert pool.stats()[0] == 0  # destroy returns every overflow slab


def test_pool_exhaustion_leaves_edge_set_unchanged(gpu):
    n = 200000
    pool, g = build(n, np.zeros(0, np.uint32), np.zeros(0, np.uint32), hints=False,
                    pool_bytes=2048 * 256)
    first = np.arange(1, 1001, dtype=np.uint32)
    g.InsertEdges(np.zeros(1000, np.uint32), first)
    before = g.export_csr()
    hub = np.zeros(150000, np.uint32)
    nbr = np.arange(1, 150001, dtype=np.uint32)
    with pytest.raises(mk.MeerkatError) as ei:
        g.InsertEdges(hub, nbr)
    assert ei.value.code == _lib.MK_ERR_POOL_FULL
    after = g.export_csr()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert g.num_edges() == 1000
    # the next batch that fits is applied: the failure count belongs to the
    # failed call only
    g.InsertEdges(np.full(10, 7, np.uint32), np.arange(100, 110, dtype=np.uint32))
    assert g.num_edges() == 1010


def test_out_of_range_batch_rejected(rmat12):
    n, s, d = rmat12
    pool, g = build(n, s, d)
    bad_s = np.array([1, 2, n], np.uint32)
    bad_d = np.array([3, 4, 5], np.uint32)
    with pytest.raises(mk.MeerkatError) as ei:
        g.InsertEdges(bad_s, bad_d)
    assert ei.value.code == _lib.MK_ERR_RANGE
    with pytest.raises(mk.MeerkatError) as ei:
        g.DeleteEdges(bad_d, bad_s)
    assert ei.value.code == _lib.MK_ERR_RANGE
    assert "1 edges" in str(ei.value)
    assert_same_adjacency(g, oracle.csr(n, s, d))
    # a later valid batch is applied (the failure belongs to its own call),
    # and a later failure is detected again
    ok_s, ok_d = np.array([1, 2], np.uint32), np.array([3, 4], np.uint32)
    g.InsertEdges(ok_s, ok_d)
    with pytest.raises(mk.MeerkatError) as ei:
        g.InsertEdges(np.array([n + 5, 1], np.uint32), np.array([0, n], np.uint32))
    assert ei.value.code == _lib.MK_ERR_RANGE and "2 edges" in str(ei.value)
    keys = np.unique(np.r_[gen.keys_of(s, d), gen.keys_of(ok_s, ok_d)])
    s2, d2 = gen.split_keys(keys)
    assert_same_adjacency(g, oracle.csr(n, s2, d2))


def test_empty_batches_and_handle_order(gpu):
    pool, g = build(100, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    g.InsertEdges(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    g.DeleteEdges(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert g.num_edges() == 0
    with pytest.raises(mk.MeerkatError) as ei:
        pool.close()  # a graph still uses it
    assert ei.value.code == _lib.MK_ERR_STATE
    g.close()
    pool.close()


@pytest.mark.parametrize("weighted", [False, True])
def test_in_edge_mirror_tracks_every_mutation(gpu, weighted):
    n, s, d, w = gen.rmat(12, 16, seed=21, weighted=True)
    w = w if weighted else None
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    rng = np.random.default_rng(5)
    mark = rng.random(len(s)) < 0.3
    g.DeleteEdges(s[mark], d[mark])
    bs, bd, bw = gen.insert_batch(gen.keys_of(s, d), n, 3000, seed=6, weighted=True)
    g.InsertEdges(bs, bd, bw if weighted else None)
    S = np.r_[s[~mark], bs]
    D = np.r_[d[~mark], bd]
    Wt = None if not weighted else np.r_[w[~mark], bw]
    fwd = oracle.csr(n, S, D, Wt)
    rev = oracle.csr(n, D, S, Wt)  # transpose, with the same weights
    assert_same_adjacency(g, fwd)
    assert_same_adjacency(g.mirror(), rev)
    assert np.array_equal(g.mirror().degrees(), rev.degrees())
    with pytest.raises(mk.MeerkatError) as ei:  # the mirror cannot be mutated directly
        _lib.call("mk_graph_insert", g.mirror().handle, None, None, None, 0, None)
    assert ei.value.code == _lib.MK_ERR_STATE


def test_mirror_rolls_back_with_the_graph(gpu):
    n = 200000
    pool = mk.SlabPool(2048 * 256)  # one block: the hub chain needs 2381 slabs
    g = mk.DynamicGraph(n, pool, 0.7, None, reverse=True)
    g.InsertEdges(np.zeros(500, np.uint32), np.arange(1, 501, dtype=np.uint32))
    before_f, before_r = g.export_csr(), g.mirror().export_csr()
    with pytest.raises(mk.MeerkatError) as ei:
        g.InsertEdges(np.zeros(150000, np.uint32), np.arange(1, 150001, dtype=np.uint32))
    assert ei.value.code == _lib.MK_ERR_POOL_FULL
    after_f, after_r = g.export_csr(), g.mirror().export_csr()
    assert np.array_equal(before_f[1], after_f[1]) and np.array_equal(before_r[1], after_r[1])
    assert np.array_equal(before_r[0], after_r[0])


@pytest.mark.parametrize("mirror", [False, True])
@pytest.mark.parametrize("count", [1, 100, 4097, 300_001])
def test_status_of_every_edge_when_the_grid_has_spare_waves(gpu, mirror, count):
    # batch sizes whose wave count leaves spare waves in the last block: every
    # new edge must report APPLIED (insert) and APPLIED again (delete), and
    # no edge may be applied twice
    n = 50_000
    rng = np.random.default_rng(count)
    keys = np.unique(rng.integers(0, n, 2 * count + 10, dtype=np.uint64) * np.uint64(n)
                     + rng.integers(0, n, 2 * count + 10, dtype=np.uint64))
    s = (keys // np.uint64(n)).astype(np.uint32)
    d = (keys % np.uint64(n)).astype(np.uint32)
    keep = s != d
    s, d = s[keep][:count], d[keep][:count]
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, None, reverse=mirror)
    sd, dd = mk.device_u32(s), mk.device_u32(d)
    g.InsertEdges(sd, dd)
    st = g.last_status()
    assert len(st) == len(s) and (st == _lib.EDGE_APPLIED).all(), np.bincount(st)
    assert g.num_edges() == len(s)
    if mirror:
        assert g.mirror().num_edges() == len(s)
    g.DeleteEdges(sd, dd)
    st = g.last_status()
    assert (st == _lib.EDGE_APPLIED).all(), np.bincount(st)
    assert g.num_edges() == 0


@pytest.mark.parametrize("weighted,lf,hints,expected", [
    # set slabs hold 63 keys: buckets = max(1, ceil(hint / (63 * lf)))
    (False, 1.0, [0, 1, 63, 64, 126, 127, 1000], [1, 1, 1, 2, 2, 3, 16]),
    (False, 0.5, [0, 31, 32, 63, 64], [1, 1, 2, 2, 3]),
    # map slabs hold 31 (key, weight) pairs
    (True, 1.0, [0, 31, 32, 62, 63], [1, 1, 2, 2, 3]),
    (True, 0.7, [21, 22, 43, 44], [1, 2, 2, 3]),
])
def test_bucket_count_rule(gpu, weighted, lf, hints, expected):
    # the reference's rule (include/Graph/SlabGraph.cuh:404-406) with the
    # keys per 256-byte slab of this layout
    n = len(hints)
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(n, pool, lf, np.array(hints, np.uint32), weighted=weighted)
    assert g.GetBucketsPerVertex().tolist() == expected
    off = g.GetVertexBucketsOffsets()
    assert off[0] == 0 and np.array_equal(np.diff(off), expected)
    g.close()
    pool.close()


def test_bucket_count_rule_random_hints(gpu):
    rng = np.random.default_rng(8)
    hints = rng.integers(0, 5000, 20000).astype(np.uint32)
    pool = mk.SlabPool(1 << 26)
    for weighted, keys in ((False, 63), (True, 31)):
        for lf in (0.3, 0.7, 1.0, 2.0):
            g = mk.DynamicGraph(len(hints), pool, lf, hints, weighted=weighted)
            exp = np.maximum(1, np.ceil(hints / (keys * lf))).astype(np.int64)
            assert np.array_equal(g.GetBucketsPerVertex().astype(np.int64), exp), (weighted, lf)
            g.close()
    pool.close()
