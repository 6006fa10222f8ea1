"""GPU parity of PageRank (FP32 ranks) against the oracle's FP64 PageRank.

Tolerance (stated per run): L1 distance between the GPU's FP32 ranks and a
from-scratch FP64 PageRank of the same graph, iterated to L1 delta 1e-14,
must be <= 1e-6 when the GPU stops at L1 delta 1e-7 (the configuration's
convergence threshold; the geometric tail after stopping is bounded by
d / (1 - d) * 1e-7 ~= 5.7e-7).

Both in-adjacency modes are covered: a graph of reversed edges with
host-given out-degrees (the reference's PageRank<RevUnweightedEdgePolicyT>
use, src/PageRank.cu:120-157), and a graph with an in-edge mirror.
"""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import gen

pytestmark = pytest.mark.gpu

TOL = 1e-6
D = 0.85
MARGIN = 1e-7


def fp64_reference(n, s, d):
    g = oracle.csr(n, s, d)
    pr, it, delta = oracle.pagerank(oracle.transpose(g), g.degrees(), D, 1e-14, 5000)
    assert delta <= 1e-14
    return pr


def l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum())


def test_reverse_graph_mode_static_and_incremental(gpu):
    # config 3 shape at reduced size: ER, mean out-degree 8, batches of new edges
    n, s, d, _ = gen.erdos_renyi(60000, 480000, seed=5)
    pool = mk.SlabPool(1 << 28)
    rev = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, d))
    rev.InsertEdges(d, s)  # reversed, as the reference driver does
    pr = mk.PageRank(rev, D, MARGIN, out_degree=gen.degree_hints(n, s), max_iter=100)
    pr.Static()
    err = l1(pr.Values(), fp64_reference(n, s, d))
    assert err <= TOL, err
    keys = gen.keys_of(s, d)
    for i in range(3):
        bs, bd, _ = gen.insert_batch(keys, n, 20000, seed=60 + i)
        bs_d, bd_d = mk.device_u32(bs), mk.device_u32(bd)
        rev.InsertEdges(bd_d, bs_d)
        pr.Incremental(bs_d, bd_d)  # same device batch: status path
        s, d = np.r_[s, bs], np.r_[d, bd]
        keys = gen.keys_of(s, d)
        err = l1(pr.Values(), fp64_reference(n, s, d))
        assert err <= TOL, (i, err)
        st = pr.last_stats()
        assert st["delta"] <= MARGIN or st["iterations"] == 100


def test_reverse_graph_mode_recount_path(gpu):
    # incremental called with a copy of the batch: out-degrees are recounted
    n, s, d, _ = gen.erdos_renyi(20000, 120000, seed=7)
    pool = mk.SlabPool(1 << 26)
    rev = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, d))
    rev.InsertEdges(d, s)
    pr = mk.PageRank(rev, D, MARGIN, out_degree=gen.degree_hints(n, s))
    pr.Static()
    bs, bd, _ = gen.insert_batch(gen.keys_of(s, d), n, 5000, seed=8)
    # keep the insert's device arrays alive so the copies get other addresses
    held = rev.InsertEdges(bd, bs)
    pr.Incremental(bs.copy(), bd.copy())
    del held
    s, d = np.r_[s, bs], np.r_[d, bd]
    assert l1(pr.Values(), fp64_reference(n, s, d)) <= TOL


def test_reduction_slots_cleared_over_dirty_device_memory(gpu):
    # device memory released with NaN bit patterns in it, then a new handle:
    # every reduction slot a compute call adds into must be cleared first
    import torch

    junk = torch.full((1 << 24,), float("nan"), device=gpu)
    torch.cuda.synchronize()
    del junk
    torch.cuda.empty_cache()
    n, s, d, _ = gen.erdos_renyi(5000, 30000, seed=11)
    pool = mk.SlabPool(1 << 24)
    rev = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, d))
    rev.InsertEdges(d, s)
    for _ in range(2):
        pr = mk.PageRank(rev, D, MARGIN, out_degree=gen.degree_hints(n, s))
        pr.Static()
        v = pr.Values()
        assert np.isfinite(v).all()
        assert l1(v, fp64_reference(n, s, d)) <= TOL
        pr.close()


@pytest.mark.parametrize("weighted", [False, True])
def test_mirror_mode_rmat_hubs_dangling_and_deletions(gpu, weighted):
    # RMAT: in-hubs (many in-buckets) and dangling vertices; out-degrees
    # come from the forward graph
    n, s, d, w = gen.rmat(15, 16, seed=11, weighted=True)
    pool = mk.SlabPool(1 << 29)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w if weighted else None)
    assert g.mirror().GetBucketsPerVertex().max() > 4  # hub path exercised
    assert (gen.degree_hints(n, s) == 0).any()  # dangling vertices present
    pr = mk.PageRank(g, D, MARGIN)
    pr.Static()
    assert l1(pr.Values(), fp64_reference(n, s, d)) <= TOL
    keys = gen.keys_of(s, d)
    ds, dd = gen.delete_batch(keys, 15000, seed=12)
    g.DeleteEdges(ds, dd)
    pr.Decremental(ds, dd)
    keep = ~np.isin(keys, gen.keys_of(ds, dd))
    s, d = s[keep], d[keep]
    assert l1(pr.Values(), fp64_reference(n, s, d)) <= TOL
    bs, bd, bw = gen.insert_batch(gen.keys_of(s, d), n, 15000, seed=13, weighted=True)
    g.InsertEdges(bs, bd, bw if weighted else None)
    pr.Incremental(bs, bd)
    s, d = np.r_[s, bs], np.r_[d, bd]
    assert l1(pr.Values(), fp64_reference(n, s, d)) <= TOL


def test_iteration_count_and_stopping_rule(gpu):
    # the loop runs while delta > margin and it < max_iter (PageRankImpl.cuh:173)
    n, s, d, _ = gen.erdos_renyi(5000, 40000, seed=21)
    pool = mk.SlabPool(1 << 24)
    rev = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, d))
    rev.InsertEdges(d, s)
    g = oracle.csr(n, s, d)
    for margin, max_iter in ((1e-3, 100), (1e-7, 100), (0.0, 7)):
        pr = mk.PageRank(rev, D, margin, out_degree=g.degrees(), max_iter=max_iter)
        pr.Static()
        st = pr.last_stats()
        _, it, delta = oracle.pagerank(oracle.transpose(g), g.degrees(), D, margin, max_iter)
        assert abs(st["iterations"] - it) <= 1, (margin, st, it)
        if margin == 0.0:
            assert st["iterations"] == 7


def test_cycle_is_uniform(gpu):
    n = 1000
    s = np.arange(n, dtype=np.uint32)
    d = ((np.arange(n) + 1) % n).astype(np.uint32)
    pool = mk.SlabPool(1 << 22)
    rev = mk.DynamicGraph(n, pool, 0.7, None)
    rev.InsertEdges(d, s)
    pr = mk.PageRank(rev, D, 1e-9, out_degree=np.ones(n, np.uint32))
    pr.Static()
    assert np.allclose(pr.Values(), 1.0 / n, rtol=1e-5)
