"""GPU parity of triangle counting against the oracle (bit-exact counts):
static, incremental and decremental on undirected graphs stored in both
directions, through both the LDS-staged and the overflow path."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from conftest import GOLDEN_NAMES, load_golden
from meerkat_amd import gen

pytestmark = pytest.mark.gpu


def sym(s, d):
    return np.r_[s, d], np.r_[d, s]


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_golden_triangles(gpu, name):
    z = load_golden(name)
    n = int(z["n"])
    S, D = sym(z["src"], z["dst"])
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(n, pool, 0.7, None)
    g.InsertEdges(S, D)
    tc = mk.TC(g)
    tc.Static()
    assert tc.GetCount() == int(z["tc"])


@pytest.mark.parametrize("weighted,hints", [(False, True), (True, True), (False, False)])
def test_static_incremental_decremental_config4_shape(gpu, weighted, hints):
    # config 4 shape at scale 14: undirected RMAT, batches of new undirected
    # edges (both directions), then deletions of existing ones
    n, s, d, w = gen.rmat(14, 8, seed=7, undirected=True, weighted=True)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s) if hints else None,
                        weighted=weighted)
    g.InsertEdges(s, d, w if weighted else None)
    tc = mk.TC(g)
    tc.Static()
    assert tc.GetCount() == oracle.tc(oracle.csr(n, s, d))
    keys = gen.keys_of(s, d)
    for i in range(3):
        bs, bd, bw = gen.insert_batch(keys, n, 4000, seed=90 + i, weighted=True)
        # drop pairs whose reverse is also drawn, so each undirected edge is new once
        S, D = sym(bs, bd)
        Wt = np.r_[bw, bw]
        sd, dd, wd = mk.device_u32(S), mk.device_u32(D), mk.device_u32(Wt)
        g.InsertEdges(sd, dd, wd if weighted else None)
        tc.Incremental(sd, dd)
        s, d = np.r_[s, S], np.r_[d, D]
        keys = np.unique(gen.keys_of(s, d))
        s, d = gen.split_keys(keys)
        assert tc.GetCount() == oracle.tc(oracle.csr(n, s, d)), i
    for i in range(3):
        ds, dd = gen.delete_batch(keys, 3000, seed=95 + i)
        S, D = sym(ds, dd)
        sd, dd_ = mk.device_u32(S), mk.device_u32(D)
        g.DeleteEdges(sd, dd_)
        tc.Decremental(sd, dd_)
        keys = keys[~np.isin(keys, gen.keys_of(S, D))]
        s, d = gen.split_keys(keys)
        assert tc.GetCount() == oracle.tc(oracle.csr(n, s, d)), i


@pytest.mark.parametrize("k", [600, 1500])
def test_clique_counts(gpu, k):
    # a k-clique: vertex 0 has k - 1 higher-ranked neighbours; at k = 1500
    # they fill more than one 1024-key LDS table; C(k, 3) triangles
    a, b = np.meshgrid(np.arange(k, dtype=np.uint32), np.arange(k, dtype=np.uint32))
    m = a != b
    s, d = a[m].ravel(), b[m].ravel()
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(k + 10, pool, 0.7, np.r_[np.full(k, k - 1), np.zeros(10)].astype(np.uint32))
    g.InsertEdges(s, d)
    tc = mk.TC(g)
    tc.Static()
    assert tc.GetCount() == k * (k - 1) * (k - 2) // 6


def test_batch_copy_without_status(gpu):
    # a copy of the batch (other device pointers): every entry counts as new,
    # which is exact for a batch of new edges
    n, s, d, _ = gen.rmat(12, 8, seed=3, undirected=True)
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    tc = mk.TC(g)
    tc.Static()
    bs, bd, _ = gen.insert_batch(gen.keys_of(s, d), n, 2000, seed=5)
    S, D = sym(bs, bd)
    g.InsertEdges(S, D)
    tc.Incremental(S.copy(), D.copy())
    keys = np.unique(np.r_[gen.keys_of(s, d), gen.keys_of(S, D)])
    s2, d2 = gen.split_keys(keys)
    assert tc.GetCount() == oracle.tc(oracle.csr(n, s2, d2))
