"""GPU parity of connected components (union-find) against the oracle:
labels bit-exact, each the smallest vertex id of its component."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from conftest import GOLDEN_NAMES, load_golden
from meerkat_amd import gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_golden_components(gpu, name):
    z = load_golden(name)
    n = int(z["n"])
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(n, pool, 0.7, None)
    g.InsertEdges(z["raw_src"], z["raw_dst"])
    cc = mk.CC(g)
    cc.Static()
    assert np.array_equal(cc.Labels(), z["cc"])


@pytest.mark.parametrize("weighted", [False, True])
def test_static_incremental_decremental_config4_shape(gpu, weighted):
    # config 4 shape at scale 16: undirected RMAT, batches of new undirected
    # edges inserted in both directions; sparse enough to have many components
    n, s, d, w = gen.rmat(16, 2, seed=7, undirected=True, weighted=True)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted)
    g.InsertEdges(s, d, w if weighted else None)
    cc = mk.CC(g)
    cc.Static()
    labels, k = oracle.cc(oracle.csr(n, s, d))
    assert k > 1000
    assert np.array_equal(cc.Labels(), labels)
    for i in range(4):
        bs, bd, bw = gen.insert_batch(gen.keys_of(s, d), n, 3000, seed=70 + i, weighted=True)
        S, D = np.r_[bs, bd], np.r_[bd, bs]
        g.InsertEdges(S, D, np.r_[bw, bw] if weighted else None)
        cc.Incremental(S, D)
        s, d = np.r_[s, S], np.r_[d, D]
        assert np.array_equal(cc.Labels(), oracle.cc(oracle.csr(n, s, d))[0]), i
    keys = gen.keys_of(s, d)
    ds, dd = gen.delete_batch(keys, 20000, seed=80)
    g.DeleteEdges(ds, dd)
    cc.Decremental(ds, dd)
    keep = ~np.isin(keys, gen.keys_of(ds, dd))
    s2, d2 = gen.split_keys(keys[keep])
    assert np.array_equal(cc.Labels(), oracle.cc(oracle.csr(n, s2, d2))[0])


def test_directed_graph_weak_components(gpu):
    # a directed chain 9 -> 8 -> ... -> 0 plus isolated vertices: one weak
    # component labelled 0, the rest singletons
    n = 20
    s = np.arange(9, 0, -1, dtype=np.uint32)
    d = np.arange(8, -1, -1, dtype=np.uint32)
    pool = mk.SlabPool(1 << 20)
    g = mk.DynamicGraph(n, pool, 0.7, None)
    g.InsertEdges(s, d)
    cc = mk.CC(g)
    cc.Static()
    lab = cc.Labels()
    assert (lab[:10] == 0).all() and np.array_equal(lab[10:], np.arange(10, 20))
    # joining the singletons 15 and 12 with one inserted edge
    cc.Incremental(np.array([15], np.uint32), np.array([12], np.uint32))
    lab = cc.Labels()
    assert lab[15] == 12 and lab[12] == 12
