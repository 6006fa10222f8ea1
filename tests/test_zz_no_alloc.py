"""mk_graph_count_slots_noalloc (the entry point that allocates nothing)
against mk_graph_count_slots, through the C ABI, on two calls in a row (the
second on a smaller graph, so stale counts of the first would show), and
its refusal of a graph larger than its fixed scratch."""
import ctypes

import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import _lib, gen

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def counts(name, g):
    live = np.full(g.n, 0xDEADBEEF, np.uint32)
    tomb = np.full(g.n, 0xDEADBEEF, np.uint32)
    _lib.call(name, g._h, live.ctypes.data_as(_vp), tomb.ctypes.data_as(_vp))
    return live, tomb


def graph_with_tombstones(scale, seed, weighted):
    n, s, d, w = gen.rmat(scale, 16, seed=seed, weighted=weighted)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted)
    g.InsertEdges(s, d, w)
    rng = np.random.default_rng(seed)
    k = rng.choice(len(s), size=len(s) // 3, replace=False)
    g.DeleteEdges(s[k], d[k])
    keep = np.ones(len(s), bool)
    keep[k] = False
    ref = oracle.csr(n, s[keep], d[keep], None if w is None else w[keep])
    return pool, g, ref, len(k)


@pytest.mark.parametrize("weighted", [False, True])
def test_noalloc_equals_count_slots_twice(gpu, weighted):
    big = graph_with_tombstones(14, 3, weighted)
    small = graph_with_tombstones(11, 4, weighted)
    try:
        for pool, g, ref, deleted in (big, small):
            live, tomb = counts("mk_graph_count_slots_noalloc", g)
            live_a, tomb_a = counts("mk_graph_count_slots", g)
            assert np.array_equal(live, live_a) and np.array_equal(tomb, tomb_a)
            assert np.array_equal(live, ref.degrees())
            assert int(tomb.sum()) == deleted
    finally:
        for pool, g, _, _ in (small, big):
            g.close()
            pool.close()


def test_noalloc_refuses_graphs_beyond_its_scratch(gpu):
    n = (1 << 22) + 1
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(n, pool, 0.7, None)
    try:
        live = np.empty(n, np.uint32)
        with pytest.raises(_lib.MeerkatError) as e:
            _lib.call("mk_graph_count_slots_noalloc", g._h, live.ctypes.data_as(_vp), None)
        assert e.value.code == _lib.MK_ERR_BAD_ARG
        assert "4194304" in str(e.value)
    finally:
        g.close()
        pool.close()
