"""Near-far static SSSP (knob sssp_delta): vertices at or past the near
bound wait in a far list until the near frontier drains.

Distances bit-exact against the oracle's Dijkstra and parents valid for
several bucket widths (1 defers almost everything, 1000 almost nothing), with
far lists too small to hold the deferred vertices (small vertices are then
expanded at once; a hub that does not fit repeats the call without
near-far), on graphs whose distances jump past several bucket widths, and
followed by incremental and decremental calls on the tree it left.
"""
import numpy as np
import pytest

import meerkat_amd as mk
from meerkat_amd import autotune, gen
from test_gpu_traversal import EdgeSet, check_sssp

pytestmark = pytest.mark.gpu

KNOBS = ("sssp_delta", "sssp_far_cap")


@pytest.fixture
def knobs():
    saved = {k: autotune.get_knob(k) for k in KNOBS}
    yield autotune.set_knob
    for k, v in saved.items():
        autotune.set_knob(k, v)


def weighted_graph(scale, seed, reverse=False):
    n, s, d, w = gen.rmat(scale, 16, seed=seed, weighted=True)
    pool = mk.SlabPool(1 << 29)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=reverse,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    return n, s, d, w, pool, g


@pytest.mark.parametrize("delta", [1, 16, 64, 1000])
def test_near_far_static_matches_dijkstra(gpu, knobs, delta):
    knobs("sssp_delta", delta)
    n, s, d, w, pool, g = weighted_graph(15, 41)
    es = EdgeSet(s, d, w)
    for src in (0, 321):
        sp = mk.SSSP(g, src)
        sp.Static()
        check_sssp(sp, es, n, src)
        sp.close()
    g.close()
    pool.close()


@pytest.mark.parametrize("cap", [1, 64, 1000])
def test_near_far_with_full_far_lists(gpu, knobs, cap):
    knobs("sssp_delta", 8)
    knobs("sssp_far_cap", cap)
    n, s, d, w, pool, g = weighted_graph(14, 42)
    es = EdgeSet(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    check_sssp(sp, es, n, 0)
    sp.close()
    g.close()
    pool.close()


def test_near_far_distance_jumps(gpu, knobs):
    # a path whose weights (1000) jump past many bucket widths of 4, and a
    # heavy shortcut that a light detour beats
    knobs("sssp_delta", 4)
    n = 64
    s = np.r_[np.arange(n - 1), 0, 0, 40].astype(np.uint32)
    d = np.r_[np.arange(1, n), 50, 40, 50].astype(np.uint32)
    w = np.r_[np.full(n - 1, 1000), 60000, 1, 2].astype(np.uint32)
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(n, pool, 0.7, None, weighted=True)
    g.InsertEdges(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    check_sssp(sp, EdgeSet(s, d, w), n, 0)
    dist, _ = sp.read()
    assert dist[50] == 3 and dist[40] == 1 and dist[63] == 3 + 13 * 1000
    sp.close()
    g.close()
    pool.close()


def test_near_far_then_incremental_and_decremental(gpu, knobs):
    knobs("sssp_delta", 32)
    n, s, d, w, pool, g = weighted_graph(15, 43, reverse=True)
    es = EdgeSet(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    check_sssp(sp, es, n, 0)
    for i in range(3):
        ds, dd = gen.delete_batch(es.keys, 5000, seed=50 + i)
        g.DeleteEdges(ds, dd)
        es.delete(ds, dd)
        sp.Decremental(ds, dd)
        check_sssp(sp, es, n, 0)
        bs, bd, bw = gen.insert_batch(es.keys, n, 5000, seed=60 + i, weighted=True, wseed=7)
        g.InsertEdges(bs, bd, bw)
        es.insert(bs, bd, bw)
        sp.Incremental(bs, bd, bw)
        check_sssp(sp, es, n, 0)
    sp.Static()
    check_sssp(sp, es, n, 0)
    sp.close()
    g.close()
    pool.close()
