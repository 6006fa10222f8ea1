"""Direction-optimising static BFS: bottom-up steps over the in-edge mirror.

Levels bit-exact against the oracle's BFS and parents valid, with the
push/pull thresholds at their defaults, switched off, and forced to pull at
almost every step; on set and map graphs, with tombstones in the mirror, and
followed by incremental and decremental calls (which build on the tree the
bottom-up steps left).  The bottom-up steps must read fewer slabs than the
top-down traversal of the same graph.
"""
import numpy as np
import pytest

import meerkat_amd as mk
from meerkat_amd import autotune, gen
from test_gpu_traversal import EdgeSet, check_bfs

pytestmark = pytest.mark.gpu


@pytest.fixture
def knobs():
    saved = {k: autotune.get_knob(k) for k in ("pull_enter", "pull_exit")}
    yield autotune.set_knob
    for k, v in saved.items():
        autotune.set_knob(k, v)


def mirrored(scale, seed, weighted=False, ef=16):
    n, s, d, w = gen.rmat(scale, ef, seed=seed, weighted=weighted)
    pool = mk.SlabPool(1 << 29)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=weighted, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    return n, s, d, w, pool, g


def static_bfs(g, src):
    b = mk.BFS(g, src)
    b.Static()
    return b


@pytest.mark.parametrize("weighted", [False, True])
def test_pull_levels_match_oracle_and_read_less(gpu, knobs, weighted):
    n, s, d, w, pool, g = mirrored(16, 11, weighted)
    es = EdgeSet(s, d)
    isolated = int(np.nonzero(gen.degree_hints(n, s) == 0)[0][0])
    for src in (0, 77, isolated):
        knobs("pull_enter", 14)
        b = static_bfs(g, src)
        check_bfs(b, es, n, src)
        pulled = b.last_stats()["slabs_read"]
        knobs("pull_enter", 0)
        t = static_bfs(g, src)
        check_bfs(t, es, n, src)
        pushed = t.last_stats()["slabs_read"]
        assert np.array_equal(b.read()[0], t.read()[0])
        if src == 0:  # the hub's frontier goes bottom-up
            assert pulled < pushed, (pulled, pushed)
        b.close()
        t.close()
    g.close()
    pool.close()


def test_pull_forced_at_nearly_every_step(gpu, knobs):
    knobs("pull_enter", 4096)
    knobs("pull_exit", 4096)
    n, s, d, w, pool, g = mirrored(13, 12, ef=8)
    es = EdgeSet(s, d)
    for src in (0, 5, 999):
        b = static_bfs(g, src)
        check_bfs(b, es, n, src)
        b.close()
    g.close()
    pool.close()


def test_pull_with_tombstones_then_incremental_and_decremental(gpu, knobs):
    knobs("pull_enter", 14)
    n, s, d, w, pool, g = mirrored(15, 13, weighted=True)
    es = EdgeSet(s, d)
    # deletions leave tombstones in the mirror's slabs
    ds, dd = gen.delete_batch(es.keys, 60_000, seed=14)
    g.DeleteEdges(ds, dd)
    es.delete(ds, dd)
    b = static_bfs(g, 0)
    check_bfs(b, es, n, 0)
    for i in range(4):
        ds, dd = gen.delete_batch(es.keys, 5000, seed=20 + i)
        g.DeleteEdges(ds, dd)
        es.delete(ds, dd)
        b.Decremental(ds, dd)
        check_bfs(b, es, n, 0)
        bs, bd, bw = gen.insert_batch(es.keys, n, 5000, seed=30 + i, weighted=True, wseed=3)
        g.InsertEdges(bs, bd, bw)
        es.insert(bs, bd)
        b.Incremental(bs, bd)
        check_bfs(b, es, n, 0)
        b.Static()  # and again from scratch on the changed graph
        check_bfs(b, es, n, 0)
    b.close()
    g.close()
    pool.close()
