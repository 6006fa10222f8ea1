"""GPU parity of frontier steps run in tail mode (a small frontier expanded
by one block, step after step, frontiers in LDS) and of every way the tail
hands the frontier back to full-grid steps: the frontier grows past the
LDS queue (spill), reaches a hub vertex, or the epochs reserved per launch
run out (a long path).  Distances bit-exact against the oracle, parents
valid."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import gen

pytestmark = pytest.mark.gpu


def check(algo, n, s, d, w=None, source=0):
    ref = oracle.csr(n, s, d, w)
    dist, par = algo.read()
    exp = oracle.dijkstra(ref, source) if w is not None else oracle.bfs(ref, source)
    assert np.array_equal(dist, exp), int((dist != exp).sum())
    assert oracle.check_parents_fast(ref, dist, par, source, weights=w is not None) == 0


@pytest.mark.parametrize("weighted", [False, True])
def test_long_path_static_decremental_incremental(gpu, weighted, persistent):
    # 1000 levels: every step's frontier is one vertex, so the tail runs
    # ~254 steps per launch and hands over when its epoch range runs out
    n = 1000
    s = np.arange(n - 1, dtype=np.uint32)
    d = s + 1
    w = (np.arange(n - 1, dtype=np.uint32) % 7 + 1) if weighted else None
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(n, pool, 0.7, None, weighted=weighted)
    g.InsertEdges(s, d, w)
    algo = mk.SSSP(g, 0) if weighted else mk.BFS(g, 0)
    algo.Static()
    check(algo, n, s, d, w)
    assert algo.last_stats()["iterations"] >= n - 1
    # cut the path at 500: the tail invalidates 499 vertices one level a step
    cs, cd = np.array([500], np.uint32), np.array([501], np.uint32)
    g.DeleteEdges(cs, cd)
    algo.Decremental(cs, cd) if not weighted else algo.Decremental(cs, cd, None)
    keep = ~((s == 500) & (d == 501))
    s1, d1 = s[keep], d[keep]
    w1 = w[keep] if weighted else None
    check(algo, n, s1, d1, w1)
    # reconnect through a detour: 499 -> 700 and 10 -> 520
    bs, bd = np.array([499, 10], np.uint32), np.array([700, 520], np.uint32)
    bw = np.array([3, 5], np.uint32) if weighted else None
    g.InsertEdges(bs, bd, bw)
    if weighted:
        algo.Incremental(bs, bd, bw)
    else:
        algo.Incremental(bs, bd)
    s2, d2 = np.r_[s1, bs], np.r_[d1, bd]
    w2 = np.r_[w1, bw] if weighted else None
    check(algo, n, s2, d2, w2)


@pytest.fixture(params=[128, 0, 1024], ids=["tail128", "no_tail", "tail1024"])
def tail(request):
    from meerkat_amd import autotune
    before = autotune.get_knob("tail_enter")
    autotune.set_knob("tail_enter", request.param)
    yield request.param
    autotune.set_knob("tail_enter", before)


@pytest.mark.parametrize("hints", [False, True])
def test_broom_spill_and_hub_handover(gpu, hints, tail):
    # 0 -> 1 -> {2 .. 3001}, each leaf -> one of 50 tail vertices: the
    # frontier {1} (tail mode) pushes 3000 leaves, more than the LDS queue
    # holds (spill); with degree hints vertex 1 is a hub (hand-over)
    k = 3000
    leaves = np.arange(2, 2 + k, dtype=np.uint32)
    tails = (2 + k + leaves % 50).astype(np.uint32)
    s = np.r_[np.uint32(0), np.full(k, 1, np.uint32), leaves]
    d = np.r_[np.uint32(1), leaves, tails]
    n = 2 + k + 50
    h = gen.degree_hints(n, s) if hints else None
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(n, pool, 0.7, h)
    g.InsertEdges(s, d)
    b = mk.BFS(g, 0)
    b.Static()
    check(b, n, s, d)
    # delete the hand edge and put it back: decremental then incremental
    es, ed = np.array([0], np.uint32), np.array([1], np.uint32)
    g.DeleteEdges(es, ed)
    b.Decremental(es, ed)
    keep = ~((s == 0) & (d == 1))
    check(b, n, s[keep], d[keep])
    g.InsertEdges(es, ed)
    b.Incremental(es, ed)
    check(b, n, s, d)


@pytest.fixture(params=[1, 0], ids=["one_launch", "launch_per_step"])
def persistent(request):
    from meerkat_amd import autotune
    before = autotune.get_knob("persistent_steps")
    autotune.set_knob("persistent_steps", request.param)
    yield request.param
    autotune.set_knob("persistent_steps", before)


def test_small_batches_on_rmat_stay_exact(gpu, repair, persistent):
    # small update batches give small frontiers (tail mode) on a skewed graph
    n, s, d, w = gen.rmat(14, 8, seed=41, weighted=True)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    keys, wt = gen.keys_of(s, d), w.copy()
    for i in range(6):
        ds, dd = gen.delete_batch(keys, 20, seed=300 + i)
        g.DeleteEdges(ds, dd)
        sp.Decremental(ds, dd, None)
        keep = ~np.isin(keys, gen.keys_of(ds, dd))
        keys, wt = keys[keep], wt[keep]
        bs, bd, bw = gen.insert_batch(keys, n, 20, seed=400 + i, weighted=True, wseed=500 + i)
        g.InsertEdges(bs, bd, bw)
        sp.Incremental(bs, bd, bw)
        keys, wt = np.r_[keys, gen.keys_of(bs, bd)], np.r_[wt, bw]
        order = np.argsort(keys)
        keys, wt = keys[order], wt[order]
        s2, d2 = gen.split_keys(keys)
        check(sp, n, s2, d2, wt)


@pytest.fixture(params=[0, 1], ids=["no_repair", "repair"])
def repair(request):
    from meerkat_amd import autotune
    before = autotune.get_knob("orphan_repair")
    autotune.set_knob("orphan_repair", request.param)
    yield request.param
    autotune.set_knob("orphan_repair", before)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("mirror", [False, True])
def test_deleted_tree_edge_with_equal_alternative(gpu, weighted, mirror, repair):
    # 0 -> 1 -> 3, 0 -> 2 -> 3 (equal lengths), then a 200-vertex chain below
    # 3: deleting whichever edge 3's tree uses leaves every distance as it
    # was (with the mirror, 3 is re-parented and the chain never invalidated)
    chain = np.arange(3, 203, dtype=np.uint32)
    s = np.r_[np.array([0, 0, 1, 2], np.uint32), chain[:-1]]
    d = np.r_[np.array([1, 2, 3, 3], np.uint32), chain[1:]]
    w = np.r_[np.array([2, 3, 4, 3], np.uint32), np.full(len(chain) - 1, 5, np.uint32)]
    n = 203
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(n, pool, 0.7, None, weighted=weighted, reverse=mirror)
    g.InsertEdges(s, d, w if weighted else None)
    algo = mk.SSSP(g, 0) if weighted else mk.BFS(g, 0)
    algo.Static()
    ww = w if weighted else None
    check(algo, n, s, d, ww)
    before = algo.Distances().copy()
    p3 = int(algo.Parents()[3])
    assert p3 in (1, 2)
    # the batch names the edge twice: the orphan is still handled once
    es, ed = np.array([p3, p3], np.uint32), np.array([3, 3], np.uint32)
    g.DeleteEdges(es, ed)
    algo.Decremental(es, ed) if not weighted else algo.Decremental(es, ed, None)
    keep = ~((s == p3) & (d == 3))
    check(algo, n, s[keep], d[keep], ww[keep] if weighted else None)
    assert np.array_equal(algo.Distances(), before)
    assert int(algo.Parents()[3]) == 3 - p3  # the other middle vertex


def test_loops_on_two_graphs_and_forced_cooperative_launch(gpu):
    # handles on two graphs run their loops on two streams (cooperative
    # launches); coop_launch=1 forces them for a single graph too
    from meerkat_amd import autotune

    graphs = []
    for seed in (71, 72):
        n, s, d, w = gen.rmat(13, 8, seed=seed, weighted=True)
        pool = mk.SlabPool(1 << 26)
        g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
        g.InsertEdges(s, d, w)
        graphs.append((n, s, d, w, pool, g, mk.SSSP(g, 0)))
    for n, s, d, w, pool, g, sp in graphs:
        sp.Static()
        check(sp, n, s, d, w)
    for n, s, d, w, pool, g, sp in graphs:
        sp.close()
    before = autotune.get_knob("coop_launch")
    autotune.set_knob("coop_launch", 1)
    try:
        n, s, d, w, pool, g, _ = graphs[0]
        b = mk.BFS(g, 0)
        b.Static()
        check(b, n, s, d)
        b.close()
    finally:
        autotune.set_knob("coop_launch", before)


def test_barrier_wait_limit_ends_the_call_and_static_recovers(gpu):
    # a zero wait limit (forced from the host) makes the first grid barrier
    # give up: every wave still exits, the call returns MK_ERR_STATE, and a
    # handle created with the normal limit recomputes exactly afterwards
    from meerkat_amd import _lib, autotune

    n, s, d, w = gen.rmat(14, 8, seed=81, weighted=True)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    before = autotune.get_knob("barrier_timeout_ms")
    autotune.set_knob("barrier_timeout_ms", 0)
    try:
        sp = mk.SSSP(g, 0)
    finally:
        autotune.set_knob("barrier_timeout_ms", before)
    with pytest.raises(mk.MeerkatError) as ei:
        sp.Static()
    assert ei.value.code == _lib.MK_ERR_STATE
    sp.close()
    sp = mk.SSSP(g, 0)
    sp.Static()
    check(sp, n, s, d, w)


@pytest.mark.parametrize("wpc", [32, 64])
def test_widest_launch_shapes_keep_loop_barriers_consistent(gpu, wpc):
    # waves_per_cu at and beyond what a CU holds: the loop grid is capped
    # one block per CU below the occupancy figure, and the relax and
    # invalidation loops, which reach different occupancy limits but share
    # one call's barrier counters, take one grid (a grid not all resident,
    # or two grids, make a barrier wait time out with MK_ERR_STATE).
    # Weighted SSSP and BFS on the same graph cover both kernel variants.
    from meerkat_amd import autotune

    n, s, d, w = gen.rmat(13, 8, seed=91, weighted=True)
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    before = autotune.get_knob("waves_per_cu")
    autotune.set_knob("waves_per_cu", wpc)
    try:
        sp, bf = mk.SSSP(g, 0), mk.BFS(g, 0)
    finally:
        autotune.set_knob("waves_per_cu", before)
    sp.Static()
    check(sp, n, s, d, w)
    bf.Static()
    keys, wt = gen.keys_of(s, d), w.copy()
    ds, dd = gen.delete_batch(keys, 2000, seed=92)
    g.DeleteEdges(ds, dd)
    sp.Decremental(ds, dd, None)
    bf.Decremental(ds, dd)
    keep = ~np.isin(keys, gen.keys_of(ds, dd))
    keys, wt = keys[keep], wt[keep]
    bs, bd, bw = gen.insert_batch(keys, n, 2000, seed=93, weighted=True, wseed=94)
    g.InsertEdges(bs, bd, bw)
    sp.Incremental(bs, bd, bw)
    bf.Incremental(bs, bd)
    keys, wt = np.r_[keys, gen.keys_of(bs, bd)], np.r_[wt, bw]
    order = np.argsort(keys)
    s2, d2 = gen.split_keys(keys[order])
    check(sp, n, s2, d2, wt[order])
    check(bf, n, s2, d2)
