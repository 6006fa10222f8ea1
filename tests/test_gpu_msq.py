"""GPU parity of multi-source BFS / SSSP: every source slot equals the
oracle's single-source BFS / Dijkstra (bit-exact), for up to 64 sources,
repeated sources, hub sources and unreachable parts."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 7, 64])
def test_msbfs_equals_single_source(gpu, k):
    n, s, d, _ = gen.rmat(14, 16, seed=31)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    ref = oracle.csr(n, s, d)
    rng = np.random.default_rng(k)
    srcs = rng.integers(0, n, k).astype(np.uint32)
    srcs[0] = 0  # the hub
    if k > 2:
        srcs[2] = srcs[1]  # a repeated source
    q = mk.MultiSource(g)
    q.run(srcs)
    for slot, v in enumerate(srcs):
        assert np.array_equal(q.distances(slot), oracle.bfs(ref, int(v))), slot


@pytest.fixture
def unroll(request):
    from meerkat_amd import autotune

    before = autotune.get_knob("ms_unroll")
    autotune.set_knob("ms_unroll", request.param)
    yield request.param
    autotune.set_knob("ms_unroll", before)


@pytest.mark.parametrize("unroll", [4, 8], indirect=True)
@pytest.mark.parametrize("k", [1, 64])
def test_mssp_equals_dijkstra(gpu, k, unroll):
    n, s, d, w = gen.rmat(13, 16, seed=32, weighted=True)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    ref = oracle.csr(n, s, d, w)
    rng = np.random.default_rng(100 + k)
    srcs = rng.integers(0, n, k).astype(np.uint32)
    srcs[0] = 0
    q = mk.MultiSource(g, sssp=True)
    q.run(srcs)
    for slot, v in enumerate(srcs):
        assert np.array_equal(q.distances(slot), oracle.dijkstra(ref, int(v))), slot


def test_msbfs_reuse_across_runs_and_updates(gpu):
    # the same handle answers new batches after graph updates; the path graph
    # has as many levels as vertices
    n = 300
    s = np.arange(n - 1, dtype=np.uint32)
    d = np.arange(1, n, dtype=np.uint32)
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(n, pool, 0.7, None)
    g.InsertEdges(s, d)
    q = mk.MultiSource(g)
    q.run(np.array([0, 150, 299], np.uint32))
    assert np.array_equal(q.distances(0), np.arange(n, dtype=np.uint32))
    assert (q.distances(2)[:299] == mk.INF).all() and q.distances(2)[299] == 0
    g.InsertEdges(np.array([0], np.uint32), np.array([200], np.uint32))
    q.run(np.array([0], np.uint32))
    ref = oracle.csr(n, np.r_[s, 0], np.r_[d, 200])
    assert np.array_equal(q.distances(0), oracle.bfs(ref, 0))
    with pytest.raises(mk.MeerkatError):
        q.distances(1)  # slot not in the last run


def test_autotune_picks_a_setting_and_keeps_parity(gpu):
    from meerkat_amd import autotune

    before = autotune.knobs()
    n, s, d, w = gen.rmat(12, 16, seed=33, weighted=True)
    pool = mk.SlabPool(1 << 26)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    try:
        r = autotune.tune(g, waves=(8, 16), unrolls=(4, 8), deltas=(0, 40), tails=(0, 128),
                          ms_waves=(16, 24), repeats=1)
        assert r["waves_per_cu"] in (8, 16) and r["ms_unroll"] in (4, 8)
        assert r["ms_waves_per_cu"] in (16, 24)
        assert r["sssp_delta"] in (0, 40) and r["tail_enter"] in (0, 128)
        assert set(r["timings"]["waves_per_cu"]) == {8, 16}
        assert set(r["timings"]["sssp_delta"]) == {0, 40}
        assert "pull_enter/pull_exit" not in r["timings"]  # no in-edge mirror
        tuned = ("waves_per_cu", "ms_unroll", "ms_waves_per_cu", "sssp_delta", "tail_enter")
        k = autotune.knobs()
        assert all(k[x] == r[x] for x in tuned)
        assert {x: k[x] for x in k if x not in tuned} == {
            x: before[x] for x in before if x not in tuned}
        # a tuned static SSSP is still exact
        sp = mk.SSSP(g, 0)
        sp.Static()
        assert np.array_equal(sp.Distances(), oracle.dijkstra(oracle.csr(n, s, d, w), 0))
        sp.close()
        srcs = np.array([0, 5, 77], np.uint32)
        q = mk.MultiSource(g, sssp=True)
        q.run(srcs)
        ref = oracle.csr(n, s, d, w)
        for slot, v in enumerate(srcs):
            assert np.array_equal(q.distances(slot), oracle.dijkstra(ref, int(v))), slot
        q.close()
        # apply=False restores what was set before the call
        autotune.set_knob("waves_per_cu", 12)
        autotune.tune(g, waves=(8,), unrolls=(4,), ms_waves=(16,), repeats=1, apply=False)
        assert autotune.get_knob("waves_per_cu") == 12
    finally:
        for k_, v in before.items():
            autotune.set_knob(k_, v)


def test_autotune_push_pull_thresholds_and_load_factor(gpu):
    from meerkat_amd import autotune

    before = autotune.knobs()
    n, s, d, _ = gen.rmat(13, 16, seed=34)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d)
    try:
        r = autotune.tune(g, waves=(16,), pulls=((14, 24), (0, 24)), tails=(128,), repeats=1,
                          sources=np.zeros(64, np.uint32))
        assert (r["pull_enter"], r["pull_exit"]) in ((14, 24), (0, 24))
        assert set(r["timings"]["pull_enter/pull_exit"]) == {(14, 24), (0, 24)}
        b = mk.BFS(g, 0)
        b.Static()
        assert np.array_equal(b.Distances(), oracle.bfs(oracle.csr(n, s, d), 0))
        b.close()
        lf = autotune.tune_load_factor(n, s, d, factors=(0.5, 0.9), reverse=True, repeats=1,
                                       pool_bytes=1 << 27)
        assert lf["load_factor"] in (0.5, 0.9)
        t = lf["timings"]
        # a lower load factor spends more head slabs
        assert t[0.5]["slabs"]["head_slabs"] >= t[0.9]["slabs"]["head_slabs"]
    finally:
        g.close()
        pool.close()
        for k_, v in before.items():
            autotune.set_knob(k_, v)


def test_frontier_loops_beside_a_running_multi_source_query(gpu):
    # a multi-source SSSP on its own stream (a thread) while single-source
    # SSSP calls run their persistent loops on the graph's stream: with
    # another of the device's streams registered, the loops are launched
    # cooperatively (every block co-resident), and both answers stay exact.
    # Every wait is bounded (barrier_timeout_ms): a loop that could not
    # gather its blocks would fail its call, not hang.
    import threading

    n, s, d, w = gen.rmat(16, 16, seed=34, weighted=True)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    ref = oracle.csr(n, s, d, w)
    q = mk.MultiSource(g, sssp=True)
    sp = mk.SSSP(g, 0)
    srcs = np.random.default_rng(9).integers(0, n, 64).astype(np.uint32)
    errors = []

    def queries():
        try:
            for _ in range(3):
                q.run(srcs)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    th = threading.Thread(target=queries)
    th.start()
    for src in (0, 5, 17, 0):
        sp.set_source(src)
        sp.Static()
    th.join()
    assert not errors, errors
    assert np.array_equal(sp.Distances(), oracle.dijkstra(ref, 0))
    for slot in (0, 31, 63):
        assert np.array_equal(q.distances(slot), oracle.dijkstra(ref, int(srcs[slot]))), slot
    q.close()
    sp.close()
    g.close()
    pool.close()
