"""Serving loop: every query answered in the continuous-batching mix equals
the same query run alone (the oracle's static BFS / Dijkstra) on the graph
version it ran on; batching beats one-at-a-time on the same stream; update
batches are admitted against free slab-pool memory."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import gen, serving

pytestmark = pytest.mark.gpu


def build(scale, seed, pool_bytes=1 << 28):
    n, s, d, w = gen.rmat(scale, 16, seed=seed, weighted=True)
    pool = mk.SlabPool(pool_bytes)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    return n, s, d, w, pool, g


def replay_versions(n, s, d, w, reqs):
    """Host edge sets (oracle CSR) after each applied update version."""
    ups = sorted((r for r in reqs if r.kind == "update"), key=lambda r: r.version)
    keys = gen.keys_of(s, d)
    wmap = dict(zip(keys.tolist(), w.tolist()))
    versions = {0: oracle.csr(n, s, d, w)}
    for r in ups:
        for k in gen.keys_of(*r.dele).tolist():
            wmap.pop(k, None)
        ins_s, ins_d, ins_w = r.ins
        for k, ww in zip(gen.keys_of(ins_s, ins_d).tolist(), ins_w.tolist()):
            if (k >> 32) != (k & 0xFFFFFFFF) and k not in wmap:
                wmap[k] = ww
        kk = np.array(sorted(wmap), dtype=np.uint64)
        vs, vd = gen.split_keys(kk)
        vw = np.array([wmap[x] for x in kk.tolist()], np.uint32)
        versions[r.version] = oracle.csr(n, vs, vd, vw)
    return versions


@pytest.mark.parametrize("lockstep", [False, True])
def test_queries_in_the_mix_equal_queries_alone(gpu, lockstep):
    # lockstep: both engines take a step, then the loop waits for both
    n, s, d, w, pool, g = build(12, 41)
    keys = gen.keys_of(s, d)
    reqs = serving.make_stream(n, keys, 150, rate=20000.0, seed=3, median=300, max_size=3000)
    loop = serving.ServingLoop(g, keep_results=True, lockstep=lockstep)
    done = loop.run(reqs, mode="batched")
    assert all(r.done is not None for r in done)
    versions = replay_versions(n, s, d, w, done)
    checked = 0
    for r in done:
        if r.kind == "update":
            continue
        ref = versions[r.version]
        exp = oracle.bfs(ref, r.src) if r.kind == "bfs" else oracle.dijkstra(ref, r.src)
        assert np.array_equal(r.result, exp), (r.kind, r.src, r.version)
        checked += 1
    assert checked > 50
    # several queries shared a launch (batching happened)
    starts = [r.start for r in done if r.kind != "update"]
    assert len(set(starts)) < len(starts)


def test_batched_beats_one_at_a_time(gpu):
    n, s, d, w, pool, g = build(14, 42)
    keys = gen.keys_of(s, d)
    mk_reqs = lambda: serving.make_stream(n, keys, 200, rate=1.0, seed=5, median=2000,
                                          max_size=20000, all_at_once=True)
    loop = serving.ServingLoop(g)
    seq = loop.run(mk_reqs(), mode="sequential")
    t_seq = max(r.done for r in seq)
    bat = loop.run(mk_reqs(), mode="batched")
    t_bat = max(r.done for r in bat)
    assert t_bat < t_seq, (t_bat, t_seq)


def test_update_admission_against_pool(gpu):
    n = 5000
    pool = mk.SlabPool(2048 * 256)  # one block of 2048 slabs
    g = mk.DynamicGraph(n, pool, 0.7, None, weighted=True, reverse=True)
    loop = serving.ServingLoop(g)
    big = serving.Request("update", 0.0,
                          ins=(np.zeros(200000, np.uint32) + 1, np.arange(200000, dtype=np.uint32) % n,
                               np.ones(200000, np.uint32)),
                          dele=(np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    with pytest.raises(RuntimeError):
        loop.run([big], mode="sequential")
    small = serving.Request("update", 0.0,
                            ins=(np.array([1, 2], np.uint32), np.array([3, 4], np.uint32),
                                 np.array([5, 6], np.uint32)),
                            dele=(np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    q = serving.Request("bfs", 0.0, src=1)
    out = loop.run([small, q], mode="batched")
    assert out[0].version == 1 and out[1].done is not None


def test_sequential_mode_answers_match_oracle(gpu):
    n, s, d, w, pool, g = build(11, 43)
    keys = gen.keys_of(s, d)
    reqs = serving.make_stream(n, keys, 40, rate=1.0, seed=8, median=200, max_size=1000,
                               all_at_once=True)
    loop = serving.ServingLoop(g, keep_results=True)
    done = loop.run(reqs, mode="sequential")
    versions = replay_versions(n, s, d, w, done)
    for r in done:
        if r.kind == "update":
            continue
        ref = versions[r.version]
        exp = oracle.bfs(ref, r.src) if r.kind == "bfs" else oracle.dijkstra(ref, r.src)
        assert np.array_equal(r.result, exp), (r.kind, r.src, r.version)


def test_queries_join_running_traversals_and_finish_on_their_own(gpu):
    # queries only, arriving faster than a traversal finishes: later ones are
    # admitted into the traversal already running, and each finishes on its
    # own step (a source without out-edges at once), exact
    n, s, d, w, pool, g = build(13, 44)
    keys = gen.keys_of(s, d)
    reqs = serving.make_stream(n, keys, 120, rate=40000.0, seed=9, update_frac=0.0)
    loop = serving.ServingLoop(g, keep_results=True)
    done = loop.run(reqs, mode="batched")
    ref = oracle.csr(n, s, d, w)
    for r in done:
        exp = oracle.bfs(ref, r.src) if r.kind == "bfs" else oracle.dijkstra(ref, r.src)
        assert np.array_equal(r.result, exp), (r.kind, r.src)
    joined = 0
    for kind in ("bfs", "sssp"):
        rs = [r for r in done if r.kind == kind]
        joined += sum(1 for a in rs for b in rs if a.start < b.start < a.done)
        # within one admission, traversals end at different steps
        by_start = {}
        for r in rs:
            by_start.setdefault(r.start, set()).add(r.done)
        assert any(len(v) > 1 for v in by_start.values())
    assert joined > 0
