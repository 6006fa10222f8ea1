"""Continuous batching in the multi-source engine: sources admitted into
free slots of a traversal that is already running (mk_msq_open / admit /
launch / poll).  Every slot equals the oracle's single-source BFS / Dijkstra
on the graph (bit-exact) whatever step it joined at, a slot is reported
finished exactly when its traversal is, and the protocol errors are
reported."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import _lib, gen

pytestmark = pytest.mark.gpu


def run_session(q, waves, max_polls=10_000):
    """Admit waves[i] (a list of (slot, source)) before the i-th poll, one
    step per poll, until every admitted slot is finished.  Returns
    {slot: (poll index at admission, poll index when reported finished)}."""
    q.open()
    seen = {}
    finished = {}
    i = 0
    while True:
        if i < len(waves) and waves[i]:
            sl = np.array([s for s, _ in waves[i]], np.uint32)
            src = np.array([v for _, v in waves[i]], np.uint32)
            q.admit(src, sl)
            for s in sl:
                seen[int(s)] = i
        q.launch(1)
        active, _ = q.poll()
        for s in seen:
            if s not in finished and not (active >> s) & 1:
                finished[s] = i
        i += 1
        if i >= len(waves) and len(finished) == len(seen):
            return {s: (seen[s], finished[s]) for s in seen}
        assert i < max_polls


@pytest.mark.parametrize("sssp", [False, True])
def test_admissions_at_different_steps_are_exact(gpu, sssp):
    n, s, d, w = gen.rmat(13, 16, seed=51, weighted=sssp)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=sssp)
    g.InsertEdges(s, d, w)
    ref = oracle.csr(n, s, d, w)
    deg = ref.degrees()
    isolated = int(np.nonzero(deg == 0)[0][0])
    rng = np.random.default_rng(52)
    pick = rng.choice(np.nonzero(deg)[0], 40, replace=False).astype(int)
    # waves: step 0 (hub, isolated, a source twice), step 2, step 3, step 7,
    # and late ones while early slots have long finished
    waves = [[(0, 0), (1, isolated), (2, pick[0]), (3, pick[0])], [],
             [(10 + j, pick[1 + j]) for j in range(8)],
             [(30, 0), (31, pick[9])], [], [], [],
             [(40 + j, pick[10 + j]) for j in range(20)]]
    waves += [[] for _ in range(12)] + [[(63, pick[30])]]
    q = mk.MultiSource(g, sssp=sssp)
    times = run_session(q, waves)
    srcs = {slot: v for wv in waves for slot, v in wv}
    for slot, v in srcs.items():
        exp = oracle.dijkstra(ref, int(v)) if sssp else oracle.bfs(ref, int(v))
        assert np.array_equal(q.distances(slot), exp), (slot, v)
    # a source without out-edges has nothing to expand: done at its first poll
    assert times[1] == (0, 0)
    # a slot cannot be reported finished before it was admitted
    assert all(f >= a for a, f in times.values())
    # the engine still answers whole runs after a session
    q.run(np.array([pick[5], 0], np.uint32))
    exp = oracle.dijkstra(ref, int(pick[5])) if sssp else oracle.bfs(ref, int(pick[5]))
    assert np.array_equal(q.distances(0), exp)
    q.close()
    g.close()
    pool.close()


def test_session_protocol_errors(gpu):
    n, s, d, _ = gen.rmat(10, 8, seed=53)
    pool = mk.SlabPool(1 << 24)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s))
    g.InsertEdges(s, d)
    q = mk.MultiSource(g)
    with pytest.raises(mk.MeerkatError) as e:
        q.admit([0], [0])  # no session
    assert e.value.code == _lib.MK_ERR_STATE
    q.open()
    q.admit([0, 5], [0, 1])
    for bad_src, bad_slot, code in [([1], [0], _lib.MK_ERR_STATE),  # slot used
                                    ([1], [64], _lib.MK_ERR_BAD_ARG),
                                    ([n], [2], _lib.MK_ERR_RANGE)]:
        with pytest.raises(mk.MeerkatError) as e:
            q.admit(bad_src, bad_slot)
        assert e.value.code == code
    with pytest.raises(mk.MeerkatError) as e:
        q.distances(7)  # never admitted
    assert e.value.code == _lib.MK_ERR_RANGE
    while True:
        q.launch(2)
        active, _ = q.poll()
        if active == 0:
            break
    ref = oracle.csr(n, s, d)
    assert np.array_equal(q.distances(1), oracle.bfs(ref, 5))
    q.close()
    g.close()
    pool.close()
