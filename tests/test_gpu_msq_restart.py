"""A multi-source BFS session whose level log fills up: on a path with 64
sources spread along it, each vertex sits at a different level for every
source before it, so the log needs about 32 entries per vertex against the
8 it holds.  mk_msq_poll then reports MK_ERR_STATE instead of finishing the
slots that lost entries; the slots reported finished before stay exact; the
next session (with a log four times larger) answers the rest exactly; and
the serving loop restarts such a session by itself (bit-exact against the
oracle's BFS)."""
import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from meerkat_amd import _lib
from meerkat_amd.serving import Request, ServingLoop
from test_gpu_sessions import run_session

pytestmark = pytest.mark.gpu

N = 256


def path_graph():
    s = np.arange(N - 1, dtype=np.uint32)
    pool = mk.SlabPool(1 << 22)
    g = mk.DynamicGraph(N, pool, 0.7, None)
    g.InsertEdges(s, s + 1)
    return g, pool, oracle.csr(N, s, s + 1)


def test_full_level_log_is_reported_then_a_new_session_is_exact(gpu):
    g, pool, ref = path_graph()
    srcs = np.arange(0, N, 4, dtype=np.uint32)  # 64 sources
    q = mk.MultiSource(g)
    q.open()
    q.admit(srcs, np.arange(64, dtype=np.uint32))
    finished, err = set(), None
    for _ in range(4 * N):
        q.launch(1)
        try:
            active, _ = q.poll()
        except mk.MeerkatError as e:
            err = e
            break
        finished |= {sl for sl in range(64) if not (active >> sl) & 1}
    assert err is not None and err.code == _lib.MK_ERR_STATE
    assert 0 < len(finished) < 64  # the sources near the end finish first
    for sl in finished:
        assert np.array_equal(q.distances(sl), oracle.bfs(ref, int(srcs[sl]))), sl
    # the rest in a new session: its log (32 entries per vertex) holds them
    rest = [int(srcs[sl]) for sl in range(64) if sl not in finished]
    run_session(q, [[(i, v) for i, v in enumerate(rest)]])
    for i, v in enumerate(rest):
        assert np.array_equal(q.distances(i), oracle.bfs(ref, v)), v
    q.close()
    g.close()
    pool.close()


def test_serving_loop_restarts_a_session_whose_log_fills(gpu):
    g, pool, ref = path_graph()
    srcs = list(range(0, N, 4))
    reqs = [Request("bfs", 0.0, src=v) for v in srcs]
    loop = ServingLoop(g, keep_results=True)
    out = loop.run(reqs, mode="batched")
    assert loop.session_restarts >= 1
    for r in out:
        assert np.array_equal(r.result, oracle.bfs(ref, r.src)), r.src
    loop.close()
    g.close()
    pool.close()
