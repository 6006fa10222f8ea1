"""Captured update steps (mk_*_capture_step): one DeleteEdges ->
Decremental -> InsertEdges -> Incremental step recorded into a hipGraph and
replayed on refilled batch buffers, synchronised after every replay and
queued back to back; distances bit-exact against the oracle after each
sync, parents valid; a failed batch reported by the sync, and the handle
usable again after Static()."""
import numpy as np
import pytest
import torch

import meerkat_amd as mk
import oracle
from meerkat_amd import _lib, gen
from test_gpu_traversal import EdgeSet, check_bfs, check_sssp

pytestmark = pytest.mark.gpu


def batches(es, n, k, size, seed, weighted):
    out = []
    keys = es.keys.copy()
    for i in range(k):
        ds, dd = gen.delete_batch(keys, size, seed=seed + i)
        keys = keys[~np.isin(keys, gen.keys_of(ds, dd))]
        bs, bd, bw = gen.insert_batch(keys, n, size, seed=seed + 100 + i, weighted=weighted,
                                      wseed=seed)
        keys = np.union1d(keys, gen.keys_of(bs, bd))
        out.append((ds, dd, bs, bd, bw))
    return out


@pytest.mark.parametrize("sssp", [True, False])
def test_replays_synced_each_step_are_exact(gpu, sssp):
    n, s, d, w = gen.rmat(14, 16, seed=61, weighted=sssp)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=sssp, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    es = EdgeSet(s, d, w)
    h = mk.SSSP(g, 0) if sssp else mk.BFS(g, 0)
    h.Static()
    bs = batches(es, n, 5, 3000, 62, sssp)
    ds, dd, is_, id_, iw = bs[0]
    step = h.capture_step(ds, dd, is_, id_, iw if sssp else None)
    check = check_sssp if sssp else check_bfs
    check(h, es, n, 0)  # capture runs nothing
    for i, (ds, dd, is_, id_, iw) in enumerate(bs):
        if i:
            buf = step.buffers
            for dst, src in zip(buf, (ds, dd, is_, id_, iw)):
                if dst is not None:
                    dst.copy_(torch.from_numpy(np.ascontiguousarray(src).view(np.int32)))
            torch.cuda.synchronize()
        step.launch()
        step.sync()
        es.delete(ds, dd)
        es.insert(is_, id_, iw if sssp else None)
        check(h, es, n, 0)
    # the API calls still work between replays
    ds, dd = gen.delete_batch(es.keys, 500, seed=70)
    g.DeleteEdges(ds, dd)
    h.Decremental(ds, dd)
    es.delete(ds, dd)
    check(h, es, n, 0)
    step.close()
    h.close()
    g.close()
    pool.close()


def test_replays_back_to_back_with_loaded_batches(gpu):
    n, s, d, w = gen.rmat(14, 16, seed=63, weighted=True)
    pool = mk.SlabPool(1 << 28)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    es = EdgeSet(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    bs = batches(es, n, 6, 2000, 64, True)
    step = sp.capture_step(*bs[0])
    dev = [tuple(mk.device_u32(x) for x in b) for b in bs]
    torch.cuda.synchronize()
    for b in dev[:3]:  # three replays queued without a host wait
        step.load(*b)
        step.launch()
    step.sync()
    for ds, dd, is_, id_, iw in bs[:3]:
        es.delete(ds, dd)
        es.insert(is_, id_, iw)
    check_sssp(sp, es, n, 0)
    for b in dev[3:]:
        step.load(*b)
        step.launch()
    step.sync()
    for ds, dd, is_, id_, iw in bs[3:]:
        es.delete(ds, dd)
        es.insert(is_, id_, iw)
    check_sssp(sp, es, n, 0)
    step.close()
    sp.close()
    g.close()
    pool.close()


def test_replay_with_a_bad_batch_is_reported(gpu):
    n, s, d, w = gen.rmat(12, 16, seed=65, weighted=True)
    pool = mk.SlabPool(1 << 27)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
    g.InsertEdges(s, d, w)
    es = EdgeSet(s, d, w)
    sp = mk.SSSP(g, 0)
    sp.Static()
    bs = batches(es, n, 2, 500, 66, True)
    step = sp.capture_step(*bs[0])
    ds, dd, is_, id_, iw = bs[1]
    bad = is_.copy()
    bad[7] = n + 5  # a vertex id out of range
    step.load(*(mk.device_u32(x) for x in (ds, dd, bad, id_, iw)))
    step.launch()
    with pytest.raises(mk.MeerkatError) as e:
        step.sync()
    assert e.value.code == _lib.MK_ERR_RANGE
    # the delete applied, the insert changed nothing; Static() makes it exact
    es.delete(ds, dd)
    sp.Static()
    check_sssp(sp, es, n, 0)
    with pytest.raises(mk.MeerkatError) as e:
        sp.close()  # the step still uses the handle
    assert e.value.code == _lib.MK_ERR_STATE
    step.close()
    sp.close()
    g.close()
    pool.close()
