"""The bench's headline path and the serving configuration checked against
the oracle at their full sizes:

- RMAT-22 and RMAT-24 (weights [1, 100]) with the in-edge mirror, deferred
  mutation checks and deferred SSSP calls, in steps of 1,000,000 updates
  (500,000 deletions of existing edges, then 500,000 new edges) applied as
  DeleteEdges -> Decremental -> InsertEdges -> Incremental from vertex 0,
  the steps queued back to back and settled only by the reads: distances
  bit-exact against the oracle's from-scratch SSSP on the updated edge set,
  parents valid, every few steps.
- The serving loop on RMAT-24 resident in HBM: a Poisson stream of update
  batches (lognormal sizes) and BFS/SSSP queries served with continuous
  batching; a sample of at least 16 queries spread over the run is checked
  against the oracle's BFS / SSSP on the graph version each ran on (the same
  query alone), bit-exact.

Progress lines are printed as each phase ends (run with -s to see them)."""
import time

import numpy as np
import pytest

import meerkat_amd as mk
import oracle
from hostedges import HostEdges
from meerkat_amd import gen, serving

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[at-size] {msg}", flush=True)


def check_sssp(sp, host, source=0):
    t = time.perf_counter()
    ref_g = host.csr()
    ref = oracle.sssp_par(ref_g, source, oracle.num_threads())
    dist, par = sp.read()
    assert np.array_equal(dist, ref), int((dist != ref).sum())
    assert oracle.check_parents_fast(ref_g, dist, par, source, weights=True) == 0
    say(f"checked {ref_g.m} edges, {int((ref != 0xFFFFFFFF).sum())} reached, "
        f"{time.perf_counter() - t:.1f}s")


@pytest.mark.parametrize("scale,steps,check_every", [(22, 6, 2), (24, 4, 2)])
def test_deferred_mirror_million_edge_steps(gpu, scale, steps, check_every):
    t0 = time.perf_counter()
    n, s, d, w = gen.rmat(scale, 16, seed=9, weighted=True)
    m = len(s)
    pool = mk.SlabPool((16 if scale >= 24 else 4) << 30)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    host = HostEdges(n, s, d, w)
    half = 500_000
    rng = np.random.default_rng(scale)
    del_idx = rng.choice(m, size=steps * half, replace=False)
    ins_s, ins_d, ins_w = gen.insert_batch(host.keys, n, steps * half, seed=scale + 1,
                                           weighted=True, wseed=scale + 2)
    sp = mk.SSSP(g, 0)
    sp.Static()
    g.set_deferred(True)
    sp.set_deferred(True)
    say(f"RMAT-{scale}: n={n} m={m}, built in {time.perf_counter() - t0:.1f}s")
    keep = []  # device batches stay alive until their calls are settled
    for i in range(steps):
        sel = del_idx[i * half:(i + 1) * half]
        ds, dd = mk.device_u32(s[sel]), mk.device_u32(d[sel])
        sl = slice(i * half, (i + 1) * half)
        bs, bd, bw = mk.device_u32(ins_s[sl]), mk.device_u32(ins_d[sl]), mk.device_u32(ins_w[sl])
        keep.append((ds, dd, bs, bd, bw))
        g.DeleteEdges(ds, dd)
        sp.Decremental(ds, dd)
        g.InsertEdges(bs, bd, bw)
        sp.Incremental(bs, bd, bw)
        host.delete(s[sel], d[sel])
        host.insert(ins_s[sl], ins_d[sl], ins_w[sl])
        if (i + 1) % check_every == 0:
            check_sssp(sp, host)  # the read settles the queued calls
            keep.clear()
    stats = sp.deferred_stats()
    say(f"deferred calls {stats}")
    assert stats["settled"] == 2 * steps
    sp.close()
    g.close()
    pool.close()


def test_c5_serving_rmat24_sampled_queries_equal_oracle(gpu):
    t0 = time.perf_counter()
    n, s, d, w = gen.rmat(24, 16, seed=9, weighted=True)
    pool = mk.SlabPool(16 << 30)
    g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True, reverse=True,
                        in_degree_hints=gen.degree_hints(n, d))
    g.InsertEdges(s, d, w)
    host = HostEdges(n, s, d, w)
    say(f"RMAT-24 resident: n={n} m={len(s)} in {time.perf_counter() - t0:.1f}s")
    # deletions drawn from a sample of the initial edges
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(len(s), size=2_000_000, replace=False))
    R = 160
    reqs = serving.make_stream(n, host.keys[pick], R, rate=400.0, seed=21)
    queries = [i for i, r in enumerate(reqs) if r.kind != "update"]
    sample = set(queries[:: max(1, len(queries) // 40)])
    ids = {id(r): i for i, r in enumerate(reqs)}
    loop = serving.ServingLoop(g, keep_results=lambda r: ids[id(r)] in sample)
    t1 = time.perf_counter()
    done = loop.run(reqs, mode="batched")
    summary = serving.summarize(done, time.perf_counter() - t1)
    say(f"served {summary}")
    assert all(r.done is not None for r in done)
    # queries shared launches: continuous batching happened
    starts = [r.start for r in done if r.kind != "update"]
    assert len(set(starts)) < len(starts)
    # the sampled queries of the versions holding most of them
    kept = [r for r in done if r.result is not None]
    by_version = {}
    for r in kept:
        by_version.setdefault(r.version, []).append(r)
    chosen, covered = [], 0
    for v in sorted(by_version, key=lambda v: -len(by_version[v])):
        if covered >= 16 or len(chosen) >= 6:
            break
        chosen.append(v)
        covered += len(by_version[v])
    assert covered >= 16, {v: len(x) for v, x in by_version.items()}
    ups = sorted((r for r in done if r.kind == "update"), key=lambda r: r.version)
    applied, checked, kinds = 0, 0, set()
    for v in sorted(chosen):
        while applied < v:
            r = ups[applied]
            assert r.version == applied + 1
            host.delete(*r.dele)
            host.insert(*r.ins)
            applied += 1
        ref_g = host.csr()
        for r in by_version[v]:
            if r.kind == "bfs":
                exp = oracle.bfs_par(ref_g, r.src, oracle.num_threads())
            else:
                exp = oracle.sssp_par(ref_g, r.src, oracle.num_threads())
            assert np.array_equal(r.result, exp), (r.kind, r.src, r.version)
            checked += 1
            kinds.add(r.kind)
        say(f"version {v}: {len(by_version[v])} queries equal the oracle")
    assert checked >= 16 and kinds == {"bfs", "sssp"}
    loop.close()
    g.close()
    pool.close()
