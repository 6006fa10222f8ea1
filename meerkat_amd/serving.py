"""Serving loop over an HBM-resident dynamic graph.

Requests arrive on a timeline (seconds from the start): update batches
(insertions + deletions) and single-source BFS / SSSP queries.  Two modes:

- ``batched`` (continuous batching): whenever the device is free, every
  update batch that has arrived and fits the slab pool's free memory is
  applied (in arrival order), then up to 64 waiting BFS queries and up to 64
  waiting SSSP queries run as two multi-source traversals on two HIP streams
  at the same time (each ``MultiSource.run`` releases the GIL, so both are in
  flight together).  Requests that arrive while a round runs join the next
  round.
- ``sequential``: one request at a time in arrival order, queries by the
  single-source static BFS / SSSP (the one-at-a-time incremental path).

Each query records the graph version (updates applied before it ran) and can
keep its distances, so tests can check that a query's answer in the mix
equals the same query run alone on that version.  No reference counterpart:
the reference has no serving loop.
"""
import collections
import threading
import time

import numpy as np

from .algorithms import BFS, SSSP, MultiSource
from .graph import device_u32


class Request:
    __slots__ = ("kind", "arrival", "src", "ins", "dele", "ins_dev", "dele_dev", "start", "done",
                 "version", "result", "admitted")

    def __init__(self, kind, arrival, src=None, ins=None, dele=None):
        self.kind = kind          # "update", "bfs" or "sssp"
        self.arrival = arrival    # seconds from the start of the run
        self.src = src            # query source vertex
        self.ins = ins            # (src, dst, w) numpy arrays of an update
        self.dele = dele          # (src, dst) numpy arrays of an update
        self.ins_dev = None       # device copies, staged before the run
        self.dele_dev = None
        self.start = None
        self.done = None
        self.version = None
        self.result = None
        self.admitted = None

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def _stage(req, device):
    """Copy an update batch to the device ahead of time (inputs resident in
    HBM before the run, as a server receiving batches into device buffers)."""
    if req.kind != "update":
        return req
    if req.ins is not None and len(req.ins[0]) and req.ins_dev is None:
        s, d, w = req.ins
        req.ins_dev = (device_u32(s, device), device_u32(d, device),
                       None if w is None else device_u32(w, device))
    if req.dele is not None and len(req.dele[0]) and req.dele_dev is None:
        s, d = req.dele
        req.dele_dev = (device_u32(s, device), device_u32(d, device))
    return req


class ServingLoop:
    MAX_BATCH = MultiSource.MAX_SOURCES

    def __init__(self, graph, keep_results=False, slab_margin=1.25):
        """keep_results: True keeps every query's distances on the host
        (``Request.result``), a callable keeps those of the requests it
        accepts (a sample, on graphs where each answer is large)."""
        self.g = graph
        self.keep_results = keep_results
        self.slab_margin = slab_margin
        self.bfs = MultiSource(graph, sssp=False)
        self.sssp = MultiSource(graph, sssp=True) if graph.weighted else None
        # one-at-a-time mode answers with the single-source engines
        self.bfs1 = BFS(graph, 0)
        self.sssp1 = SSSP(graph, 0) if graph.weighted else None
        self.version = 0
        self.rejected_rounds = 0
        self.phase_s = {"updates": 0.0, "queries": 0.0}  # wall time per phase, last run

    def _keeps(self, r):
        k = self.keep_results
        return k(r) if callable(k) else bool(k)

    def close(self):
        """Release the loop's traversal handles (before the graph)."""
        for h in (self.bfs, self.sssp, self.bfs1, self.sssp1):
            if h is not None:
                h.close()

    # ------------------------------------------------------------ updates
    def _slabs_needed(self, n_ins):
        """Upper estimate of new slabs an insert batch may claim: one per
        31 keys (map slab) per adjacency (forward and mirror), with margin."""
        per = 31 if self.g.weighted else 63
        copies = 2 if self.g.reverse else 1
        return int(np.ceil(n_ins / per * copies * self.slab_margin)) + 64

    def _admit(self, req):
        n_ins = 0 if req.ins is None else len(req.ins[0])
        used, total = self.g.pool.stats()
        return total - used >= self._slabs_needed(n_ins)

    def _apply(self, req, now):
        t = time.perf_counter()
        req.start = now()
        if req.dele_dev is not None:
            self.g.DeleteEdges(*req.dele_dev)
        if req.ins_dev is not None:
            s, d, w = req.ins_dev
            self.g.InsertEdges(s, d, w if self.g.weighted else None)
        self.version += 1
        req.version = self.version
        req.done = now()
        self.phase_s["updates"] += time.perf_counter() - t

    # ------------------------------------------------------------ queries
    def _run_queries(self, engine, reqs, now):
        if not reqs:
            return
        t_start = now()
        for r in reqs:
            r.start = t_start  # one launch for the whole batch
            r.version = self.version
        engine.run(np.array([r.src for r in reqs], np.uint32))
        t = now()
        for slot, r in enumerate(reqs):
            if self._keeps(r):
                r.result = engine.distances(slot)
            r.done = t

    def _run_single(self, r, now):
        eng = self.bfs1 if r.kind == "bfs" else self.sssp1
        r.start = now()
        r.version = self.version
        eng.set_source(r.src)
        eng.Static()
        if self._keeps(r):
            r.result = eng.Distances()
        r.done = now()

    # ------------------------------------------------------------ modes
    def run(self, requests, mode="batched"):
        """Serve `requests` (sorted by arrival) and return them, completed."""
        reqs = [_stage(r, self.g.device) for r in sorted(requests, key=lambda r: r.arrival)]
        self.phase_s = {"updates": 0.0, "queries": 0.0}
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        if mode == "sequential":
            for r in reqs:
                while now() < r.arrival:
                    pass
                if r.kind == "update":
                    if not self._admit(r):
                        raise RuntimeError("update batch does not fit the slab pool")
                    r.admitted = now()
                    self._apply(r, now)
                else:
                    tq = time.perf_counter()
                    self._run_single(r, now)
                    self.phase_s["queries"] += time.perf_counter() - tq
            return reqs

        pending_u = collections.deque()
        pending = {"bfs": collections.deque(), "sssp": collections.deque()}
        i, n_done, total = 0, 0, len(reqs)
        while n_done < total:
            t = now()
            while i < total and reqs[i].arrival <= t:
                (pending_u if reqs[i].kind == "update" else pending[reqs[i].kind]).append(reqs[i])
                i += 1
            if not pending_u and not pending["bfs"] and not pending["sssp"]:
                if i < total:
                    while now() < reqs[i].arrival:
                        pass
                continue
            # updates first, in arrival order, while they fit the pool
            while pending_u:
                r = pending_u[0]
                if not self._admit(r):
                    self.rejected_rounds += 1
                    if not pending["bfs"] and not pending["sssp"]:
                        raise RuntimeError("update batch does not fit the slab pool")
                    break
                pending_u.popleft()
                r.admitted = now()
                self._apply(r, now)
                n_done += 1
            # then one round of queries: BFS and SSSP batches overlap
            qb = [pending["bfs"].popleft() for _ in range(min(self.MAX_BATCH, len(pending["bfs"])))]
            qs = [pending["sssp"].popleft()
                  for _ in range(min(self.MAX_BATCH, len(pending["sssp"])))]
            tq = time.perf_counter()
            if qb and qs:
                th = threading.Thread(target=self._run_queries, args=(self.sssp, qs, now))
                th.start()
                self._run_queries(self.bfs, qb, now)
                th.join()
            else:
                self._run_queries(self.bfs, qb, now)
                self._run_queries(self.sssp, qs, now)
            self.phase_s["queries"] += time.perf_counter() - tq
            n_done += len(qb) + len(qs)
        return reqs


def make_stream(n, keys, count, rate, seed=9, update_frac=0.3, median=10_000, max_size=1_000_000,
                weighted=True, size_scale=1.0, all_at_once=False):
    """A Poisson request stream: `count` requests at `rate` per second;
    `update_frac` of them update batches with lognormal sizes (median
    `median`, clipped at `max_size`, scaled by `size_scale`), half
    insertions of new random edges and half deletions sampled from `keys`
    (the current edge set, uint64 src << 32 | dst, consumed as used); the
    rest BFS or SSSP queries (50/50) from uniform random sources."""
    from . import gen

    rng = np.random.default_rng(seed)
    t = 0.0 if all_at_once else None
    reqs = []
    keys = np.array(keys, dtype=np.uint64, copy=True)
    perm = rng.permutation(len(keys))
    del_cursor = 0
    arrival = 0.0
    for k in range(count):
        if not all_at_once:
            arrival += rng.exponential(1.0 / rate)
        a = t if all_at_once else arrival
        if rng.random() < update_frac:
            size = int(min(max_size, rng.lognormal(np.log(median), 1.0)) * size_scale)
            size = max(size, 2)
            n_ins = size // 2
            n_del = size - n_ins
            ins = gen.insert_batch(None, n, n_ins, seed=seed * 1000003 + k, weighted=weighted,
                                   wseed=seed + 17)
            take = perm[del_cursor:del_cursor + n_del]
            del_cursor += len(take)
            ds, dd = gen.split_keys(keys[take])
            reqs.append(Request("update", a, ins=ins, dele=(ds, dd)))
        else:
            kind = "bfs" if (rng.random() < 0.5 or not weighted) else "sssp"
            reqs.append(Request(kind, a, src=int(rng.integers(0, n))))
    return reqs


def summarize(reqs, elapsed):
    lat = np.array([r.latency for r in reqs if r.kind != "update"]) * 1e3
    ulat = np.array([r.latency for r in reqs if r.kind == "update"]) * 1e3
    out = {
        "requests": len(reqs),
        "elapsed_s": elapsed,
        "requests_per_s": len(reqs) / elapsed if elapsed > 0 else 0.0,
        "queries": int(len(lat)),
        "updates": int(len(ulat)),
    }
    if len(lat):
        out.update(query_p50_ms=float(np.percentile(lat, 50)),
                   query_p99_ms=float(np.percentile(lat, 99)))
    if len(ulat):
        out.update(update_p50_ms=float(np.percentile(ulat, 50)),
                   update_p99_ms=float(np.percentile(ulat, 99)))
    return out
