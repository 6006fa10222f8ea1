"""Serving loop over an HBM-resident dynamic graph.

Requests arrive on a timeline (seconds from the start): update batches
(insertions + deletions) and single-source BFS / SSSP queries.  Two modes:

- ``batched`` (continuous batching): BFS and SSSP queries are admitted
  into free slots (up to 64) of a multi-source traversal of their kind that
  is already running, joining its next step, and each is done on the step
  its own traversal finishes; the two kinds' steps run on two HIP streams
  at once.  Update batches are applied, in arrival order and while they fit
  the slab pool's free memory, once no traversal runs; while one waits, no
  query is admitted, so the running traversals drain first.
- ``sequential``: one request at a time in arrival order, queries by the
  single-source static BFS / SSSP (the one-at-a-time incremental path).

Each query records the graph version (updates applied before it ran) and can
keep its distances, so tests can check that a query's answer in the mix
equals the same query run alone on that version.  No reference counterpart:
the reference has no serving loop.
"""
import collections
import time

import numpy as np

from ._lib import MK_ERR_STATE, MeerkatError
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

    def __init__(self, graph, keep_results=False, slab_margin=1.25, lockstep=False):
        """keep_results: True keeps every query's distances on the host
        (``Request.result``), a callable keeps those of the requests it
        accepts (a sample, on graphs where each answer is large)."""
        self.g = graph
        self.keep_results = keep_results
        self.slab_margin = slab_margin
        # lockstep: both engines take one step, then the loop waits for both;
        # otherwise each is polled when its own step has run
        self.lockstep = lockstep
        self.bfs = MultiSource(graph, sssp=False)
        self.sssp = MultiSource(graph, sssp=True) if graph.weighted else None
        # one-at-a-time mode answers with the single-source engines
        self.bfs1 = BFS(graph, 0)
        self.sssp1 = SSSP(graph, 0) if graph.weighted else None
        self.version = 0
        self.rejected_rounds = 0
        self.session_restarts = 0  # sessions whose running slots started over (last run)
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

        return self._run_batched(reqs, now)

    def _run_batched(self, reqs, now):
        """Continuous batching: each query kind has a multi-source engine
        running a session into which waiting queries are admitted, up to 64
        slots, while it runs (mk_msq_admit between steps); every query is
        done on the step its traversal finishes (mk_msq_poll), not when the
        whole batch does.  Each engine keeps one step in flight on its own
        stream, so BFS and SSSP work overlaps on the device and neither kind
        waits for the other's steps.  An update batch waits until
        no session runs (a session answers on one graph version), and while
        one waits no query is admitted, so the sessions drain; the updates
        waiting then are applied in arrival order while they fit the slab
        pool, and admission resumes."""
        engines = {"bfs": _Session(self.bfs)}
        if self.sssp is not None:
            engines["sssp"] = _Session(self.sssp)
        self.session_restarts = 0
        pending_u = collections.deque()
        waiting = {k: collections.deque() for k in engines}
        i, n_done, total = 0, 0, len(reqs)
        while n_done < total:
            t = now()
            while i < total and reqs[i].arrival <= t:
                r = reqs[i]
                (pending_u if r.kind == "update" else waiting[r.kind]).append(r)
                i += 1
            running = [e for e in engines.values() if e.running]
            blocked = False  # the next update does not fit the pool's free slabs
            if pending_u and not running:
                tu = time.perf_counter()
                while pending_u:
                    r = pending_u[0]
                    if not self._admit(r):
                        self.rejected_rounds += 1
                        if not any(waiting.values()):
                            raise RuntimeError("update batch does not fit the slab pool")
                        blocked = True
                        break
                    pending_u.popleft()
                    r.admitted = now()
                    self._apply(r, now)
                    n_done += 1
                self.phase_s["updates"] += time.perf_counter() - tu
            if not pending_u or blocked:
                for kind, e in engines.items():
                    if waiting[kind] and e.free:
                        e.admit(waiting[kind], self.version, now)
            running = [e for e in engines.values() if e.running]
            if not running:
                if i < total and not pending_u and not any(waiting.values()):
                    while now() < reqs[i].arrival:
                        pass
                continue
            # each engine keeps one step in flight and is polled when that
            # step has run: a kind with short steps (BFS) does not wait for
            # the other kind's long ones
            tq = time.perf_counter()
            if self.lockstep:
                # one step of every engine, then wait for each
                for e in running:
                    e.engine.launch(1)
                    e.inflight = True
                for e in running:
                    n_done += e.poll(now, self._keeps)
            else:
                for e in running:
                    if not e.inflight:
                        e.engine.launch(1)
                        e.inflight = True
                    elif e.engine.ready():
                        n_done += e.poll(now, self._keeps)
            self.phase_s["queries"] += time.perf_counter() - tq
        self.session_restarts = sum(e.total_restarts for e in engines.values())
        return reqs


class _Session:
    """One engine's continuous-batching session: slot bookkeeping."""

    def __init__(self, engine):
        self.engine = engine
        self.slots = {}       # slot -> request still running
        self.used = 0         # slots taken since the session opened
        self.running = False
        self.inflight = False  # a launched step not yet polled
        self.restarts = 0     # sessions restarted in a row without a slot finishing
        self.total_restarts = 0

    @property
    def free(self):
        return MultiSource.MAX_SOURCES - self.used if self.running else MultiSource.MAX_SOURCES

    def admit(self, queue, version, now):
        if not self.running:
            self.engine.open()
            self.running = True
            self.used = 0
        take = [queue.popleft() for _ in range(min(self.free, len(queue)))]
        slots = list(range(self.used, self.used + len(take)))
        self.engine.admit(np.array([r.src for r in take], np.uint32), np.array(slots, np.uint32))
        t = now()
        for s, r in zip(slots, take):
            r.start = t
            r.version = version
            self.slots[s] = r
        self.used += len(take)

    def restart(self):
        """The session cannot finish its running slots (its BFS level log
        filled up: sources admitted over many steps put each vertex in many
        levels): they run again in a new session, from their sources,
        admitted together, keeping their start times.  The engine gives the
        new session a log four times larger, up to what any session needs,
        so at most two restarts in a row find no slot finished."""
        self.restarts += 1
        self.total_restarts += 1
        if self.restarts > 2:
            raise RuntimeError("multi-source session restarted twice without progress")
        run = [self.slots[s] for s in sorted(self.slots)]
        self.engine.open()
        self.slots = {s: r for s, r in enumerate(run)}
        self.used = len(run)
        self.engine.admit(np.array([r.src for r in run], np.uint32),
                          np.arange(len(run), dtype=np.uint32))

    def poll(self, now, keeps):
        """Read the launched step's outcome (it has run); complete the slots
        that finished.  Returns how many did; the session ends when none is
        left running."""
        self.inflight = False
        try:
            active, _ = self.engine.poll()
        except MeerkatError as e:
            if e.code != MK_ERR_STATE:
                raise
            self.restart()
            return 0
        t = now()
        done = [s for s in self.slots if not (active >> s) & 1]
        if done:
            self.restarts = 0
        for s in done:
            r = self.slots.pop(s)
            if keeps(r):
                r.result = self.engine.distances(s)
            r.done = t
        if not self.slots:
            self.running = False
        return len(done)


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
