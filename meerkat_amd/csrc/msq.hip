// Multi-source BFS / SSSP: up to 64 queries answered by one traversal.
//
// No reference counterpart (the reference runs one source per BFS/SSSP
// object, include/Algorithms/BFS.cuh:34).  Source slot s (0..63) is bit s of
// 64-bit per-vertex masks:
//  - BFS: seen[v] holds the sources that reached v; F[L][v] the sources for
//    which v is at level L.  A level step expands every frontier vertex u
//    once for all its sources: per edge (u, v) new = F[L][u] & ~seen[v] is
//    claimed with one 64-bit atomicOr, so 64 BFS advance together.  Every
//    frontier entry (v, F[L][v]) stays in an append-only per-level log, from
//    which the distances of one source are materialised on demand (each
//    (v, s) bit is set at exactly one level).
//  - SSSP: dist is stored source-minor (dist[v * 64 + s]) so the 64 lanes of
//    a wave relax one edge for all sources with one coalesced 256-byte
//    access; F[i][v] holds the sources whose distance to v improved in
//    step i (label-correcting, as the single-source SSSP).
// Frontier bookkeeping follows the single-source traversal: counters in
// three rotating slots read on the device, hub vertices (more than
// kMsSmall buckets) listed one entry per kHubChunk buckets and spread over
// all waves, and the
// host reading counters only every few steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "profile.hpp"
#include "slabhash.hpp"

namespace mk {

constexpr uint32_t kMsSmall = 2;

// BFS rotates three slots (its queue is the append-only log); SSSP rotates
// four, because a step clears the masks of the previous queue (slot i-1)
// and the counters of the one before (slot i-2).
constexpr uint32_t kMsSlots = 4;

struct MsCounts {
  uint32_t cnt[kMsSlots];          // small entries per slot
  uint32_t steps;                  // non-empty steps
  uint32_t pad0;
  unsigned long long large[kMsSlots];  // (hub-list entries << 32) | buckets per slot
  uint32_t overflow;               // log capacity exceeded
  uint32_t pad;
  unsigned long long edges;        // adjacency entries read
  unsigned long long slabs;        // slabs read
  unsigned long long pairs;        // SSSP: (entry, active source) relaxations examined
  // per slot: the sources with an entry in that slot's queue (the OR of the
  // bits the step writing it set, and of admissions into it); a source
  // missing from the queue of the next step has converged
  unsigned long long active[kMsSlots];
};

struct MsView {
  uint64_t *seen;      // BFS
  uint64_t *F[kMsSlots];  // per-step masks
  uint32_t *dist;      // SSSP, n * 64
  uint32_t *logv;      // BFS log / SSSP queues (kMsSlots * n)
  uint64_t *logm;      // BFS log masks
  uint32_t *seg;       // BFS: first log index of each level
  uint32_t log_cap;
  uint2 *ql[kMsSlots];  // hub lists
  MsCounts *c;
  uint32_t *off;       // BFS: level at which each source slot was admitted
  uint32_t n;
};

MK_DEV uint64_t wave_or64(uint64_t x) {
  for (int o = 32; o >= 1; o >>= 1) {
    const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(x)), o, 64));
    const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(x >> 32)), o, 64));
    x |= (uint64_t(hi) << 32) | lo;
  }
  return x;
}

MK_DEV uint64_t read_lane64(uint64_t x, uint32_t l) {
  return (uint64_t(read_lane(uint32_t(x >> 32), l)) << 32) | read_lane(uint32_t(x), l);
}

constexpr uint32_t kMsWaveQ = 512;  // LDS staging entries per wave

// Per-wave LDS staging of list appends: one global atomicAdd per flush.
struct MsQ {
  uint32_t *buf;
  uint32_t n;  // wave-uniform
};

MK_DEV void ms_flush(const MsView &S, MsQ &q, uint32_t *list, uint32_t base, uint32_t *cnt,
                     uint32_t cap) {
  if (q.n == 0) return;
  __builtin_amdgcn_wave_barrier();
  uint32_t pos = 0;
  if (lane_id() == 0) pos = atomicAdd(cnt, q.n);
  pos = read_lane(pos, 0);
  for (uint32_t i = lane_id(); i < q.n; i += 64) {
    if (base + pos + i < cap) list[base + pos + i] = q.buf[i];
    else atomicOr(&S.c->overflow, 1u);
  }
  __builtin_amdgcn_wave_barrier();
  q.n = 0;
}

// Append v (lanes with `push`) to the list (through the wave's staging) and
// hub vertices also to the hub list `ql` / `large`; `log_all` keeps hubs in
// the list too.  Entries past `cap` set the overflow flag instead of being
// written.  Wave-collective.
MK_DEV void ms_push(const GraphView &G, const MsView &S, MsQ &q, bool push, uint32_t v,
                    uint32_t *list, uint32_t base, uint32_t *cnt, uint2 *ql,
                    unsigned long long *large, uint32_t cap, bool log_all) {
  const uint64_t pm = ballot(push);
  if (pm == 0) return;
  uint32_t nb = 0;
  if (push) nb = G.bucket_off[v + 1] - G.bucket_off[v];
  const bool hub = push && nb > kMsSmall;
  const bool to_list = push && (log_all || !hub);
  const uint64_t lm = ballot(to_list);
  const uint32_t c = uint32_t(__popcll(lm));
  if (q.n + c > kMsWaveQ) ms_flush(S, q, list, base, cnt, cap);
  if (to_list) q.buf[q.n + lanes_below(lm)] = v;
  q.n += c;
  push_hubs(G, ql, large, hub, v, nb);
}

// ------------------------------------------------------------------ BFS
// BFS work on one slab of u (mask mu): each neighbour takes the source bits
// of mu it has not seen before this level; a vertex entering level L+1 is
// pushed once (first bits).  The seen update does not return: two waves
// that both take bit s of x in this level both put it into the same level's
// mask, so each (vertex, source) bit still lands in exactly one level.
MK_DEV void msbfs_slab(const GraphView &G, const MsView &S, MsQ &q, uint64_t key_mask_all,
                       uint32_t x, uint64_t mu, uint64_t *Fn, uint32_t nbase, uint32_t nslot,
                       uint32_t &edges, uint64_t &acc) {
  const uint32_t lane = lane_id();
  const bool live = ((key_mask_all >> lane) & 1ull) && x < kTomb;
  edges += uint32_t(__popcll(ballot(live)));
  bool push = false;
  if (live) {
    const uint64_t nw = mu & ~S.seen[x];
    if (nw) {
      atomicOr((unsigned long long *)(S.seen + x), (unsigned long long)nw);
      acc |= nw;
      push = atomicOr((unsigned long long *)(Fn + x), (unsigned long long)nw) == 0;
    }
  }
  ms_push(G, S, q, push, x, S.logv, nbase, &S.c->cnt[nslot], S.ql[nslot], &S.c->large[nslot],
          S.log_cap, true);
}

template <bool MAP>
MK_DEV void msbfs_chain(const GraphView &G, const MsView &S, MsQ &q, uint32_t x, uint64_t mu,
                        uint64_t *Fn, uint32_t nbase, uint32_t nslot, uint32_t &edges,
                        uint32_t &slabs, uint64_t &acc) {
  for (;;) {
    ++slabs;
    msbfs_slab(G, S, q, key_mask<MAP>(), x, mu, Fn, nbase, nslot, edges, acc);
    const uint32_t nx = read_lane(x, kNextLane);
    if (nx == kEmpty) break;
    x = slab_ptr(G.pool, nx)[lane_id()];
  }
}

// Level L: log segment [seg[L], seg[L] + cnt[L%3]) plus the hub list of
// slot L%3.  Appends to the log after it and to the hub list of slot
// (L+1)%3; clears slot (L+2)%3 and the masks of level L-1.
template <bool MAP>
__global__ void __launch_bounds__(kBlock) k_msbfs_level(GraphView G, MsView S, uint32_t L) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t slot = L % 3, nslot = (L + 1) % 3, pslot = (L + 2) % 3;
  const uint32_t start = S.seg[L];
  const uint32_t count = S.c->cnt[slot];
  const uint32_t nbase = start + count;
  const unsigned long long lg = S.c->large[slot];
  const uint32_t n_large = uint32_t(lg >> 32);  // hub-list entries
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.seg[L + 1] = nbase;
    S.c->cnt[pslot] = 0;
    S.c->large[pslot] = 0ull;
    S.c->active[pslot] = 0ull;
    if (count) atomicAdd(&S.c->steps, 1u);
  }
  uint64_t *Fc = S.F[slot], *Fn = S.F[nslot], *Fp = S.F[pslot];
  // masks of level L-1 are no longer read: clear them for level L+2 (also
  // when level L is empty: a source admitted later makes level L+1 run, and
  // level L+2's masks must then start clear)
  if (L > 0) {
    const uint32_t ps = S.seg[L - 1];
    for (uint32_t i = ps + gw * 64 + lane; i < min(start, S.log_cap); i += W * 64)
      Fp[S.logv[i]] = 0ull;
  }
  if (count == 0) return;
  // after a log overflow (the run is repeated) only stored entries are read
  const uint32_t end = min(nbase, S.log_cap);
  __shared__ uint32_t lqbuf[kBlock / 64][kMsWaveQ];
  MsQ q{lqbuf[threadIdx.x >> 6], 0u};
  uint32_t edges = 0, slabs = 0;
  uint64_t acc = 0;  // source bits this lane set in the next level's masks
  // hubs: their list entries
  const uint2 *ql = S.ql[slot];
  // lane l takes bucket l % kHubChunk of hub-list entry l / kHubChunk
  const uint32_t eper = min(per_wave(n_large, W), 64u / kHubChunk);
  for (uint32_t e0 = gw * eper; e0 < n_large; e0 += W * eper) {
    const uint32_t el = lane / kHubChunk, ei = e0 + el;
    bool act = el < eper && ei < n_large;
    uint32_t b = 0;
    uint64_t mu = 0;
    if (act) {
      const uint2 ent = ql[ei];
      b = ent.y + lane % kHubChunk;
      act = b < G.bucket_off[ent.x + 1];
      if (act) mu = Fc[ent.x];
    }
    uint64_t m = ballot(act);
    while (m) {
      uint32_t js[4];
      uint32_t c = 0;
#pragma unroll
      for (uint32_t qq = 0; qq < 4; ++qq) {
        js[qq] = m ? ffs64(m) : 0u;
        if (m) {
          m &= m - 1;
          c = qq + 1;
        }
      }
      uint32_t x[4];
#pragma unroll
      for (uint32_t qq = 0; qq < 4; ++qq)
        x[qq] = qq < c ? G.heads[size_t(read_lane(b, js[qq])) * kSlabWords + lane] : kEmpty;
#pragma unroll
      for (uint32_t qq = 0; qq < 4; ++qq)
        if (qq < c)
          msbfs_chain<MAP>(G, S, q, x[qq], read_lane64(mu, js[qq]), Fn, nbase, nslot, edges, slabs,
                           acc);
    }
  }
  // log entries: record the mask; expand non-hub vertices
  const uint32_t per = per_wave(end - start, W);
  for (uint32_t i0 = start + gw * per; i0 < end; i0 += W * per) {
    const uint32_t i = i0 + lane;
    bool act = lane < per && i < end;
    uint32_t u = 0, b0 = 0, nb = 0;
    uint64_t mu = 0;
    if (act) {
      u = S.logv[i];
      mu = Fc[u];
      S.logm[i] = mu;
      b0 = G.bucket_off[u];
      nb = G.bucket_off[u + 1] - b0;
      if (nb > kMsSmall) act = false;
    }
    for (uint32_t k = 0; ballot(act && nb > k); ++k) {
      uint64_t m = ballot(act && nb > k);
      while (m) {
        uint32_t js[4];
        uint32_t c = 0;
#pragma unroll
        for (uint32_t qq = 0; qq < 4; ++qq) {
          js[qq] = m ? ffs64(m) : 0u;
          if (m) {
            m &= m - 1;
            c = qq + 1;
          }
        }
        uint32_t x[4];
#pragma unroll
        for (uint32_t qq = 0; qq < 4; ++qq)
          x[qq] = qq < c ? G.heads[size_t(read_lane(b0, js[qq]) + k) * kSlabWords + lane] : kEmpty;
#pragma unroll
        for (uint32_t qq = 0; qq < 4; ++qq)
          if (qq < c)
            msbfs_chain<MAP>(G, S, q, x[qq], read_lane64(mu, js[qq]), Fn, nbase, nslot, edges,
                             slabs, acc);
      }
    }
  }
  ms_flush(S, q, S.logv, nbase, &S.c->cnt[nslot], S.log_cap);
  acc = wave_or64(acc);
  if (lane == 0 && acc) atomicOr(&S.c->active[nslot], (unsigned long long)acc);
  if (lane == 0 && edges) atomicAdd(&S.c->edges, (unsigned long long)edges);
  if (lane == 0 && slabs) atomicAdd(&S.c->slabs, (unsigned long long)slabs);
}

__global__ void k_msbfs_init(MsView S, const uint32_t *sources, uint32_t k) {
  // one thread: sources may repeat (same vertex in two slots)
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint32_t cntv = 0;
  for (uint32_t s = 0; s < k; ++s) {
    const uint32_t v = sources[s];
    const uint64_t bit = 1ull << s;
    if (S.F[0][v] == 0) S.logv[cntv++] = v;
    S.F[0][v] |= bit;
    S.seen[v] |= bit;
  }
  S.seg[0] = 0;
  S.c->cnt[0] = cntv;
}

// Hub vertices among the sources go to the hub list of slot 0 (one wave).
__global__ void k_ms_source_hubs(GraphView G, MsView S) {
  const uint32_t i = threadIdx.x;
  uint32_t v = 0, nb = 0;
  if (i < S.c->cnt[0]) {
    v = S.logv[i];
    nb = G.bucket_off[v + 1] - G.bucket_off[v];
  }
  push_hubs(G, S.ql[0], &S.c->large[0], nb > kMsSmall, v, nb);
}

// Distances of slot s from the log: the level whose mask holds bit s.
__global__ void k_msbfs_read(MsView S, uint32_t s, uint32_t total, uint32_t levels,
                             uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  if (!((S.logm[i] >> s) & 1ull)) return;
  uint32_t lo = 0, hi = levels;  // last level with seg <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (S.seg[mid] <= i) lo = mid;
    else hi = mid;
  }
  out[S.logv[i]] = lo - S.off[s];
}

// ------------------------------------------------------------------ SSSP
// Relax one slab of u for all sources in mu: per live key (neighbour v,
// weight w), lanes are sources; four neighbours' distance rows are loaded at
// a time.
// A frontier vertex with at most this many active sources relaxes its slabs
// key-parallel (lanes = neighbours, one pass per source) instead of
// row-parallel (lanes = sources, neighbours a few at a time).
constexpr uint32_t kMsSparse = 4;

template <bool MAP, uint32_t U>
MK_DEV void mssp_slab(const GraphView &G, const MsView &S, MsQ &q, uint32_t x, uint32_t du_lane,
                      uint64_t mu, uint64_t *Fn, uint32_t *qn, uint32_t nslot, uint32_t &edges,
                      uint64_t &pairs, uint64_t &acc) {
  const uint32_t lane = lane_id();
  const uint64_t km = key_mask<MAP>();
  uint64_t live = ballot(x < kTomb) & km;
  edges += uint32_t(__popcll(live));
  pairs += uint64_t(__popcll(live)) * uint64_t(__popcll(mu));
  if (uint32_t(__popcll(mu)) <= kMsSparse) {
    // key-parallel: lane j relaxes key j of this slab for each active source
    const bool lv = (live >> lane) & 1ull;
    const uint32_t wv = MAP ? __shfl_down(x, 1, 64) : 1u;
    for (uint64_t m = mu; m; m &= m - 1) {
      const uint32_t src = ffs64(m);
      const uint32_t du = read_lane(du_lane, src);
      if (du == kInf) continue;
      bool push = false, imp = false;
      if (lv) {
        const uint32_t cand = du + wv;
        uint32_t *dp = S.dist + size_t(x) * 64 + src;
        if (cand >= du && cand < *dp && cand < atomicMin(dp, cand)) {
          imp = true;
          push = atomicOr((unsigned long long *)(Fn + x), 1ull << src) == 0;
        }
      }
      if (ballot(imp)) acc |= 1ull << src;
      ms_push(G, S, q, push, x, qn, 0, &S.c->cnt[nslot], S.ql[nslot], &S.c->large[nslot], S.n,
              false);
    }
    return;
  }
  const bool mine = (mu >> lane) & 1ull;
  while (live) {
    uint32_t ks[U];
    uint32_t c = 0;
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq) {
      ks[qq] = live ? ffs64(live) : 0u;
      if (live) {
        live &= live - 1;
        c = qq + 1;
      }
    }
    // only the lanes of sources active at u read their distance word, so a
    // sparse mask touches one line of the 256-byte row, not both
    const bool reads = mine && du_lane != kInf;
    uint32_t v[U], w[U], cur[U];
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq) {
      v[qq] = read_lane(x, ks[qq]);
      w[qq] = MAP ? read_lane(x, ks[qq] + 1) : 1u;
      cur[qq] = (qq < c && reads) ? S.dist[size_t(v[qq]) * 64 + lane] : 0u;
    }
    // the U neighbours' atomics are issued back to back (their round trips
    // overlap), then the mask updates, then one push for all of them
    uint32_t cand[U], old[U];
    bool imp[U];
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq) {
      cand[qq] = du_lane + w[qq];
      imp[qq] = qq < c && reads && cand[qq] >= du_lane && cand[qq] < cur[qq];
    }
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq)
      old[qq] = imp[qq] ? atomicMin(S.dist + size_t(v[qq]) * 64 + lane, cand[qq]) : 0u;
    uint64_t im[U];
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq) {
      im[qq] = ballot(imp[qq] && cand[qq] < old[qq]);
      acc |= im[qq];
    }
    unsigned long long po[U];
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq)
      po[qq] = (im[qq] && lane == 0)
                   ? atomicOr((unsigned long long *)(Fn + v[qq]), (unsigned long long)im[qq])
                   : 1ull;
    // neighbour qq is pushed (by lane qq) when its mask was empty before
    uint32_t pm = 0;
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq)
      if (im[qq] && read_lane(uint32_t(po[qq] == 0), 0)) pm |= 1u << qq;
    uint32_t vv = 0;
#pragma unroll
    for (uint32_t qq = 0; qq < U; ++qq)
      if (lane == qq) vv = v[qq];
    ms_push(G, S, q, lane < U && ((pm >> lane) & 1u), vv, qn, 0, &S.c->cnt[nslot], S.ql[nslot],
            &S.c->large[nslot], S.n, false);
  }
}

template <bool MAP, uint32_t U>
MK_DEV void mssp_chain(const GraphView &G, const MsView &S, MsQ &q, uint32_t x, uint32_t u,
                       uint64_t mu, uint64_t *Fn, uint32_t *qn, uint32_t nslot, uint32_t &edges,
                       uint32_t &slabs, uint64_t &pairs, uint64_t &acc) {
  const uint32_t du = S.dist[size_t(u) * 64 + lane_id()];
  for (;;) {
    ++slabs;
    mssp_slab<MAP, U>(G, S, q, x, du, mu, Fn, qn, nslot, edges, pairs, acc);
    const uint32_t nx = read_lane(x, kNextLane);
    if (nx == kEmpty) break;
    x = slab_ptr(G.pool, nx)[lane_id()];
  }
}

// SSSP step i: queue slot i%4 (vertices, masks F[i%4]); appends to slot
// (i+1)%4; clears the masks of queue i-1 (slot (i+3)%4) and the counters of
// slot (i+2)%4 (queue i-2, whose masks step i-1 cleared).
template <bool MAP, uint32_t U>
__global__ void __launch_bounds__(kBlock) k_mssp_step(GraphView G, MsView S, uint32_t i) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t slot = i % kMsSlots, nslot = (i + 1) % kMsSlots;
  const uint32_t zslot = (i + 2) % kMsSlots, pslot = (i + 3) % kMsSlots;
  const uint32_t count = S.c->cnt[slot];
  const unsigned long long lg = S.c->large[slot];
  const uint32_t n_large = uint32_t(lg >> 32);  // hub-list entries
  uint32_t *qc = S.logv + size_t(slot) * S.n, *qn = S.logv + size_t(nslot) * S.n;
  const uint32_t *qp = S.logv + size_t(pslot) * S.n;
  uint64_t *Fc = S.F[slot], *Fn = S.F[nslot], *Fp = S.F[pslot];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.c->cnt[zslot] = 0;
    S.c->large[zslot] = 0ull;
    S.c->active[zslot] = 0ull;
  }
  // queue i-1 is done: clear its masks before its slot is reused
  const uint32_t pcount = i > 0 ? S.c->cnt[pslot] : 0u;
  const uint32_t plarge = i > 0 ? uint32_t(S.c->large[pslot] >> 32) : 0u;
  for (uint32_t k = gw * 64 + lane; k < pcount; k += W * 64) Fp[qp[k]] = 0ull;
  for (uint32_t k = gw * 64 + lane; k < plarge; k += W * 64) Fp[S.ql[pslot][k].x] = 0ull;
  if (count + n_large == 0) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&S.c->steps, 1u);
  __shared__ uint32_t lqbuf[kBlock / 64][kMsWaveQ];
  MsQ q{lqbuf[threadIdx.x >> 6], 0u};
  uint32_t edges = 0, slabs = 0;
  uint64_t pairs = 0, acc = 0;  // acc: source bits improved into the next queue
  const uint2 *ql = S.ql[slot];
  // lane l takes bucket l % kHubChunk of hub-list entry l / kHubChunk
  const uint32_t eper = min(per_wave(n_large, W), 64u / kHubChunk);
  for (uint32_t e0 = gw * eper; e0 < n_large; e0 += W * eper) {
    const uint32_t el = lane / kHubChunk, ei = e0 + el;
    bool act = el < eper && ei < n_large;
    uint32_t b = 0, u = 0;
    uint64_t mu = 0;
    if (act) {
      const uint2 ent = ql[ei];
      u = ent.x;
      b = ent.y + lane % kHubChunk;
      act = b < G.bucket_off[u + 1];
      if (act) mu = Fc[u];
    }
    uint64_t m = ballot(act);
    while (m) {
      uint32_t js[4];
      uint32_t c = 0;
#pragma unroll
      for (uint32_t qq = 0; qq < 4; ++qq) {
        js[qq] = m ? ffs64(m) : 0u;
        if (m) {
          m &= m - 1;
          c = qq + 1;
        }
      }
      uint32_t x[4];
#pragma unroll
      for (uint32_t qq = 0; qq < 4; ++qq)
        x[qq] = qq < c ? G.heads[size_t(read_lane(b, js[qq])) * kSlabWords + lane] : kEmpty;
      // one copy of the expansion for the four heads (their loads above are
      // already in flight): fewer registers than four inlined copies
#pragma unroll 1
      for (uint32_t qq = 0; qq < c; ++qq) {
        const uint32_t xq = qq == 0 ? x[0] : qq == 1 ? x[1] : qq == 2 ? x[2] : x[3];
        const uint32_t jq = qq == 0 ? js[0] : qq == 1 ? js[1] : qq == 2 ? js[2] : js[3];
        mssp_chain<MAP, U>(G, S, q, xq, read_lane(u, jq), read_lane64(mu, jq), Fn, qn, nslot, edges,
                           slabs, pairs, acc);
      }
    }
  }
  const uint32_t per = per_wave(count, W);
  for (uint32_t k0 = gw * per; k0 < count; k0 += W * per) {
    const uint32_t k = k0 + lane;
    bool act = lane < per && k < count;
    uint32_t u = 0, b0 = 0, nb = 0;
    uint64_t mu = 0;
    if (act) {
      u = qc[k];
      mu = Fc[u];
      b0 = G.bucket_off[u];
      nb = G.bucket_off[u + 1] - b0;
      if (nb > kMsSmall) act = false;  // hubs are expanded from the hub list
    }
    for (uint32_t kk = 0; ballot(act && nb > kk); ++kk) {
      uint64_t m = ballot(act && nb > kk);
      while (m) {
        const uint32_t j = ffs64(m);
        m &= m - 1;
        const uint32_t x = G.heads[size_t(read_lane(b0, j) + kk) * kSlabWords + lane];
        mssp_chain<MAP, U>(G, S, q, x, read_lane(u, j), read_lane64(mu, j), Fn, qn, nslot, edges,
                           slabs, pairs, acc);
      }
    }
  }
  ms_flush(S, q, qn, 0, &S.c->cnt[nslot], S.n);
  acc = wave_or64(acc);  // wave-uniform already; the reduction keeps it exact
  if (lane == 0 && acc) atomicOr(&S.c->active[nslot], (unsigned long long)acc);
  if (lane == 0 && edges) atomicAdd(&S.c->edges, (unsigned long long)edges);
  if (lane == 0 && slabs) atomicAdd(&S.c->slabs, (unsigned long long)slabs);
  if (lane == 0 && pairs) atomicAdd(&S.c->pairs, (unsigned long long)pairs);
}

__global__ void k_mssp_init(MsView S, const uint32_t *sources, uint32_t k) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint32_t cntv = 0;
  for (uint32_t s = 0; s < k; ++s) {
    const uint32_t v = sources[s];
    if (S.F[0][v] == 0) S.logv[cntv++] = v;
    S.F[0][v] |= 1ull << s;
    S.dist[size_t(v) * 64 + s] = 0;
  }
  S.c->cnt[0] = cntv;
}

// Sources admitted into free slots of a running traversal (continuous
// batching), before step `step`: the sources join that step's frontier
// (queue slot `qs`: step % kMsSlots for SSSP, step % 3 for BFS) and their
// bits mark the slot active.  SSSP: distance 0 in the source's lane.  BFS:
// the source's bit in seen; its level 0 is `step`, recorded in off[slot].
// Passed by value: no copy to stage.
struct MsAdmit {
  uint32_t v[64];
  uint32_t slot[64];
  uint32_t k;
};

// One wave: lane 0 appends in order (a vertex named twice, or already in
// the frontier, is queued once), then the new hub vertices go to the hub
// list together.
__global__ void __launch_bounds__(64) k_ms_admit(GraphView G, MsView S, MsAdmit a, uint32_t step,
                                                 uint32_t sssp) {
  __shared__ uint32_t fresh[64];
  const uint32_t t = threadIdx.x;
  const uint32_t qs = sssp ? step % kMsSlots : step % 3;
  if (t == 0) {
    uint32_t cnt = S.c->cnt[qs];
    const uint32_t base = sssp ? qs * S.n : S.seg[step];
    uint64_t act = 0;
    for (uint32_t i = 0; i < a.k; ++i) {
      const uint32_t v = a.v[i], s = a.slot[i];
      const uint64_t bit = 1ull << s;
      fresh[i] = 0u;
      if (S.F[qs][v] == 0) {
        if (sssp || base + cnt < S.log_cap) S.logv[base + cnt] = v;
        else S.c->overflow = 1u;
        ++cnt;
        fresh[i] = 1u;
      }
      S.F[qs][v] |= bit;
      if (sssp) {
        S.dist[size_t(v) * 64 + s] = 0;
      } else {
        S.seen[v] |= bit;
        S.off[s] = step;
      }
      act |= bit;
    }
    S.c->cnt[qs] = cnt;
    S.c->active[qs] |= act;
  }
  __syncthreads();
  uint32_t v = 0, nb = 0;
  bool hub = false;
  if (t < a.k && fresh[t]) {
    v = a.v[t];
    nb = G.bucket_off[v + 1] - G.bucket_off[v];
    hub = nb > kMsSmall;
  }
  push_hubs(G, S.ql[qs], &S.c->large[qs], hub, v, nb);
}

__global__ void k_mssp_read(const uint32_t *dist, uint32_t n, uint32_t s, uint32_t *out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) out[v] = dist[size_t(v) * 64 + s];
}

// ------------------------------------------------------------------ host
struct MultiSource {
  mk_graph *g = nullptr;
  bool sssp = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t *seen = nullptr;
  uint64_t *F[kMsSlots] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *dist = nullptr;
  uint32_t *logv = nullptr;
  uint64_t *logm = nullptr;
  uint32_t *seg = nullptr;
  uint32_t log_cap = 0;
  uint2 *ql[kMsSlots] = {nullptr, nullptr, nullptr, nullptr};
  MsCounts *c = nullptr;
  uint32_t *src_dev = nullptr;
  uint32_t *off = nullptr;  // BFS: admission level per slot
  HostBuf host;
  // continuous batching (open / admit / launch / poll): the session's next
  // step, the slots admitted since open, and whether a session is open
  bool session = false;
  bool grow_log = false;  // a session's log filled up: the next one gets a larger log
  uint32_t step = 0;
  uint64_t admitted = 0;
  uint32_t k = 0;
  uint32_t last_steps = 0;
  uint64_t last_edges = 0;
  uint32_t log_total = 0;
  int max_waves = 0;
  int unroll = 4;

  MsView view() const {
    MsView v;
    v.seen = seen;
    for (uint32_t i = 0; i < kMsSlots; ++i) {
      v.F[i] = F[i];
      v.ql[i] = ql[i];
    }
    v.dist = dist;
    v.logv = logv;
    v.logm = logm;
    v.seg = seg;
    v.log_cap = log_cap;
    v.c = c;
    v.off = off;
    v.n = g->n;
    return v;
  }

  // BFS level log of `entries` (vertex, mask) entries; no run or session
  // needs more than 64 per vertex (each source bit of a vertex is set at one
  // level) plus the admissions
  int alloc_log(uint64_t entries) {
    (void)hipFree(logv);
    (void)hipFree(logm);
    logv = nullptr;
    logm = nullptr;
    entries = std::min<uint64_t>(entries, 64 * uint64_t(g->n) + 64);
    entries = std::min<uint64_t>(entries, 0xFFFFFFF0ull);
    MK_HIP(hipMalloc(&logv, entries * 4));
    MK_HIP(hipMalloc(&logm, entries * 8));
    log_cap = uint32_t(entries);
    return MK_OK;
  }

  int init(mk_graph *graph, bool is_sssp) {
    g = graph;
    sssp = is_sssp;
    const size_t n = g->n;
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipStreamCreate(&stream));
    loop_stream_add(g->device, stream, 1);
    MK_HIP(hipEventCreate(&ev0));
    MK_HIP(hipEventCreate(&ev1));
    for (uint32_t i = 0; i < kMsSlots; ++i) {
      MK_HIP(hipMalloc(&F[i], n * 8));
      // one entry per kHubChunk buckets of each listed hub
      MK_HIP(hipMalloc(&ql[i], (n + size_t(g->total_buckets) / kHubChunk + 64) * 8));
      fill_u64(F[i], 0ull, n, stream);
    }
    MK_HIP(hipMalloc(&c, sizeof(MsCounts)));
    MK_HIP(hipMalloc(&src_dev, 64 * 4));
    MK_HIP(hipMalloc(&off, 64 * 4));
    if (sssp) {
      MK_HIP(hipMalloc(&dist, n * 64 * 4));
      MK_HIP(hipMalloc(&logv, kMsSlots * n * 4));  // one queue per slot
      log_cap = uint32_t(std::min<uint64_t>(kMsSlots * n, 0xFFFFFFF0ull));
    } else {
      MK_HIP(hipMalloc(&seen, n * 8));
      MK_HIP(hipMalloc(&seg, (n + 2) * 4));
      if (int r = alloc_log(8 * uint64_t(n) + 64)) return r;
    }
    if (int r = host.ensure(sizeof(MsCounts))) return r;
    MK_HIP(hipStreamSynchronize(stream));
    unroll = int(tune_get("ms_unroll", 8));
    if (int r = size_grid()) return r;
    return MK_OK;
  }

  void destroy() {
    (void)hipSetDevice(g->device);
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(seen);
    for (uint32_t i = 0; i < kMsSlots; ++i) {
      (void)hipFree(F[i]);
      (void)hipFree(ql[i]);
    }
    (void)hipFree(dist);
    (void)hipFree(logv);
    (void)hipFree(logm);
    (void)hipFree(seg);
    (void)hipFree(c);
    (void)hipFree(src_dev);
    (void)hipFree(off);
    host.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) {
      loop_stream_add(g->device, stream, -1);
      (void)hipStreamDestroy(stream);
    }
  }

  int read_counts(MsCounts &h) {
    MK_HIP(hipMemcpyAsync(host.ptr, c, sizeof(MsCounts), hipMemcpyDeviceToHost, stream));
    MK_HIP(hipStreamSynchronize(stream));
    h = *host.as<MsCounts>();
    return MK_OK;
  }

  // The step kernel this handle launches.
  const void *step_kernel() const {
    if (!sssp)
      return g->weighted ? reinterpret_cast<const void *>(&k_msbfs_level<true>)
                         : reinterpret_cast<const void *>(&k_msbfs_level<false>);
    return unroll == 8 ? reinterpret_cast<const void *>(&k_mssp_step<true, 8>)
                       : reinterpret_cast<const void *>(&k_mssp_step<true, 4>);
  }

  // The steps are latency-bound: `ms_waves_per_cu` waves per CU (more than
  // the frontier loops use keep more rows in flight), capped at what a CU
  // holds of the step kernel.
  int size_grid() {
    MK_HIP(resident_waves(g->device, step_kernel(), int(tune_get("ms_waves_per_cu", 24)),
                          &max_waves));
    return MK_OK;
  }

  void reset_state() {
    const size_t n = g->n;
    for (uint32_t i = 0; i < kMsSlots; ++i) fill_u64(F[i], 0ull, n, stream);
    fill_u32(reinterpret_cast<uint32_t *>(c), 0u, sizeof(MsCounts) / 4, stream);
    fill_u32(off, 0u, 64, stream);
    if (sssp) fill_u32(dist, kInf, n * 64, stream);
    else fill_u64(seen, 0ull, n, stream);
  }

  void launch_step(const GraphView &G, const MsView &S, uint32_t s, unsigned grid) {
    // multi-source SSSP needs a weighted graph (mk_msq_create): map slabs
    if (sssp) {
      if (unroll == 8) k_mssp_step<true, 8><<<grid, kBlock, 0, stream>>>(G, S, s);
      else k_mssp_step<true, 4><<<grid, kBlock, 0, stream>>>(G, S, s);
    } else {
      if (g->weighted) k_msbfs_level<true><<<grid, kBlock, 0, stream>>>(G, S, s);
      else k_msbfs_level<false><<<grid, kBlock, 0, stream>>>(G, S, s);
    }
  }

  // ---- continuous batching
  int open() {
    MK_HIP(hipSetDevice(g->device));
    if (grow_log) {
      MK_HIP(hipStreamSynchronize(stream));
      if (int r = alloc_log(uint64_t(log_cap) * 4)) return r;
      grow_log = false;
    }
    reset_state();
    if (!sssp) fill_u32(seg, 0u, 1, stream);  // level 0 starts the log
    MK_HIP(hipGetLastError());
    session = true;
    step = 0;
    admitted = 0;
    k = 0;
    last_steps = 0;
    last_edges = 0;
    return MK_OK;
  }

  int admit(const uint32_t *sources, const uint32_t *slots, uint32_t n) {
    if (!session) return fail(MK_ERR_STATE, "mk_msq_admit: no open session (mk_msq_open)");
    MsAdmit a;
    a.k = n;
    uint64_t bits = 0;
    for (uint32_t i = 0; i < n; ++i) {
      if (slots[i] >= 64) return fail(MK_ERR_BAD_ARG, "mk_msq_admit: slot >= 64");
      if (sources[i] >= g->n) return fail(MK_ERR_RANGE, "mk_msq_admit: source >= n");
      const uint64_t bit = 1ull << slots[i];
      if ((admitted | bits) & bit)
        return fail(MK_ERR_STATE, "mk_msq_admit: slot already used in this session");
      bits |= bit;
      a.v[i] = sources[i];
      a.slot[i] = slots[i];
    }
    if (n == 0) return MK_OK;
    if (!sssp && step + 1 >= g->n + 2) return fail(MK_ERR_STATE, "mk_msq_admit: session too long");
    MK_HIP(hipSetDevice(g->device));
    k_ms_admit<<<1, 64, 0, stream>>>(g->view(), view(), a, step, sssp ? 1u : 0u);
    MK_HIP(hipGetLastError());
    admitted |= bits;
    return MK_OK;
  }

  // Queue `steps` more steps of the session (they are no-ops once its
  // frontier is empty); poll() waits for them.
  int launch(uint32_t steps) {
    if (!session) return fail(MK_ERR_STATE, "mk_msq_launch: no open session (mk_msq_open)");
    MK_HIP(hipSetDevice(g->device));
    const GraphView G = g->view();
    const MsView S = view();
    const unsigned grid = blocks_for_waves(uint64_t(max_waves));
    for (uint32_t t = 0; t < steps; ++t, ++step) {
      ProfScope prof(stream, sssp ? "k_mssp_step" : "k_msbfs_level", 0.0);
      if (!sssp && step + 1 >= g->n + 2)
        return fail(MK_ERR_STATE, "multi-source BFS: session too long");
      launch_step(G, S, step, grid);
    }
    MK_HIP(hipGetLastError());
    return MK_OK;
  }

  // Wait for the queued steps; `active` = admitted slots with work left
  // (an admitted slot not in it has its final distances).
  int poll(uint64_t *active, uint32_t *steps_done) {
    if (!session) return fail(MK_ERR_STATE, "mk_msq_poll: no open session (mk_msq_open)");
    MsCounts h{};
    if (int r = read_counts(h)) return r;
    if (!sssp && h.overflow) {
      grow_log = true;
      return fail(MK_ERR_STATE, "multi-source BFS: the level log is full; the slots still "
                                "active are incomplete (open a new session for them: it gets a "
                                "log four times larger)");
    }
    last_steps = h.steps;
    last_edges = h.edges;
    *active = h.active[step % (sssp ? kMsSlots : 3)] & admitted;
    if (steps_done) *steps_done = step;
    return MK_OK;
  }

  int run(const uint32_t *sources, uint32_t nsrc, float *ms) {
    MK_HIP(hipSetDevice(g->device));
    session = false;
    for (int attempt = 0; attempt < 4; ++attempt) {
      MK_HIP(hipEventRecord(ev0, stream));
      reset_state();
      MK_HIP(hipMemcpyAsync(src_dev, sources, nsrc * 4, hipMemcpyHostToDevice, stream));
      const GraphView G = g->view();
      const MsView S = view();
      if (sssp) k_mssp_init<<<1, 64, 0, stream>>>(S, src_dev, nsrc);
      else k_msbfs_init<<<1, 64, 0, stream>>>(S, src_dev, nsrc);
      k_ms_source_hubs<<<1, 64, 0, stream>>>(G, S);
      MK_HIP(hipGetLastError());
      const unsigned grid = blocks_for_waves(uint64_t(max_waves));
      uint32_t step = 0, group = 2;
      const uint64_t limit = uint64_t(g->n) * (sssp ? 1 : 64) + 8;
      MsCounts h{};
      for (;;) {
        for (uint32_t q = 0; q < group; ++q, ++step) {
          ProfScope prof(stream, sssp ? "k_mssp_step" : "k_msbfs_level", 0.0);
          if (!sssp && step + 1 >= g->n + 2) break;  // seg has n + 2 entries
          launch_step(G, S, step, grid);
        }
        MK_HIP(hipGetLastError());
        if (int r = read_counts(h)) return r;
        const uint32_t slot = step % (sssp ? kMsSlots : 3);
        if (h.overflow || (h.cnt[slot] == 0 && (h.large[slot] >> 32) == 0)) break;
        if (step > limit || (!sssp && step + 1 >= g->n + 2))
          return fail(MK_ERR_STATE, "multi-source traversal did not drain");
        group = std::min<uint32_t>(group * 2, 8);
      }
      if (h.overflow) {
        // the BFS log ran out of room: enlarge and run again
        if (int r = alloc_log(uint64_t(log_cap) * 4)) return r;
        continue;
      }
      MK_HIP(hipEventRecord(ev1, stream));
      MK_HIP(hipEventSynchronize(ev1));
      float t = 0.f;
      MK_HIP(hipEventElapsedTime(&t, ev0, ev1));
      if (ms) *ms = t;
      k = nsrc;
      last_steps = h.steps;
      last_edges = h.edges;
      if (profiling_on())
        // BFS: 256 B per slab, 8 B per entry (the neighbour's seen mask);
        // SSSP: 256 B per slab, 4 B per (entry, active source) distance
        prof_add_bytes(sssp ? "k_mssp_step" : "k_msbfs_level",
                       256.0 * double(h.slabs) +
                           (sssp ? 4.0 * double(h.pairs) : 8.0 * double(h.edges)));
      if (!sssp) {
        // levels 0..steps-1 are non-empty; seg[steps] is the log length
        uint32_t segs[1];
        MK_HIP(hipMemcpy(segs, seg + h.steps, 4, hipMemcpyDeviceToHost));
        log_total = segs[0];
      }
      return MK_OK;
    }
    return fail(MK_ERR_STATE, "multi-source BFS log kept overflowing");
  }

  int read(uint32_t slot, uint32_t *out) {
    if (session ? (slot >= 64 || !((admitted >> slot) & 1ull)) : slot >= k)
      return fail(MK_ERR_RANGE, session ? "mk_msq_read: slot not admitted in this session"
                                        : "mk_msq_read: slot not in the last run");
    MK_HIP(hipSetDevice(g->device));
    DevBuf b;
    if (int r = b.ensure(size_t(g->n) * 4)) return r;
    if (session && !sssp) {
      // the log so far: levels 0 .. step - 1 are written
      uint32_t total = 0;
      MK_HIP(hipMemcpyAsync(&total, seg + step, 4, hipMemcpyDeviceToHost, stream));
      MK_HIP(hipStreamSynchronize(stream));
      log_total = std::min(total, log_cap);
      last_steps = step;
    }
    if (sssp) {
      k_mssp_read<<<blocks_for_threads(g->n), kBlock, 0, stream>>>(dist, g->n, slot,
                                                                   b.as<uint32_t>());
    } else {
      fill_u32(b.as<uint32_t>(), kInf, g->n, stream);
      if (log_total)
        k_msbfs_read<<<blocks_for_threads(log_total), kBlock, 0, stream>>>(
            view(), slot, log_total, last_steps + 1, b.as<uint32_t>());
    }
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipMemcpy(out, b.ptr, size_t(g->n) * 4, hipMemcpyDeviceToHost);
    b.release();
    if (e != hipSuccess) return hip_fail(e, "mk_msq_read");
    return MK_OK;
  }
};

}  // namespace mk

struct mk_msq {
  mk::MultiSource m;
};

using namespace mk;

extern "C" {

int mk_msq_create(mk_graph *g, int sssp, mk_msq **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "mk_msq_create: out is NULL");
  *out = nullptr;
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_msq_create: NULL graph");
  if (sssp && !g->weighted) return fail(MK_ERR_BAD_ARG, "multi-source SSSP needs a weighted graph");
  mk_msq *h = new mk_msq();
  int r = h->m.init(g, sssp != 0);
  if (r) {
    h->m.destroy();
    delete h;
    return r;
  }
  g->users++;
  *out = h;
  return MK_OK;
}

int mk_msq_destroy(mk_msq *q) {
  if (!q) return fail(MK_ERR_BAD_ARG, "mk_msq_destroy: NULL handle");
  q->m.g->users--;
  q->m.destroy();
  delete q;
  return MK_OK;
}

int mk_msq_run(mk_msq *q, const uint32_t *sources_host, uint32_t k, float *ms) {
  if (!q || (k && !sources_host)) return fail(MK_ERR_BAD_ARG, "mk_msq_run: NULL argument");
  if (k == 0 || k > 64) return fail(MK_ERR_BAD_ARG, "mk_msq_run: 1..64 sources");
  for (uint32_t i = 0; i < k; ++i)
    if (sources_host[i] >= q->m.g->n) return fail(MK_ERR_RANGE, "mk_msq_run: source >= n");
  if (int r = graph_settle(q->m.g)) return r;
  return q->m.run(sources_host, k, ms);
}

int mk_msq_open(mk_msq *q) {
  if (!q) return fail(MK_ERR_BAD_ARG, "mk_msq_open: NULL handle");
  if (int r = graph_settle(q->m.g)) return r;
  return q->m.open();
}

int mk_msq_admit(mk_msq *q, const uint32_t *sources_host, const uint32_t *slots_host, uint32_t k) {
  if (!q || (k && (!sources_host || !slots_host)))
    return fail(MK_ERR_BAD_ARG, "mk_msq_admit: NULL argument");
  if (k > 64) return fail(MK_ERR_BAD_ARG, "mk_msq_admit: at most 64 sources");
  return q->m.admit(sources_host, slots_host, k);
}

int mk_msq_launch(mk_msq *q, uint32_t steps) {
  if (!q) return fail(MK_ERR_BAD_ARG, "mk_msq_launch: NULL handle");
  return q->m.launch(steps);
}

int mk_msq_ready(mk_msq *q, int *ready) {
  if (!q || !ready) return fail(MK_ERR_BAD_ARG, "mk_msq_ready: NULL argument");
  MK_HIP(hipSetDevice(q->m.g->device));
  const hipError_t e = hipStreamQuery(q->m.stream);
  if (e == hipErrorNotReady) {
    *ready = 0;
    return MK_OK;
  }
  if (e != hipSuccess) return hip_fail(e, "mk_msq_ready");
  *ready = 1;
  return MK_OK;
}

int mk_msq_poll(mk_msq *q, uint64_t *active, uint32_t *steps) {
  if (!q || !active) return fail(MK_ERR_BAD_ARG, "mk_msq_poll: NULL argument");
  return q->m.poll(active, steps);
}

int mk_msq_read(mk_msq *q, uint32_t slot, uint32_t *dist_host) {
  if (!q || !dist_host) return fail(MK_ERR_BAD_ARG, "mk_msq_read: NULL argument");
  return q->m.read(slot, dist_host);
}

int mk_msq_last_stats(mk_msq *q, uint32_t *steps, uint64_t *edges) {
  if (!q) return fail(MK_ERR_BAD_ARG, "mk_msq_last_stats: NULL handle");
  if (steps) *steps = q->m.last_steps;
  if (edges) *edges = q->m.last_edges;
  return MK_OK;
}

}  // extern "C"
