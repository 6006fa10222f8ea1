// Frontier-driven BFS and SSSP over the slab-hash graph: static,
// incremental (after InsertEdges) and decremental (after DeleteEdges).
//
// Reference: include/Algorithms/impl/BFSImpl.cuh (SetLevelStatic :498-588,
// SetLevelDynamicIncremental :270-382, SetLevelDynamicDecremental :384-496,
// Invalidate :3-15, PropogateInvalidation :17-42) and impl/SSSPImpl.cuh
// (:243-605).  State per vertex is the packed 64-bit (distance << 32 |
// parent) of include/Algorithms/definitions.cuh:12-17, relaxed with a 64-bit
// atomicMin, so the stored pair is the lexicographic minimum.
//
// Differences in method (same results):
//  - The frontier holds vertices, not edges.  A wave takes up to 64 frontier
//    vertices at once (their queue entries, bucket ranges and distances are
//    read by all lanes in parallel) and then loads their bucket slabs four at
//    a time, so several independent slab loads are in flight per wave.
//    Vertices with more than kSmallBuckets buckets are queued separately
//    with a reserved range of "bucket work items" (one 64-bit atomicAdd
//    returns the entry index and the range start together, so entries are
//    sorted by range start) and expanded bucket-by-bucket by all waves, which
//    balances hub vertices.
//  - Duplicate pushes within an iteration are dropped with a per-vertex
//    epoch stamp instead of the reference's Queued flags.
//  - Frontier counters are triple-buffered in device memory: step i reads
//    slot i%3, appends to slot (i+1)%3 and clears slot (i+2)%3, so the host
//    launches several steps back to back and reads the counters only to
//    detect the empty frontier (steps past it are no-ops).
//  - Decremental invalidation walks the shortest-path tree downward from the
//    deleted tree edges (children of x are out-neighbours whose packed parent
//    is x), O(edges of invalidated vertices), instead of every vertex walking
//    its ancestor chain.  Re-seeding scans the graph once for edges from a
//    valid vertex into an invalidated one, as the reference's frontier build
//    does (BFSImpl.cuh:406-440), but reads only buckets of reached vertices
//    with live edges and tests targets against an L2-resident bitmap.
//  - Every frontier buffer is sized for n vertices once, at create time
//    (the reference sizes them from the batch and can overflow).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "profile.hpp"
#include "slabhash.hpp"

namespace mk {

constexpr uint32_t kSmallBuckets = 2;
constexpr uint32_t kBatch = 4;       // slab loads in flight per wave

constexpr uint32_t kWaveQ = 512;     // LDS staging entries per wave
constexpr uint32_t kStatSlots = 128;  // statistics spread over this many 128-byte lines
constexpr uint32_t kTraceSteps = 64;  // steps of a loop launch recorded while profiling
enum { kStatSlabs = 0, kStatEdges = 1, kStatVisited = 2, kStatPushed = 3 };
// statistics sets, one per kind of kernel: frontier steps relaxing (MODE 0),
// invalidating (MODE 1), pull/bucket re-seeds (MODE 2), batch seeds
enum { kSetRelax = 0, kSetInvalidate = 1, kSetReseed = 2, kSetSeed = 3, kSetHubs = 4, kSets = 5 };

// Device counters.  Frontier slots rotate per step; the statistics
// accumulate over one compute call, each block adding to slot
// blockIdx % kStatSlots so no single address is contended.
struct Counts {
  uint32_t small[3];
  uint32_t steps;                 // steps that found a non-empty frontier
  uint32_t ninv;                  // entries of the invalidated-vertex list
  uint32_t abort;                 // set when a grid-barrier wait timed out
  uint32_t loop_end[2];           // step index at which k_expand_loop<MODE> stopped
  uint32_t bar_done;              // grid barriers completed (k_expand_loop)
  // near-far static SSSP (S.delta != 0): entries of the three rotating far
  // lists, the complement of each list's smallest distance (atomicMax of
  // ~d: zero means empty), and the near bound and split count a relax loop
  // leaves for the launch that continues it (0: a fresh start)
  uint32_t far_cnt[3];
  uint32_t far_minx[3];
  uint32_t nf_T, nf_split;
  uint32_t far_over;              // a deferred hub did not fit its far list
  unsigned long long large[3];    // (entries << 32) | bucket items
  unsigned long long pull_large;  // k_reseed_pull's hub list (entries << 32) | items
  unsigned long long stats[kStatSlots][kSets * 4];  // [slot][set * 4 + which], a line per slot
  // step trace of k_expand_loop<MODE> while profiling: the device clock at
  // the start of step k (k = steps: the loop's end) and its frontier
  unsigned long long trace_t[2][kTraceSteps + 1];
  uint32_t trace_small[2][kTraceSteps];
  uint32_t trace_items[2][kTraceSteps];
  uint32_t trace_n[2];
  // device clock at the start and the end of the last k_expand_loop<MODE>
  // launch (block 0) while profiling: the launch's duration
  unsigned long long loop_t[2][2];
};

// The counters a call's host side reads, written by the device straight
// into pinned host memory (`out`: HostBuf::dev of the host copy): the
// fields before the statistics, the statistics summed over their slots
// (into slot 0; the host copy's other slots stay zero), and the step trace
// while tracing; then, once those are visible, `seq` in the word after the
// copy (the host waits for it: host_wait_word).  One block.
__global__ void __launch_bounds__(256) k_publish_counts(const Counts *c, Counts *out,
                                                        uint32_t trace, uint32_t seq) {
  const uint32_t t = threadIdx.x;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(c);
  uint32_t *dst = reinterpret_cast<uint32_t *>(out);
  constexpr uint32_t head = offsetof(Counts, stats) / 4;
  for (uint32_t i = t; i < head; i += blockDim.x) dst[i] = src[i];
  if (t < kSets * 4) {
    unsigned long long x = 0;
    for (uint32_t k = 0; k < kStatSlots; ++k) x += c->stats[k][t];
    out->stats[0][t] = x;
  }
  if (trace) {
    constexpr uint32_t t0 = offsetof(Counts, trace_t) / 4, t1 = sizeof(Counts) / 4;
    for (uint32_t i = t0 + t; i < t1; i += blockDim.x) dst[i] = src[i];
  }
  __threadfence_system();
  __syncthreads();
  if (t == 0) {  // the word after the copy: seen last by the host
    reinterpret_cast<uint32_t *>(out + 1)[0] = seq;
    __threadfence_system();
  }
}

// A captured call's outcome flags (mk_step), ORed into `sticky` (device):
// 1 a grid-barrier wait timed out, 2 the relax loop used up the steps it was
// given (relax_end), 4 the invalidation loop did not drain (inval_end).
__global__ void k_step_check(const Counts *c, uint32_t relax_end, uint32_t inval_end,
                             uint32_t *sticky) {
  uint32_t f = c->abort ? 1u : 0u;
  if (relax_end && c->loop_end[0] == relax_end) f |= 2u;
  if (inval_end && c->loop_end[1] >= inval_end) f |= 4u;
  if (f) atomicOr(sticky, f);
}

// Replay bases of a captured update step: set (before the first replay) or
// advanced by one replay's epochs (the step graph's last node).
__global__ void k_step_bases(uint32_t *mut_base, uint32_t *trav_base, uint32_t m, uint32_t t,
                             uint32_t add) {
  if (add) {
    *mut_base += m;
    *trav_base += t;
  } else {
    *mut_base = m;
    *trav_base = t;
  }
}

// Host-side totals of a Counts snapshot.
struct Totals {
  uint32_t small[3];
  uint32_t steps;
  uint32_t abort, loop_end[2];
  uint32_t far_over, nf_split;
  unsigned long long large[3];
  unsigned long long slabs, edges, visited, pushed;  // over every set
  unsigned long long set[kSets][4];                 // [set][which]
};

MK_DEV void add_stat(Counts *c, uint32_t set, uint32_t which, uint32_t v) {
  if (v) atomicAdd(&c->stats[blockIdx.x % kStatSlots][set * 4 + which], (unsigned long long)v);
}

// Per-wave frontier staging in LDS: pushes are appended to the wave's LDS
// buffer and copied out with one global atomicAdd per kWaveQ entries.
struct WaveQ {
  uint32_t *buf;
  uint32_t n;       // wave-uniform fill
  uint32_t pushed;  // wave-uniform statistics
};

// Frontier queue entries and counters cross the steps of one k_expand_loop
// launch, between blocks on different XCDs (each with its own L2): they are
// written and read with agent-scope coherent accesses, so a grid barrier
// needs no L2 write-back or invalidation.
MK_DEV void qstore(uint32_t *p, uint32_t v) { coherent_store(p, v); }
MK_DEV void qstore2(uint2 *p, uint2 v) {
  __hip_atomic_store(reinterpret_cast<uint64_t *>(p), (uint64_t(v.y) << 32) | v.x,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MK_DEV uint32_t qload(const uint32_t *p) { return coherent_load(p); }
MK_DEV uint2 qload2(const uint2 *p) {
  const uint64_t x = coherent_load64(reinterpret_cast<const uint64_t *>(p));
  return make_uint2(uint32_t(x), uint32_t(x >> 32));
}
MK_DEV uint32_t qload2_y(const uint2 *p) { return qload(reinterpret_cast<const uint32_t *>(p) + 1); }

MK_DEV void wq_flush(WaveQ &q, uint32_t *queue, uint32_t *count) {
  if (q.n == 0) return;
  __builtin_amdgcn_wave_barrier();
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(count, q.n);
  base = read_lane(base, 0);
  for (uint32_t i = lane_id(); i < q.n; i += 64) qstore(queue + base + i, q.buf[i]);
  __builtin_amdgcn_wave_barrier();
  q.n = 0;
}

// Grid-barrier state of k_expand_loop, one 128-byte line per counter so
// that arrivals on different counters do not serialise: blocks arrive in
// groups of kBarGroup on their group's line; the last of a group arrives on
// the top line and, once all groups are in, writes the barrier's number to
// its group's release line.  Counts only grow (zeroed with the frontier
// counters at the start of a compute call); barrier k is complete when the
// top count reaches k x groups.
constexpr uint32_t kBarGroup = 32;
constexpr uint32_t kMaxBarGroups = 128;  // 4096 blocks
struct BarLines {
  uint32_t top[32];
  uint32_t grp[kMaxBarGroups][32];
  uint32_t rel[kMaxBarGroups][32];
};

struct TravView {
  uint64_t *dist;
  uint32_t *stamp;
  uint32_t *invbits;
  uint32_t *inv_list;  // invalidated vertices, compacted from invbits
  uint32_t *q_small[2];
  uint2 *q_large[2];
  Counts *cnt;
  struct BarLines *bar;  // grid-barrier counters (k_expand_loop)
  uint32_t tail_enter, tail_stay;  // tail-mode thresholds (0: no tail mode)
  uint32_t trace;                  // record the step trace (profiling)
  // Bottom-up (pull) steps of a static BFS over the in-edge mirror R
  // (vbits[0] == nullptr: none).  vbits[side] marks the vertices pushed to
  // queue side `side` (the frontier the step reading that side expands);
  // a bitmap also keeps bits of earlier frontiers of the same parity, which
  // is harmless: an unvisited vertex has no in-neighbour at a level below
  // the current one.  A step runs bottom-up once the frontier holds more
  // than 1/pull_enter of the graph's buckets, and stays so until it holds
  // fewer than 1/pull_exit of the vertices.
  GraphView R;
  uint32_t *vbits[2];
  uint32_t pull_enter, pull_exit;
  // Near-far static SSSP (delta == 0: plain label-correcting).  A frontier
  // vertex whose distance is at or above the near bound T is not expanded
  // but deferred to the current far list (vertex, distance); once the near
  // frontier is empty, a split step raises T to the smallest deferred
  // distance + delta and queues the far entries now below it (entries whose
  // vertex has since improved are dropped: a later entry stands for it).
  // Label-correcting is exact in any order, so a full far list only means
  // expanding a vertex early.
  uint2 *far[3];
  uint32_t far_cap;
  uint32_t delta;
  // added to every stamp epoch a kernel is given: zero for calls made
  // through the API, the replay's base while a captured update step
  // (mk_step) runs, so each replay stamps with epochs of its own
  const uint32_t *ep_base;
};

MK_DEV uint32_t wave_min_u32(uint32_t x) {
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t o = uint32_t(__shfl_xor(int(x), off, 64));
    x = o < x ? o : x;
  }
  return x;
}

// Append (v, d) of the lanes with `put` to far list `f`, one atomic per
// wave; returns whether this lane's entry was stored (false past far_cap).
// Wave-collective.
MK_DEV bool far_append(const TravView &S, uint32_t f, bool put, uint32_t v, uint32_t d) {
  const uint64_t m = ballot(put);
  if (m == 0) return false;
  const uint32_t lane = lane_id(), leader = ffs64(m);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&S.cnt->far_cnt[f], uint32_t(__popcll(m)));
  base = read_lane(base, leader);
  const uint32_t pos = base + lanes_below(m);
  const bool stored = put && pos < S.far_cap;
  if (stored) qstore2(S.far[f] + pos, make_uint2(v, d));
  const uint32_t mn = wave_min_u32(stored ? d : kInf);
  if (lane == leader && mn != kInf) atomicMax(&S.cnt->far_minx[f], ~mn);
  return stored;
}

MK_DEV void mark_frontier(const TravView &S, uint32_t side, bool push, uint32_t v) {
  if (push && S.vbits[0]) atomicOr(S.vbits[side] + (v >> 5), 1u << (v & 31u));
}

// Append v (lanes with `push`) to queue side `side`, counter slot `slot`:
// small vertices through the wave's LDS staging, hub vertices directly.
// Wave-collective.
MK_DEV void enqueue_nb(const GraphView &G, const TravView &S, WaveQ &q, uint32_t side,
                       uint32_t slot, bool push, uint32_t v, uint32_t nb) {
  const uint64_t pm = ballot(push);
  if (pm == 0) return;
  mark_frontier(S, side, push, v);
  const bool small = push && nb <= kSmallBuckets;
  const uint64_t sm = ballot(small);
  const uint32_t c = uint32_t(__popcll(sm));
  if (q.n + c > kWaveQ) wq_flush(q, S.q_small[side], &S.cnt->small[slot]);
  if (small) q.buf[q.n + lanes_below(sm)] = v;
  q.n += c;
  push_hubs(G, S.q_large[side], &S.cnt->large[slot], push && !small, v, nb);
  q.pushed += uint32_t(__popcll(pm));
}

// Bucket count of v for the lanes that push it (0 elsewhere).
MK_DEV uint32_t push_buckets(const GraphView &G, bool push, uint32_t v) {
  return push ? G.bucket_off[v + 1] - G.bucket_off[v] : 0u;
}

MK_DEV void enqueue(const GraphView &G, const TravView &S, WaveQ &q, uint32_t side,
                    uint32_t slot, bool push, uint32_t v) {
  if (ballot(push) == 0) return;
  enqueue_nb(G, S, q, side, slot, push, v, push_buckets(G, push, v));
}

// wq_flush for every wave of the block at once, one atomic per block;
// `sh` holds blockDim.x / 64 + 1 words of LDS.  Called by every thread.
MK_DEV void block_flush(WaveQ &q, uint32_t *queue, uint32_t *count, uint32_t *sh) {
  const uint32_t NW = blockDim.x >> 6;
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) sh[w] = q.n;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t k = 0; k < NW; ++k) {
      const uint32_t c = sh[k];
      sh[k] = tot;
      tot += c;
    }
    sh[NW] = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  const uint32_t base = sh[NW] + sh[w];
  for (uint32_t i = lane_id(); i < q.n; i += 64) qstore(queue + base + i, q.buf[i]);
  q.n = 0;
}

// Push targets of the expansion helpers below (wave-collective calls).
// GlobalSink appends to a step's global queues through the wave's LDS
// staging (enqueue); TailSink appends to a block's LDS queue (tail mode of
// k_expand), spilling to the global queues of the next step when the LDS
// queue is full or the vertex is a hub.
struct GlobalSink {
  const TravView &S;
  WaveQ &q;
  uint32_t side, slot;
  MK_DEV void operator()(const GraphView &G, bool push, uint32_t v) {
    enqueue(G, S, q, side, slot, push, v);
  }
  // the pushing lanes' bucket counts already loaded
  MK_DEV void operator()(const GraphView &G, bool push, uint32_t v, uint32_t nb) {
    enqueue_nb(G, S, q, side, slot, push, v, nb);
  }
};

constexpr uint32_t kTailCap = 2048;    // LDS queue entries per tail-mode block
// Frontiers of up to S.tail_enter vertices (default 128, the knob
// "tail_enter") run in tail mode; a tail frontier above S.tail_stay (1.5x)
// is handed back.
constexpr uint32_t kEpochStride = 256; // epochs reserved per launch (tail steps)

struct TailSink {
  const TravView &S;
  uint32_t *lq;     // LDS queue, kTailCap entries
  uint32_t *ctl;    // LDS: [0] entries claimed, [1] handover flag
  uint32_t side, slot;  // global queues of the next step
  uint32_t pushed;      // wave-uniform statistics
  MK_DEV void operator()(const GraphView &G, bool push, uint32_t v) {
    if (ballot(push) == 0) return;
    (*this)(G, push, v, push_buckets(G, push, v));
  }
  // the pushing lanes' bucket counts already loaded
  MK_DEV void operator()(const GraphView &G, bool push, uint32_t v, uint32_t nb) {
    const uint64_t pm = ballot(push);
    if (pm == 0) return;
    mark_frontier(S, side, push, v);
    const uint32_t lane = lane_id();
    const bool small = push && nb <= kSmallBuckets;
    const uint64_t sm = ballot(small);
    const uint32_t c = uint32_t(__popcll(sm));
    if (c) {
      const uint32_t leader = ffs64(sm);
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(ctl, c);
      base = read_lane(base, leader);
      if (base + c <= kTailCap) {
        if (small) lq[base + lanes_below(sm)] = v;
      } else {
        // LDS queue full: these go to the global queue (positions claimed
        // in LDS are marked empty) and the block hands the frontier back
        uint32_t gb = 0;
        if (lane == leader) gb = atomicAdd(&S.cnt->small[slot], c);
        gb = read_lane(gb, leader);
        if (small) {
          const uint32_t pos = base + lanes_below(sm);
          if (pos < kTailCap) lq[pos] = kEmpty;
          qstore(S.q_small[side] + gb + lanes_below(sm), v);
        }
        if (lane == leader) ctl[1] = 1u;
      }
    }
    // hubs go to the next step's bucket work list
    push_hubs(G, S.q_large[side], &S.cnt->large[slot], push && !small, v, nb);
    if (push && !small) ctl[1] = 1u;
    pushed += uint32_t(__popcll(pm));
  }
};

// Block-aggregated append for kernels with one item per thread: one global
// atomicAdd per block.  `nb`: bucket count of v on the pushing lanes
// (loaded by the caller alongside its atomics).  Called by every thread of
// the block.
template <uint32_t NW>
MK_DEV void block_enqueue(const GraphView &G, const TravView &S, uint32_t side, uint32_t slot,
                          bool push, uint32_t v, uint32_t nb, uint32_t *sh) {
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  const bool small = push && nb <= kSmallBuckets;
  const uint64_t sm = ballot(small);
  if (lane == 0) sh[w] = uint32_t(__popcll(sm));
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t k = 0; k < NW; ++k) {
      const uint32_t c = sh[k];
      sh[k] = tot;
      tot += c;
    }
    sh[NW] = tot ? atomicAdd(&S.cnt->small[slot], tot) : 0u;
  }
  __syncthreads();
  if (small) qstore(S.q_small[side] + sh[NW] + sh[w] + lanes_below(sm), v);
  push_hubs(G, S.q_large[side], &S.cnt->large[slot], push && !small, v, nb);
  const uint64_t pm = ballot(push);
  if (lane == 0) add_stat(S.cnt, kSetSeed, kStatPushed, uint32_t(__popcll(pm)));
}

MK_DEV uint32_t wave_sum(uint32_t x) {
  for (int off = 32; off >= 1; off >>= 1) x += uint32_t(__shfl_xor(int(x), off, 64));
  return x;
}

MK_DEV bool inv_bit(const uint32_t *bits, uint32_t v) { return (bits[v >> 5] >> (v & 31u)) & 1u; }

// Per-slab work of the three passes.  `v` is this lane's word of one slab of
// vertex u (distance du).  Wave-collective; returns live keys seen.
template <bool MAP, bool WEIGHTS, int MODE, typename Sink>
MK_DEV uint32_t slab_work(const GraphView &G, const TravView &S, Sink &sink, uint32_t epoch,
                          uint32_t u, uint32_t du, uint32_t v) {
  const uint32_t lane = lane_id();
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  const bool live = key_lane && v < kTomb;
  const uint32_t wv = MAP ? __shfl_down(v, 1, 64) : 1u;
  bool push = false;
  if (MODE == 0) {  // relax u -> v
    if (live) {
      const uint32_t D = du + (WEIGHTS ? wv : 1u);
      if (D >= du && D < dist_of(S.dist[v])) {
        const uint64_t old = atomicMin((unsigned long long *)(S.dist + v),
                                       (unsigned long long)pack(D, u));
        if (D < dist_of(old)) push = atomicExch(S.stamp + v, epoch) != epoch;
      }
    }
  } else if (MODE == 1) {  // invalidate tree children of u
    if (live) {
      // values only move to INF during this pass, so a plain load suffices
      const uint64_t cur = S.dist[v];
      if (cur != kInfInf && parent_of(cur) == u) {
        if (atomicCAS((unsigned long long *)(S.dist + v), (unsigned long long)cur,
                      (unsigned long long)kInfInf) == cur) {
          atomicOr(S.invbits + (v >> 5), 1u << (v & 31u));
          push = true;
        }
      }
    }
  } else {  // re-seed invalidated v from valid u
    if (live && inv_bit(S.invbits, v)) {
      const uint32_t D = du + (WEIGHTS ? wv : 1u);
      if (D >= du) {
        const uint64_t old = atomicMin((unsigned long long *)(S.dist + v),
                                       (unsigned long long)pack(D, u));
        if (D < dist_of(old)) push = atomicExch(S.stamp + v, epoch) != epoch;
      }
    }
  }
  sink(G, push, v);
  return uint32_t(__popcll(ballot(live)));
}

// Work of a whole chain given its first slab's word `v` (already loaded).
template <bool MAP, bool WEIGHTS, int MODE, typename Sink>
MK_DEV void chain_work(const GraphView &G, const TravView &S, Sink &sink, uint32_t epoch,
                       uint32_t u, uint32_t du, uint32_t v, uint32_t &slabs, uint32_t &edges) {
  for (;;) {
    ++slabs;
    edges += slab_work<MAP, WEIGHTS, MODE>(G, S, sink, epoch, u, du, v);
    const uint32_t nx = read_lane(v, kNextLane);
    if (nx == kEmpty) break;
    v = slab_ptr(G.pool, nx)[lane_id()];
  }
}

// slab_work over c <= kBatch slabs at once (slab k of vertex u[k] with
// distance du[k], this lane's word v[k]), phase by phase, so that the
// dependent memory operations of the slabs overlap instead of following
// one another: every distance (invalid bit) load, then every atomic (with
// the bucket count of each candidate vertex), then every stamp (bitmap)
// update, then the pushes.  Wave-collective; returns live keys seen.
template <bool MAP, bool WEIGHTS, int MODE, typename Sink>
MK_DEV uint32_t batch_slab_work(const GraphView &G, const TravView &S, Sink &sink, uint32_t epoch,
                                uint32_t c, const uint32_t (&u)[kBatch],
                                const uint32_t (&du)[kBatch], const uint32_t (&v)[kBatch]) {
  const uint32_t lane = lane_id();
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  bool live[kBatch], ok[kBatch];
  uint32_t D[kBatch];
  uint64_t cur[kBatch];
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) {
    live[k] = k < c && key_lane && v[k] < kTomb;
    const uint32_t wv = MAP ? __shfl_down(v[k], 1, 64) : 1u;
    D[k] = du[k] + (WEIGHTS ? wv : 1u);
    ok[k] = live[k] && (MODE == 1 || D[k] >= du[k]);
    cur[k] = kInfInf;
    if (MODE == 2) {
      if (ok[k]) ok[k] = inv_bit(S.invbits, v[k]);
    } else if (ok[k]) {
      // MODE 1: values only move to INF during this pass, so a plain load
      // suffices; MODE 0: a stale value only lets an extra atomic through
      cur[k] = S.dist[v[k]];
    }
  }
  // the atomics, with the bucket counts of their lanes loaded alongside
  // (read only where the vertex is then pushed)
  bool push[kBatch];
  uint32_t nb[kBatch];
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) {
    push[k] = false;
    const bool cand = MODE == 1 ? ok[k] && cur[k] != kInfInf && parent_of(cur[k]) == u[k]
                                : ok[k] && (MODE == 2 || D[k] < dist_of(cur[k]));
    nb[k] = push_buckets(G, cand, v[k]);
    if (MODE == 1) {
      if (cand)
        push[k] = atomicCAS((unsigned long long *)(S.dist + v[k]), (unsigned long long)cur[k],
                            (unsigned long long)kInfInf) == cur[k];
    } else if (cand) {
      cur[k] = atomicMin((unsigned long long *)(S.dist + v[k]), (unsigned long long)pack(D[k], u[k]));
      push[k] = D[k] < dist_of(cur[k]);  // improved
    }
  }
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) {
    if (MODE == 1) {
      if (push[k]) atomicOr(S.invbits + (v[k] >> 5), 1u << (v[k] & 31u));
    } else if (push[k]) {
      push[k] = atomicExch(S.stamp + v[k], epoch) != epoch;
    }
  }
  uint32_t edges = 0;
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) {
    if (k < c) sink(G, push[k], v[k], nb[k]);
    edges += uint32_t(__popcll(ballot(live[k])));
  }
  return edges;
}

// Process the buckets named by lanes of `m`: lane j owns bucket bk (its
// vertex uj with distance dj).  Head slabs are loaded and worked kBatch at
// a time (batch_slab_work); chained slabs follow one by one.
template <bool MAP, bool WEIGHTS, int MODE, typename Sink>
MK_DEV void batched_buckets(const GraphView &G, const TravView &S, Sink &sink, uint32_t epoch,
                            uint64_t m, uint32_t bk, uint32_t uj, uint32_t dj, uint32_t &slabs,
                            uint32_t &edges) {
  const uint32_t lane = lane_id();
  while (m) {
    uint32_t js[kBatch];
    uint32_t c = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      js[k] = m ? ffs64(m) : 0u;
      if (m) {
        m &= m - 1;
        c = k + 1;
      }
    }
    uint32_t v[kBatch], us[kBatch], ds[kBatch];
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      v[k] = k < c ? G.heads[size_t(read_lane(bk, js[k])) * kSlabWords + lane] : kEmpty;
      us[k] = read_lane(uj, js[k]);
      ds[k] = read_lane(dj, js[k]);
    }
    slabs += c;
    edges += batch_slab_work<MAP, WEIGHTS, MODE>(G, S, sink, epoch, c, us, ds, v);
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      const uint32_t nx = k < c ? read_lane(v[k], kNextLane) : kEmpty;
      if (nx != kEmpty)
        chain_work<MAP, WEIGHTS, MODE>(G, S, sink, epoch, us[k], ds[k],
                                       slab_ptr(G.pool, nx)[lane], slabs, edges);
    }
  }
}

// Expand frontier vertices queue[0..n) (non-hub entries), waves wave,
// wave + NW, ... of the caller taking up to 64 each.
template <bool MAP, bool WEIGHTS, int MODE, typename Sink>
MK_DEV void expand_rows(const GraphView &G, const TravView &S, Sink &sink, uint32_t epoch,
                        const uint32_t *queue, uint32_t n, uint32_t wave, uint32_t NW,
                        uint32_t &slabs, uint32_t &edges, uint32_t &visited, uint32_t T = kInf,
                        uint32_t fcur = 0) {
  const uint32_t lane = lane_id();
  const uint32_t per = per_wave(n, NW);
  for (uint32_t i0 = wave * per; i0 < n; i0 += NW * per) {
    const uint32_t i = i0 + lane;
    bool act = lane < per && i < n;
    uint32_t u = 0, b0 = 0, nb = 0, du = 0;
    if (act) {
      u = qload(queue + i);
      b0 = G.bucket_off[u];
      nb = G.bucket_off[u + 1] - b0;
      du = MODE == 1 ? 0u : dist_of(coherent_load64(S.dist + u));
      if (MODE == 0 && du == kInf) act = false;
    }
    if constexpr (MODE == 0 && WEIGHTS) {
      // near-far: vertices at or past the near bound wait in the far list
      if (S.delta && far_append(S, fcur, act && du >= T, u, du)) act = false;
    }
    visited += uint32_t(__popcll(ballot(act)));
    batched_buckets<MAP, WEIGHTS, MODE>(G, S, sink, epoch, ballot(act), b0, u, du, slabs, edges);
    // further buckets (at most kSmallBuckets - 1 for queued vertices)
    for (uint32_t k = 1; ballot(act && nb > k); ++k)
      batched_buckets<MAP, WEIGHTS, MODE>(G, S, sink, epoch, ballot(act && nb > k), b0 + k, u,
                                          du, slabs, edges);
  }
}

// Tail mode of a frontier step: block 0 alone runs this step and the
// following ones, frontiers ping-ponging between two LDS queues, while they
// stay small; step j uses epoch + j.  The frontier is handed back (written
// to the global queues of the next step, where the next launch reads it)
// when it grows past S.tail_stay, overflows the LDS queue, reaches a hub, or
// the epoch range reserved for this launch runs out.
template <bool MAP, bool WEIGHTS, int MODE>
MK_DEV void expand_tail(const GraphView &G, const TravView &S, uint32_t side, uint32_t slot,
                        uint32_t epoch, uint32_t n0, uint32_t *lds, uint32_t T = kInf,
                        uint32_t fcur = 0) {
  uint32_t *qa = lds, *qb = lds + kTailCap, *ctl = lds + 2 * kTailCap;  // ctl: 2 x {count, flag}
  const uint32_t nside = side ^ 1u, nslot = slot == 2 ? 0 : slot + 1;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6, NW = blockDim.x >> 6;
  if (tid < 4) ctl[tid] = 0u;
  for (uint32_t i = tid; i < n0; i += blockDim.x) qa[i] = qload(S.q_small[side] + i);
  uint32_t nin = n0, slabs = 0, edges = 0, visited = 0, pushed = 0;
  for (uint32_t j = 0;; ++j) {
    __syncthreads();  // the previous step is complete everywhere; qa holds its frontier
    uint32_t *c = ctl + 2 * (j & 1u);
    if (tid == 0) {  // the other parity is read next step; nobody reads it now
      ctl[2 * ((j + 1) & 1u)] = 0u;
      ctl[2 * ((j + 1) & 1u) + 1] = 0u;
    }
    TailSink sink{S, qb, c, nside, nslot, 0u};
    expand_rows<MAP, WEIGHTS, MODE>(G, S, sink, epoch + j, qa, nin, wave, NW, slabs, edges,
                                    visited, T, fcur);
    pushed += sink.pushed;
    __syncthreads();
    const uint32_t nb = c[0], handover = c[1];
    if (nb == 0 && handover == 0) break;  // the traversal is done
    if (handover || nb > S.tail_stay || j + 2 >= kEpochStride) {
      // append the LDS part of the frontier to the next step's global queue
      const uint32_t m = min(nb, kTailCap);
      for (uint32_t i0 = wave * 64; i0 < m; i0 += NW * 64) {
        const uint32_t i = i0 + lane;
        const uint32_t v = i < m ? qb[i] : kEmpty;
        const uint64_t vm = ballot(v != kEmpty);
        if (vm == 0) continue;
        const uint32_t leader = ffs64(vm);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&S.cnt->small[nslot], uint32_t(__popcll(vm)));
        base = read_lane(base, leader);
        if (v != kEmpty) qstore(S.q_small[nside] + base + lanes_below(vm), v);
      }
      break;
    }
    if (tid == 0) atomicAdd(&S.cnt->steps, 1u);  // one more step runs here
    uint32_t *t = qa;
    qa = qb;
    qb = t;
    nin = nb;
  }
  if (lane == 0) {
    add_stat(S.cnt, MODE, kStatSlabs, slabs);
    add_stat(S.cnt, MODE, kStatEdges, edges);
    add_stat(S.cnt, MODE, kStatVisited, visited);
    add_stat(S.cnt, MODE, kStatPushed, pushed);
  }
}

// One frontier step: read queue side `side`, counter slot `slot`; append to
// side^1 / slot+1; clear slot+2.  MODE 0 relaxes, MODE 1 invalidates.  A
// small frontier without hub entries is run in tail mode by block 0.
// Called by every thread of every block of the grid; `lds` holds
// kExpandLds words (kLoopLds for blocks of kWideBlock threads).
constexpr uint32_t kWaveLds = (kBlock / 64) * kWaveQ + kBlock / 64 + 1;  // staging + block_flush
// Blocks of the persistent loop are four times larger: a quarter of the
// blocks arrive at each grid barrier and flush their staged frontier with
// one atomic each on the same counter (such atomics serialise).
constexpr uint32_t kWideBlock = 1024;
constexpr uint32_t kWideWaveLds = (kWideBlock / 64) * kWaveQ + kWideBlock / 64 + 1;
constexpr uint32_t kLoopLds = 2 * kTailCap + 4 > kWideWaveLds ? 2 * kTailCap + 4 : kWideWaveLds;
constexpr uint32_t kExpandLds = 2 * kTailCap + 4 > kWaveLds ? 2 * kTailCap + 4 : kWaveLds;

template <bool MAP, bool WEIGHTS, int MODE>
MK_DEV void expand_step(const GraphView &G, const TravView &S, uint32_t side, uint32_t slot,
                        uint32_t epoch, uint32_t *lds, uint32_t T = kInf, uint32_t fcur = 0) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t nslot = slot == 2 ? 0 : slot + 1;
  const uint32_t cslot = nslot == 2 ? 0 : nslot + 1;
  const uint32_t nside = side ^ 1u;
  const uint32_t n_small = coherent_load(&S.cnt->small[slot]);
  const unsigned long long lg = coherent_load64(reinterpret_cast<const uint64_t *>(&S.cnt->large[slot]));
  const uint32_t n_large = uint32_t(lg >> 32);  // hub-list entries
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    coherent_store(&S.cnt->small[cslot], 0u);
    __hip_atomic_store(reinterpret_cast<uint64_t *>(&S.cnt->large[cslot]), uint64_t(0),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (n_small + n_large) atomicAdd(&S.cnt->steps, 1u);
  }
  if (n_small + n_large == 0) return;
  if (n_large == 0 && n_small <= S.tail_enter) {
    if (blockIdx.x == 0)
      expand_tail<MAP, WEIGHTS, MODE>(G, S, side, slot, epoch, n_small, lds, T, fcur);
    return;
  }
  WaveQ q{lds + (threadIdx.x >> 6) * kWaveQ, 0u, 0u};
  GlobalSink sink{S, q, nside, nslot};
  uint32_t slabs = 0, edges = 0, visited = 0;
  // hub vertices: their list entries (kHubChunk buckets each) spread over
  // all waves; lane l takes bucket l % kHubChunk of entry l / kHubChunk
  const uint2 *ql = S.q_large[side];
  const uint32_t eper = min(per_wave(n_large, W), 64u / kHubChunk);
  for (uint32_t e0 = gw * eper; e0 < n_large; e0 += W * eper) {
    const uint32_t el = lane / kHubChunk, ei = e0 + el;
    bool act = el < eper && ei < n_large;
    uint32_t u = 0, b = 0, du = 0;
    if (act) {
      const uint2 ent = qload2(ql + ei);
      u = ent.x;
      b = ent.y + lane % kHubChunk;
      act = b < G.bucket_off[u + 1];
      if (act) {
        du = MODE == 1 ? 0u : dist_of(coherent_load64(S.dist + u));
        if (MODE == 0 && du == kInf) act = false;
      }
    }
    if constexpr (MODE == 0 && WEIGHTS) {
      if (S.delta) {
        // near-far: a hub at or past the near bound waits in the far list,
        // deferred once, by the lane of its first bucket (the split queues
        // the whole vertex again); its other entries' lanes just skip
        const bool later = act && du >= T;
        const bool first = later && b == G.bucket_off[u];
        const bool stored = far_append(S, fcur, first, u, du);
        if (later) act = false;
        // a hub that did not fit (full far list) has lost buckets other
        // lanes skipped: the call is repeated without near-far
        if (first && !stored) atomicExch(&S.cnt->far_over, 1u);
      }
    }
    batched_buckets<MAP, WEIGHTS, MODE>(G, S, sink, epoch, ballot(act), b, u, du, slabs, edges);
  }
  // small vertices: up to 64 per wave step, spread over all waves when few
  expand_rows<MAP, WEIGHTS, MODE>(G, S, sink, epoch, S.q_small[side], n_small, gw, W, slabs, edges,
                                  visited, T, fcur);
  block_flush(q, S.q_small[nside], &S.cnt->small[nslot], lds + (blockDim.x / 64) * kWaveQ);
  if (lane == 0) {
    add_stat(S.cnt, MODE, kStatSlabs, slabs);
    add_stat(S.cnt, MODE, kStatEdges, edges);
    add_stat(S.cnt, MODE, kStatVisited, visited);
    add_stat(S.cnt, MODE, kStatPushed, q.pushed);
  }
}

// One bottom-up step of a static BFS (no reference counterpart: the
// reference's BFS is top-down only, BFSImpl.cuh:498-588; its Rev* policies,
// SlabGraph.cuh:113-132, are the in-edge view used here).  The frontier
// (queue side `side`, level L) is marked in S.vbits[side]; every unvisited
// vertex v that ever had an in-edge (R.vdeg[v] != 0) scans its in-slabs in
// the mirror R for a marked in-neighbour u and stops at the first: v takes
// level L + 1 with parent u and is pushed to side^1 / slot+1 (marked for the
// next step).  An unvisited vertex has no in-neighbour below level L, so
// every marked in-neighbour is at level L and the levels are those of the
// top-down step; the parent is a valid one (any frontier in-neighbour), not
// necessarily the smallest.  Each wave takes 64 consecutive vertices; its
// eight 8-lane groups scan one candidate each at a time.  Called by every
// thread of every block; `lds` as expand_step.
template <bool MAP>
MK_DEV void pull_step(const GraphView &G, const TravView &S, uint32_t side, uint32_t slot,
                      uint32_t *lds) {
  const uint32_t lane = lane_id(), gl = group_lane(), grp = lane / kGroupLanes;
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t nslot = slot == 2 ? 0 : slot + 1;
  const uint32_t cslot = nslot == 2 ? 0 : nslot + 1;
  const uint32_t nside = side ^ 1u;
  const uint32_t n_small = coherent_load(&S.cnt->small[slot]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    coherent_store(&S.cnt->small[cslot], 0u);
    __hip_atomic_store(reinterpret_cast<uint64_t *>(&S.cnt->large[cslot]), uint64_t(0),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(&S.cnt->steps, 1u);
  }
  // the frontier's level, from any of its entries (the caller saw one)
  const uint32_t u0 = n_small ? qload(S.q_small[side]) : qload2(S.q_large[side]).x;
  const uint32_t L = dist_of(coherent_load64(S.dist + u0));
  const uint32_t *vb = S.vbits[side];
  const GraphView &R = S.R;
  WaveQ q{lds + (threadIdx.x >> 6) * kWaveQ, 0u, 0u};
  GlobalSink sink{S, q, nside, nslot};
  uint32_t slabs = 0, edges = 0, visited = 0;
  for (uint64_t v0 = uint64_t(gw) * 64; v0 < G.n; v0 += uint64_t(W) * 64) {
    const uint32_t v = uint32_t(v0) + lane;
    bool cand = v < G.n && R.vdeg[v] != 0u;
    if (cand) cand = dist_of(coherent_load64(S.dist + v)) == kInf;
    uint64_t cm = ballot(cand);
    visited += uint32_t(__popcll(cm));
    while (cm) {
      // candidate k of this round (k = group) is the k-th lowest set bit
      uint64_t m = cm;
      for (uint32_t k = 0; k < grp && m; ++k) m &= m - 1;
      const bool has = m != 0;
      const uint32_t vl = has ? ffs64(m) : 0u;  // its lane
      for (uint32_t k = 0; k < kGroupsPerWave && cm; ++k) cm &= cm - 1;
      const uint32_t vk = uint32_t(v0) + vl;
      uint32_t parent = kEmpty;
      if (has) {
        const uint32_t b1 = R.bucket_off[vk + 1];
        for (uint32_t b = R.bucket_off[vk]; b < b1 && parent == kEmpty; ++b) {
          const uint32_t *slab = R.heads + size_t(b) * kSlabWords;
          for (;;) {
            uint32_t w[kWordsPerLane];
            load_words(slab + kWordsPerLane * gl, w);
            if (gl == 0) ++slabs;
            // every live key's bitmap word first (independent loads), then the tests
            uint32_t bits[kWordsPerLane];
#pragma unroll
            for (uint32_t k = 0; k < kWordsPerLane; ++k) {
              const bool key = key_word<MAP>(kWordsPerLane * gl + k) && w[k] < kTomb;
              bits[k] = key ? coherent_load(vb + (w[k] >> 5)) : 0u;
              edges += key ? 1u : 0u;
            }
            uint32_t hit = kEmpty;
#pragma unroll
            for (uint32_t k = 0; k < kWordsPerLane; ++k)
              if (hit == kEmpty && key_word<MAP>(kWordsPerLane * gl + k) && w[k] < kTomb &&
                  ((bits[k] >> (w[k] & 31u)) & 1u))
                hit = w[k];
            const uint32_t hb = group_bits(ballot(hit != kEmpty));
            if (hb) {
              parent = group_read(hit, ffs32(hb));
              break;
            }
            const uint32_t nx = group_read(w[kWordsPerLane - 1], kGroupLanes - 1);
            if (nx == kEmpty) break;
            slab = slab_ptr(R.pool, nx);
          }
        }
        if (gl == 0 && parent != kEmpty) {
          __hip_atomic_store(reinterpret_cast<uint64_t *>(S.dist + vk), pack(L + 1, parent),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      sink(G, gl == 0 && has && parent != kEmpty, vk);
    }
  }
  block_flush(q, S.q_small[nside], &S.cnt->small[nslot], lds + (blockDim.x / 64) * kWaveQ);
  // slabs were counted on group lane 0, keys on every lane
  slabs = wave_sum(slabs);
  edges = wave_sum(edges);
  if (lane == 0) {
    add_stat(S.cnt, kSetRelax, kStatSlabs, slabs);
    add_stat(S.cnt, kSetRelax, kStatEdges, edges);
    add_stat(S.cnt, kSetRelax, kStatVisited, visited);
    add_stat(S.cnt, kSetRelax, kStatPushed, q.pushed);
  }
}

// The split step of near-far SSSP, run in place of a frontier step when the
// near frontier is empty and far list `fin` is not: with the near bound
// already raised to T, each entry (v, d) of `fin` whose vertex still has
// distance d (else a later entry stands for it) is queued for the next step
// (queue side^1, slot+1, once per vertex: stamp `epoch`) when d < T, and
// kept in far list `fout` otherwise (queued instead if `fout` is full).
// Far list `fclr` (consumed by the previous split) is cleared for the next.
// Called by every thread of every block; `lds` as expand_step.
MK_DEV void split_step(const GraphView &G, const TravView &S, uint32_t side, uint32_t slot,
                       uint32_t epoch, uint32_t T, uint32_t fin, uint32_t fout, uint32_t fclr,
                       uint32_t *lds) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t nslot = slot == 2 ? 0 : slot + 1;
  const uint32_t cslot = nslot == 2 ? 0 : nslot + 1;
  const uint32_t nside = side ^ 1u;
  const uint32_t fc = min(coherent_load(&S.cnt->far_cnt[fin]), S.far_cap);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    coherent_store(&S.cnt->small[cslot], 0u);
    __hip_atomic_store(reinterpret_cast<uint64_t *>(&S.cnt->large[cslot]), uint64_t(0),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    coherent_store(&S.cnt->far_cnt[fclr], 0u);
    coherent_store(&S.cnt->far_minx[fclr], 0u);
    atomicAdd(&S.cnt->steps, 1u);
  }
  WaveQ q{lds + (threadIdx.x >> 6) * kWaveQ, 0u, 0u};
  GlobalSink sink{S, q, nside, nslot};
  uint32_t visited = 0;
  for (uint32_t i0 = gw * 64; i0 < fc; i0 += W * 64) {
    const uint32_t i = i0 + lane;
    bool live = i < fc;
    uint2 e = make_uint2(0u, 0u);
    if (live) {
      e = qload2(S.far[fin] + i);
      live = dist_of(coherent_load64(S.dist + e.x)) == e.y;
    }
    visited += uint32_t(__popcll(ballot(live)));
    const bool kept = far_append(S, fout, live && e.y >= T, e.x, e.y);
    bool push = live && !kept;
    if (push) push = atomicExch(S.stamp + e.x, epoch) != epoch;
    sink(G, push, e.x);
  }
  block_flush(q, S.q_small[nside], &S.cnt->small[nslot], lds + (blockDim.x / 64) * kWaveQ);
  if (lane == 0) {
    add_stat(S.cnt, kSetRelax, kStatVisited, visited);
    add_stat(S.cnt, kSetRelax, kStatPushed, q.pushed);
  }
}

template <bool MAP, bool WEIGHTS, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_expand(GraphView G, TravView S, uint32_t side, uint32_t slot, uint32_t epoch) {
  epoch += *S.ep_base;
  __shared__ uint32_t lds[kExpandLds];
  expand_step<MAP, WEIGHTS, MODE>(G, S, side, slot, epoch, lds);
}

// Grid-wide barrier number `k` (1, 2, ... since the counters were zeroed,
// counted over every launch in between, all of one grid size) of a kernel
// whose blocks are all resident.  Every
// wave first waits for its own memory operations to complete (the data that
// crosses steps is written with coherent stores and atomics, so complete
// means visible to every XCD); thread 0 of each block then arrives as
// BarLines describes.  A wait longer than `timeout` wall-clock ticks sets
// cnt->abort instead (so every wave still exits); returns false once abort
// is set.
MK_DEV bool barrier_wait(const uint32_t *p, uint32_t target, uint64_t timeout, Counts *c) {
  const uint64_t t0 = wall_clock64();
  while (coherent_load(p) < target) {
    __builtin_amdgcn_s_sleep(2);
    if (wall_clock64() - t0 > timeout) {
      atomicExch(&c->abort, 1u);
      return false;
    }
  }
  return true;
}

MK_DEV bool grid_barrier(BarLines *bl, Counts *c, uint32_t k, uint64_t timeout, uint32_t *flag) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t g = blockIdx.x / kBarGroup;
    const uint32_t ng = (gridDim.x + kBarGroup - 1) / kBarGroup;
    const uint32_t gs = min(kBarGroup, gridDim.x - g * kBarGroup);
    bool ok = true;
    if (atomicAdd(&bl->grp[g][0], 1u) + 1u == k * gs) {
      // last of the group: wait for every group, then release this one
      if (atomicAdd(&bl->top[0], 1u) + 1u < k * ng) ok = barrier_wait(bl->top, k * ng, timeout, c);
      coherent_store(&bl->rel[g][0], k);
    } else {
      ok = barrier_wait(bl->rel[g], k, timeout, c);
    }
    if (coherent_load(&c->abort)) ok = false;
    *flag = ok ? 1u : 0u;
  }
  __syncthreads();
  return *flag != 0u;
}

// Frontier steps j0, j0 + 1, ... < j_end in one launch (all blocks
// resident), a grid barrier between steps: step j reads queue side j & 1,
// counter slot j % 3, as k_expand does.  Stops before the first step whose
// frontier is empty and records its index in cnt->loop_end[MODE] (j_end
// when the range ran out first).  MODE 0 step j uses epochs epoch0 +
// (j - j0) * kEpochStride onwards.  cnt->bar_done counts the grid barriers
// completed since the barrier counters were zeroed (read at the start,
// advanced at the end), so launches queued back to back need no host.
// Housekeeping folded in to save launches: words zero[0..nzero) are
// cleared at the start (no step reads them), and with `clear_exit` the
// frontier counters of every slot and cnt->ninv are cleared at the end
// (after the last step every block reads only the empty slot's counts).
template <bool MAP, bool WEIGHTS, int MODE>
__global__ void __launch_bounds__(kWideBlock)
    k_expand_loop(GraphView G, TravView S, uint32_t j0, uint32_t j_end, uint32_t epoch0,
                  uint64_t timeout, uint32_t *zero, uint32_t nzero, uint32_t clear_exit) {
  if (MODE == 0) epoch0 += *S.ep_base;
  __shared__ uint32_t lds[kLoopLds];
  __shared__ uint32_t flag;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nzero; i += gridDim.x * blockDim.x)
    zero[i] = 0u;
  // written by an earlier launch on this stream; every block reads it
  // before its first barrier arrival, so block 0's update at the end
  // cannot race with it
  const uint32_t bdone = coherent_load(&S.cnt->bar_done);
  if (S.trace && blockIdx.x == 0 && threadIdx.x == 0) S.cnt->loop_t[MODE][0] = wall_clock64();
  uint32_t j = j0;
  bool pulling = false;  // bottom-up steps (static BFS with the in-edge mirror)
  // near-far (static SSSP): the near bound and the number of splits so far,
  // as the launch this one continues left them
  uint32_t T = kInf, nsplit = 0;
  if constexpr (MODE == 0 && WEIGHTS) {
    if (S.delta) {
      nsplit = coherent_load(&S.cnt->nf_split);
      T = coherent_load(&S.cnt->nf_T);
      if (T == 0) T = S.delta;
    }
  }
  for (; j < j_end; ++j) {
    const uint32_t slot = j % 3;
    // every block reads the same settled counts: the uniform stop decision
    const uint32_t ns = coherent_load(&S.cnt->small[slot]);
    const uint64_t nl = coherent_load64(reinterpret_cast<const uint64_t *>(&S.cnt->large[slot]));
    if (S.trace && blockIdx.x == 0 && threadIdx.x == 0 && j - j0 < kTraceSteps) {
      S.cnt->trace_t[MODE][j - j0] = wall_clock64();
      S.cnt->trace_small[MODE][j - j0] = ns;
      S.cnt->trace_items[MODE][j - j0] = uint32_t(nl);
    }
    const uint32_t ep = MODE == 0 ? epoch0 + (j - j0) * kEpochStride : 0u;
    if (ns == 0 && (nl >> 32) == 0) {
      // near frontier empty: done, unless near-far has deferred vertices
      bool split = false;
      if constexpr (MODE == 0 && WEIGHTS) {
        const uint32_t fin = nsplit % 3;
        if (S.delta && coherent_load(&S.cnt->far_cnt[fin]) != 0) {
          const uint32_t fmin = ~coherent_load(&S.cnt->far_minx[fin]);
          const uint32_t base = fmin != kInf && fmin > T ? fmin : T;
          T = base >= kInf - 1 - S.delta ? kInf - 1 : base + S.delta;
          split_step(G, S, j & 1u, slot, ep, T, fin, (nsplit + 1) % 3, (nsplit + 2) % 3, lds);
          ++nsplit;
          split = true;
        }
      }
      if (!split) break;
      if (!grid_barrier(S.bar, S.cnt, bdone + (j - j0) + 1, timeout, &flag)) return;
      continue;
    }
    if constexpr (MODE == 0 && !WEIGHTS) {
      if (S.vbits[0]) {
        // the same counts in every block: the same choice everywhere
        const uint64_t fbuckets = uint64_t(ns) + uint32_t(nl);
        const uint64_t fverts = uint64_t(ns) + (nl >> 32);
        pulling = pulling ? fverts * S.pull_exit >= G.n
                          : fbuckets * S.pull_enter > G.total_buckets;
      }
    }
    if (pulling)
      pull_step<MAP>(G, S, j & 1u, slot, lds);
    else
      expand_step<MAP, WEIGHTS, MODE>(G, S, j & 1u, slot, ep, lds, T, nsplit % 3);
    if (!grid_barrier(S.bar, S.cnt, bdone + (j - j0) + 1, timeout, &flag)) return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (S.trace) {
      const uint32_t k = min(j - j0, kTraceSteps);
      const unsigned long long now = wall_clock64();
      S.cnt->trace_t[MODE][k] = now;
      S.cnt->trace_n[MODE] = k;
      S.cnt->loop_t[MODE][1] = now;
    }
    S.cnt->loop_end[MODE] = j;
    S.cnt->bar_done = bdone + (j - j0);
    if (MODE == 0 && S.delta) {
      S.cnt->nf_T = T;
      S.cnt->nf_split = nsplit;
    }
    if (clear_exit) {
      for (uint32_t k = 0; k < 3; ++k) {
        S.cnt->small[k] = 0u;
        S.cnt->large[k] = 0ull;
      }
      S.cnt->ninv = 0u;
    }
  }
}

// Re-seed invalidated vertices from valid in-neighbours: waves stream over
// all buckets 64 at a time; a bucket's slabs are read only if its owner is
// reached, valid and has live edges.
template <bool MAP, bool WEIGHTS>
__global__ void __launch_bounds__(kBlock)
    k_reseed(GraphView G, TravView S, uint32_t epoch) {
  epoch += *S.ep_base;
  if (S.cnt->steps == 0) return;  // no invalidation step ran: nothing to re-seed
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  __shared__ uint32_t lq[kWaveLds];
  WaveQ q{lq + (threadIdx.x >> 6) * kWaveQ, 0u, 0u};
  uint32_t slabs = 0, edges = 0;
  for (uint64_t b0 = uint64_t(gw) * 64; b0 < G.total_buckets; b0 += uint64_t(W) * 64) {
    const uint64_t b = b0 + lane;
    bool act = b < G.total_buckets;
    uint32_t u = 0, du = kInf;
    if (act) {
      u = G.bucket_owner[b];
      act = G.vdeg[u] != 0;
      if (act) {
        du = dist_of(S.dist[u]);
        act = du != kInf && !inv_bit(S.invbits, u);
      }
    }
    GlobalSink sink{S, q, 0u, 0u};
    batched_buckets<MAP, WEIGHTS, 2>(G, S, sink, epoch, ballot(act), uint32_t(b), u, du, slabs,
                                     edges);
  }
  block_flush(q, S.q_small[0], &S.cnt->small[0], lq + (kBlock / 64) * kWaveQ);
  if (lane == 0) {
    add_stat(S.cnt, kSetReseed, kStatSlabs, slabs);
    add_stat(S.cnt, kSetReseed, kStatEdges, edges);
    add_stat(S.cnt, kSetReseed, kStatPushed, q.pushed);
  }
}

// Minimum of a 64-bit value over the wave (all lanes active).
MK_DEV uint64_t wave_min_u64(uint64_t x) {
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t lo = __shfl_xor(uint32_t(x), off, 64);
    const uint32_t hi = __shfl_xor(uint32_t(x >> 32), off, 64);
    const uint64_t o = (uint64_t(hi) << 32) | lo;
    x = o < x ? o : x;
  }
  return x;
}

// Pull one in-chain of invalidated vertex tv (first slab word xv loaded):
// the wave's minimum of dist[p] + w over valid, reached in-neighbours p is
// applied to tv with one atomicMin.  Wave-collective.
template <bool MAP, bool WEIGHTS>
MK_DEV void pull_chain(const GraphView &G, const GraphView &R, const TravView &S, WaveQ &q,
                       uint32_t epoch, uint32_t tv, uint32_t xv, uint32_t &slabs,
                       uint32_t &edges) {
  const uint32_t lane = lane_id();
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  for (;;) {
    ++slabs;
    const bool live = key_lane && xv < kTomb;
    edges += uint32_t(__popcll(ballot(live)));
    const uint32_t wv = MAP ? __shfl_down(xv, 1, 64) : 1u;
    uint64_t cand = kInfInf;
    if (live && !inv_bit(S.invbits, xv)) {
      const uint32_t dp = dist_of(S.dist[xv]);
      const uint32_t D = dp + (WEIGHTS ? wv : 1u);
      if (dp != kInf && D >= dp) cand = pack(D, xv);
    }
    cand = wave_min_u64(cand);
    bool push = false;
    if (lane == 0 && cand != kInfInf) {
      const uint64_t old = atomicMin((unsigned long long *)(S.dist + tv), (unsigned long long)cand);
      if (dist_of(cand) < dist_of(old)) push = atomicExch(S.stamp + tv, epoch) != epoch;
    }
    enqueue(G, S, q, 0, 0, push, tv);
    const uint32_t nx = read_lane(xv, kNextLane);
    if (nx == kEmpty) break;
    xv = slab_ptr(R.pool, nx)[lane];
  }
}

// pull_chain over the head slabs of c <= kBatch in-buckets at once (slab k
// of invalidated vertex tv[k], this lane's word x[k]), phase by phase so
// that their dependent memory operations overlap: every in-neighbour's
// invalid bit and distance are loaded together (with each target's
// current value), the c wave minima are taken and merged per target, then
// lane k applies slab k's minimum to tv[k] where it is lower (all atomics
// in flight at once) and the improved vertices are pushed.  Chained slabs
// follow one by one.  Wave-collective.
template <bool MAP, bool WEIGHTS>
MK_DEV void batch_pull_work(const GraphView &G, const GraphView &R, const TravView &S, WaveQ &q,
                            uint32_t epoch, uint32_t c, const uint32_t (&tv)[kBatch],
                            const uint32_t (&x)[kBatch], uint32_t &slabs, uint32_t &edges) {
  const uint32_t lane = lane_id();
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  // lane k < c: target tv[k] and its current value (a stale value only
  // lets an extra atomic through: values only fall during the pull)
  uint32_t t = 0;
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k)
    if (lane == k) t = tv[k];
  const uint64_t tcur = lane < c ? S.dist[t] : 0ull;
  bool live[kBatch];
  uint32_t bits[kBatch], wv[kBatch];
  uint64_t dp[kBatch];
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) {
    live[k] = k < c && key_lane && x[k] < kTomb;
    wv[k] = MAP ? __shfl_down(x[k], 1, 64) : 1u;
    bits[k] = 0u;
    dp[k] = kInfInf;
    if (live[k]) {
      bits[k] = S.invbits[x[k] >> 5];
      dp[k] = S.dist[x[k]];
    }
  }
  uint64_t cand[kBatch];
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) {
    cand[k] = kInfInf;
    if (live[k] && !((bits[k] >> (x[k] & 31u)) & 1u)) {
      const uint32_t d = dist_of(dp[k]);
      const uint32_t D = d + (WEIGHTS ? wv[k] : 1u);
      if (d != kInf && D >= d) cand[k] = pack(D, x[k]);
    }
    edges += uint32_t(__popcll(ballot(live[k])));
  }
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) cand[k] = wave_min_u64(cand[k]);
  // slabs of one vertex (a hub's buckets come in order) give one atomic:
  // atomics on one address serialise
#pragma unroll
  for (uint32_t k = kBatch - 1; k > 0; --k)
    if (k < c && tv[k - 1] == tv[k]) {
      cand[k - 1] = cand[k] < cand[k - 1] ? cand[k] : cand[k - 1];
      cand[k] = kInfInf;
    }
  uint64_t mine = kInfInf;
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k)
    if (lane == k) mine = cand[k];
  bool push = false;
  if (mine < tcur) {
    const uint64_t old = atomicMin((unsigned long long *)(S.dist + t), (unsigned long long)mine);
    push = dist_of(mine) < dist_of(old);
  }
  if (push) push = atomicExch(S.stamp + t, epoch) != epoch;
  enqueue(G, S, q, 0, 0, push, t);
  slabs += c;
#pragma unroll
  for (uint32_t k = 0; k < kBatch; ++k) {
    const uint32_t nx = k < c ? read_lane(x[k], kNextLane) : kEmpty;
    if (nx != kEmpty)
      pull_chain<MAP, WEIGHTS>(G, R, S, q, epoch, tv[k], slab_ptr(R.pool, nx)[lane], slabs, edges);
  }
}

// Pull the in-buckets named by lanes of `m` (lane j: vertex tj, in-bucket
// bk), kBatch head slabs at a time.
template <bool MAP, bool WEIGHTS>
MK_DEV void pull_buckets(const GraphView &G, const GraphView &R, const TravView &S, WaveQ &q,
                         uint32_t epoch, uint64_t m, uint32_t bk, uint32_t tj, uint32_t &slabs,
                         uint32_t &edges) {
  const uint32_t lane = lane_id();
  while (m) {
    uint32_t js[kBatch];
    uint32_t c = 0;
#pragma unroll
    for (uint32_t t = 0; t < kBatch; ++t) {
      js[t] = m ? ffs64(m) : 0u;
      if (m) {
        m &= m - 1;
        c = t + 1;
      }
    }
    uint32_t x[kBatch], tv[kBatch];
#pragma unroll
    for (uint32_t t = 0; t < kBatch; ++t) {
      x[t] = t < c ? R.heads[size_t(read_lane(bk, js[t])) * kSlabWords + lane] : kEmpty;
      tv[t] = read_lane(tj, js[t]);
    }
    batch_pull_work<MAP, WEIGHTS>(G, R, S, q, epoch, c, tv, x, slabs, edges);
  }
}

// Pull re-seed with the in-edge mirror R: every invalidated vertex v (a set
// bit of invbits) takes the minimum of dist[p] + w over its valid, reached
// in-neighbours p.  Waves take 64 bitmap words at a time; each set bit
// becomes one lane's vertex.  Vertices with more than kSmallBuckets
// in-buckets are listed (queue side 1 storage, counter cnt->pull_large, as
// bucket work items) for
// k_pull_hubs, which spreads their buckets over all waves.
template <bool MAP, bool WEIGHTS>
__global__ void __launch_bounds__(kBlock)
    k_reseed_pull(GraphView G, GraphView R, TravView S, uint32_t epoch) {
  epoch += *S.ep_base;
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  __shared__ uint32_t lq[kWaveLds];
  WaveQ q{lq + (threadIdx.x >> 6) * kWaveQ, 0u, 0u};
  const uint32_t total = S.cnt->ninv;
  const uint32_t per = per_wave(total, W);
  uint32_t slabs = 0, edges = 0, visited = 0;
  for (uint32_t i0 = gw * per; i0 < total; i0 += W * per) {
    const uint32_t i = i0 + lane;
    bool act = lane < per && i < total;
    uint32_t v = 0, rb0 = 0, rnb = 0;
    if (act) {
      v = S.inv_list[i];
      rb0 = R.bucket_off[v];
      rnb = R.bucket_off[v + 1] - rb0;
    }
    const bool hub = act && rnb > kSmallBuckets;
    push_hubs(R, S.q_large[1], &S.cnt->pull_large, hub, v, rnb);
    if (hub) act = false;
    visited += uint32_t(__popcll(ballot(act)));
    for (uint32_t k = 0; ballot(act && rnb > k); ++k)
      pull_buckets<MAP, WEIGHTS>(G, R, S, q, epoch, ballot(act && rnb > k), rb0 + k, v, slabs,
                                 edges);
  }
  block_flush(q, S.q_small[0], &S.cnt->small[0], lq + (kBlock / 64) * kWaveQ);
  if (lane == 0) {
    add_stat(S.cnt, kSetReseed, kStatSlabs, slabs);
    add_stat(S.cnt, kSetReseed, kStatEdges, edges);
    add_stat(S.cnt, kSetReseed, kStatVisited, visited);
    add_stat(S.cnt, kSetReseed, kStatPushed, q.pushed);
  }
}

// Compact the invalidation bitmap into S.inv_list (count in S.cnt->ninv).
// Each thread scans kInvWords consecutive bitmap words; a block-wide prefix
// sum of their popcounts (wave scans plus one LDS pass) places the entries
// with a single atomicAdd per block.
constexpr uint32_t kInvWords = 8;

__global__ void __launch_bounds__(kBlock) k_inv_compact(TravView S, uint32_t nwords) {
  if (S.cnt->steps == 0) return;  // no invalidation step ran: the bitmap is clear
  __shared__ uint32_t wsum[kBlock / 64 + 1];
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  const uint64_t first = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * kInvWords;
  uint32_t words[kInvWords];
  uint32_t c = 0;
#pragma unroll
  for (uint32_t k = 0; k < kInvWords; ++k) {
    words[k] = first + k < nwords ? S.invbits[first + k] : 0u;
    c += uint32_t(__popc(words[k]));
  }
  uint32_t x = c;  // inclusive prefix sum over the wave
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t k = 0; k < kBlock / 64; ++k) {
      const uint32_t v = wsum[k];
      wsum[k] = t;
      t += v;
    }
    wsum[kBlock / 64] = t ? atomicAdd(&S.cnt->ninv, t) : 0u;
  }
  __syncthreads();
  uint32_t pos = wsum[kBlock / 64] + wsum[w] + x - c;
#pragma unroll
  for (uint32_t k = 0; k < kInvWords; ++k) {
    uint32_t word = words[k];
    while (word) {
      S.inv_list[pos++] = uint32_t(first + k) * 32u + uint32_t(__ffs(word) - 1);
      word &= word - 1;
    }
  }
}

// Pull re-seed of the hub vertices listed by k_reseed_pull: their in-buckets
// are work items spread over all waves (same layout as the expand hub list).
template <bool MAP, bool WEIGHTS>
__global__ void __launch_bounds__(kBlock)
    k_pull_hubs(GraphView G, GraphView R, TravView S, uint32_t epoch) {
  epoch += *S.ep_base;
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t n_large = uint32_t(S.cnt->pull_large >> 32);  // list entries
  if (n_large == 0) return;
  __shared__ uint32_t lq[kWaveLds];
  WaveQ q{lq + (threadIdx.x >> 6) * kWaveQ, 0u, 0u};
  uint32_t slabs = 0, edges = 0, items = 0;
  const uint2 *ql = S.q_large[1];
  // lane l takes in-bucket l % kHubChunk of entry l / kHubChunk
  const uint32_t eper = min(per_wave(n_large, W), 64u / kHubChunk);
  for (uint32_t e0 = gw * eper; e0 < n_large; e0 += W * eper) {
    const uint32_t el = lane / kHubChunk, ei = e0 + el;
    bool act = el < eper && ei < n_large;
    uint32_t v = 0, b = 0;
    if (act) {
      const uint2 ent = ql[ei];
      v = ent.x;
      b = ent.y + lane % kHubChunk;
      act = b < R.bucket_off[v + 1];
    }
    items += uint32_t(__popcll(ballot(act)));
    pull_buckets<MAP, WEIGHTS>(G, R, S, q, epoch, ballot(act), b, v, slabs, edges);
  }
  block_flush(q, S.q_small[0], &S.cnt->small[0], lq + (kBlock / 64) * kWaveQ);
  if (lane == 0) {
    add_stat(S.cnt, kSetHubs, kStatSlabs, slabs);
    add_stat(S.cnt, kSetHubs, kStatEdges, edges);
    add_stat(S.cnt, kSetHubs, kStatVisited, items);
    add_stat(S.cnt, kSetHubs, kStatPushed, q.pushed);
  }
}

// Relax the batch edges (u, v) into side 0 / slot 0, thread per edge.  When
// this batch was the graph's last insert (`status` non-NULL), only the edges
// it applied are relaxed, with the batch's weight: an edge it found present
// was already relaxed by the exact distances before the call, and one it
// did not apply is not in the graph.  Without `status`, every edge is taken
// as present (the reference's protocol, BFSImpl.cuh:270-382), its weight
// searched in the graph (which keeps an existing edge's weight), one
// wave-serial search per edge.
constexpr uint32_t kItemBlock = 1024;  // one item per thread kernels

template <bool MAP, bool WEIGHTS>
__global__ void __launch_bounds__(kItemBlock)
    k_seed_batch(GraphView G, TravView S, const uint32_t *src, const uint32_t *dst,
                 const uint32_t *wts, const uint8_t *status, uint32_t n, uint32_t epoch) {
  epoch += *S.ep_base;
  __shared__ uint32_t sh[kItemBlock / 64 + 1];
  const uint32_t lane = lane_id();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t u = 0, v = 0, wgt = 1;
  uint64_t xu = kInfInf, xv = kInfInf;
  bool act = false, need = false;
  if (i < n) {
    // the edge, its status and batch weight, then both distances, each
    // group of loads issued together
    u = src[i];
    v = dst[i];
    const bool applied = !status || status[i] == MK_EDGE_APPLIED;
    const uint32_t bw = WEIGHTS && wts ? wts[i] : 0u;
    act = applied && u != v && u < G.n && v < G.n;
    if (act) {
      // a stale dist[v] only lets an extra atomic through: values only fall
      xu = S.dist[u];
      xv = S.dist[v];
    }
    act = act && dist_of(xu) != kInf;
    if (act && WEIGHTS) {
      if (status) wgt = bw;  // applied by this batch: its weight (NULL weights: 0)
      else need = true;
    }
  }
  if (WEIGHTS) {
    uint64_t pending = ballot(need);
    while (pending) {
      const uint32_t j = ffs64(pending);
      pending &= pending - 1;
      uint32_t wj = 0;
      const int r = slab_search<MAP>(G, read_lane(u, j), read_lane(v, j), &wj);
      if (lane == j) {
        if (r == 1) wgt = wj;
        else act = false;  // not in the graph: nothing to relax
      }
    }
  }
  bool push = false;
  uint32_t nb = 0;
  if (act) {
    const uint32_t du = dist_of(xu);
    const uint32_t D = du + (WEIGHTS ? wgt : 1u);
    if (D >= du && D < dist_of(xv)) {
      nb = G.bucket_off[v + 1] - G.bucket_off[v];  // alongside the atomic
      const uint64_t old = atomicMin((unsigned long long *)(S.dist + v),
                                     (unsigned long long)pack(D, u));
      if (D < dist_of(old)) push = atomicExch(S.stamp + v, epoch) != epoch;
    }
  }
  block_enqueue<kItemBlock / 64>(G, S, 0, 0, push, v, nb, sh);
}

// Decremental seed: batch edge (u, v) was deleted; if it was v's tree edge,
// v becomes invalid (Invalidate, BFSImpl.cuh:3-15).  With `collect`, such a
// v (an orphan) gets its invalid bit and is appended once to S.inv_list
// instead (count S.cnt->ninv) for k_repair, which looks for another parent
// before invalidating.  With `wts` (a weight update: edge (u, v) now weighs
// wts[i]) a tree edge breaks only if it became heavier, dist[u] + wts[i] >
// dist[v]; one that kept or lowered v's distance is left to the relaxation
// that follows.  (A u invalidated concurrently reaches v through the
// invalidation walk, v being its tree child.)
__global__ void __launch_bounds__(kItemBlock)
    k_invalidate_batch(GraphView G, TravView S, const uint32_t *src, const uint32_t *dst,
                       const uint32_t *wts, uint32_t n, uint32_t source, uint32_t collect) {
  __shared__ uint32_t sh[kItemBlock / 64 + 1];
  const uint32_t lane = lane_id();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool push = false, orphan = false;
  uint32_t v = 0, nb = 0;
  if (i < n) {
    const uint32_t u = src[i];
    v = dst[i];
    const uint32_t wn = wts ? wts[i] : 0u;
    if (u != v && u < G.n && v < G.n && v != source) {
      const uint64_t cur = S.dist[v];
      bool broken = cur != kInfInf && parent_of(cur) == u;
      if (broken && wts) {
        const uint32_t du = dist_of(S.dist[u]);
        broken = du != kInf && uint64_t(du) + wn > uint64_t(dist_of(cur));
      }
      if (broken) {
        if (collect) {
          // the invalid bit marks v as listed: a batch naming (u, v) twice
          // lists v once (the list holds at most n vertices)
          const uint32_t bit = 1u << (v & 31u);
          orphan = (atomicOr(S.invbits + (v >> 5), bit) & bit) == 0u;
        } else {
          nb = G.bucket_off[v + 1] - G.bucket_off[v];  // alongside the CAS
          if (atomicCAS((unsigned long long *)(S.dist + v), (unsigned long long)cur,
                        (unsigned long long)kInfInf) == cur) {
            atomicOr(S.invbits + (v >> 5), 1u << (v & 31u));
            push = true;
          }
        }
      }
    }
  }
  if (collect) {
    const uint64_t om = ballot(orphan);
    if (om) {
      const uint32_t leader = ffs64(om);
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(&S.cnt->ninv, uint32_t(__popcll(om)));
      base = read_lane(base, leader);
      if (orphan) S.inv_list[base + lanes_below(om)] = v;
    }
    return;
  }
  block_enqueue<kItemBlock / 64>(G, S, 0, 0, push, v, nb, sh);
}

// Orphans with at most this many in-buckets are searched for another parent.
constexpr uint32_t kRepairBuckets = 8;

// Re-parent orphans: an orphan v (tree edge deleted) keeps its distance if
// an in-neighbour p (read from the in-edge mirror R) has dist[p] + w(p, v)
// == dist[v]; v's parent becomes p and its subtree is left alone.  This is
// safe even when p is itself invalidated later in the same call: the
// invalidation walk then reaches v as p's child.  Only edges of weight >= 1
// qualify, so dist[p] < dist[v]: p is not in v's subtree, and the new parent
// pointers form no cycle (each repaired pointer strictly lowers the
// distance, tree pointers never raise it).  A repaired orphan's invalid bit
// is cleared; one without such a p keeps it, is set to INF and is pushed as
// k_invalidate_batch does.  One 8-lane group per orphan (each listed once).
template <bool MAP, bool WEIGHTS>
__global__ void __launch_bounds__(kBlock) k_repair(GraphView G, GraphView R, TravView S) {
  __shared__ uint32_t lq[kWaveLds];
  WaveQ q{lq + (threadIdx.x >> 6) * kWaveQ, 0u, 0u};
  const uint32_t gl = group_lane();
  const uint32_t gid = (blockIdx.x * blockDim.x + threadIdx.x) / kGroupLanes;
  const uint32_t ng = (gridDim.x * blockDim.x) / kGroupLanes;
  const uint32_t total = S.cnt->ninv;
  // every group runs the same number of rounds (wave-collective enqueue)
  const uint32_t rounds = (total + ng - 1) / ng;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t k = r * ng + gid;
    bool push = false;
    uint32_t v = 0;
    if (k < total) {
      v = S.inv_list[k];
      // one read, broadcast: the group's decisions must agree
      const uint64_t c0 = S.dist[v];
      const uint64_t cur = (uint64_t(group_read(uint32_t(c0 >> 32), 0)) << 32) |
                           group_read(uint32_t(c0), 0);
      const uint32_t dv = dist_of(cur);
      uint64_t repl = kInfInf;  // new packed value: repaired, or INF
      if (cur != kInfInf) {
        const uint32_t rb0 = R.bucket_off[v], rnb = R.bucket_off[v + 1] - rb0;
        uint32_t found = kEmpty;
        for (uint32_t bk = 0; bk < rnb && rnb <= kRepairBuckets && found == kEmpty; ++bk) {
          const uint32_t *slab = R.heads + size_t(rb0 + bk) * kSlabWords;
          for (;;) {
            uint32_t x[kWordsPerLane];
            load_words(slab + kWordsPerLane * gl, x);
            uint32_t mine = kEmpty;
#pragma unroll
            for (uint32_t j = 0; j < kWordsPerLane; ++j) {
              const uint32_t word = kWordsPerLane * gl + j;
              if (!key_word<MAP>(word) || x[j] >= kTomb || mine != kEmpty) continue;
              const uint32_t p = x[j];
              const uint32_t w = WEIGHTS ? x[j + 1 < kWordsPerLane ? j + 1 : j] : 1u;
              const uint32_t dp = dist_of(S.dist[p]);
              if (w != 0u && p != parent_of(cur) && dp != kInf && uint64_t(dp) + w == dv)
                mine = p;
            }
            const uint32_t hb = group_bits(ballot(mine != kEmpty));
            if (hb) {
              found = group_read(mine, ffs32(hb));
              break;
            }
            const uint32_t nx = group_read(x[kWordsPerLane - 1], kGroupLanes - 1);
            if (nx == kEmpty) break;
            slab = slab_ptr(R.pool, nx);
          }
        }
        if (found != kEmpty) repl = pack(dv, found);
        if (gl == 0) {
          const uint64_t old = atomicCAS((unsigned long long *)(S.dist + v),
                                         (unsigned long long)cur, (unsigned long long)repl);
          if (old == cur && repl == kInfInf)
            push = true;
          else
            atomicAnd(S.invbits + (v >> 5), ~(1u << (v & 31u)));
        }
      }
    }
    enqueue(G, S, q, 0, 0, push, v);
  }
  block_flush(q, S.q_small[0], &S.cnt->small[0], lq + (kBlock / 64) * kWaveQ);
}


// Start a static run: the source at distance 0 is the frontier (side 0,
// slot 0), in the queue its bucket count selects.
// The source alone in the frontier of step 0 (queue side 0, slot 0), and
// marked in `vbits0` (side 0's frontier bitmap) when bottom-up steps may run.
__global__ void k_set_source(GraphView G, uint64_t *dist, uint32_t source, uint32_t *q_small,
                             uint2 *q_large, Counts *cnt, uint32_t *vbits0) {
  dist[source] = pack(0, source);
  if (vbits0) vbits0[source >> 5] |= 1u << (source & 31u);
  const uint32_t b0 = G.bucket_off[source], nb = G.bucket_off[source + 1] - b0;
  if (nb <= kSmallBuckets) {
    q_small[0] = source;
    cnt->small[0] = 1;
  } else {  // hub-list entries, kHubChunk buckets each
    const uint32_t nc = (nb + kHubChunk - 1) / kHubChunk;
    for (uint32_t c = 0; c < nc; ++c) q_large[c] = make_uint2(source, b0 + c * kHubChunk);
    cnt->large[0] = (uint64_t(nc) << 32) | nb;
  }
}

__global__ void k_unpack(const uint64_t *dist, uint32_t n, uint32_t *d, uint32_t *p) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t x = dist[i];
  if (d) d[i] = dist_of(x);
  if (p) p[i] = parent_of(x);
}

// ---------------------------------------------------------------- host
// Streams of the library's handles, per device (BFS/SSSP handles share
// their graph's stream; multi-source, PageRank, TC and CC handles have their
// own).  A frontier loop's grid barrier needs every block resident.  While
// a single stream of the device is registered, nothing else of the
// library's can run beside a loop and hold compute units, and a grid sized
// to what the device keeps resident is enough: loops are launched like any
// kernel.  Once handles of the device run on two streams (another graph's
// traversals, a multi-source query, PageRank, TC or CC work queued from
// another thread), a partly resident loop grid could wait on kernels that
// hold the rest of the device, so loops are launched cooperatively (the
// runtime checks co-residency).
static std::mutex g_loop_mu;
static std::map<int, std::map<hipStream_t, int>> g_loop_streams;  // device -> stream -> handles

void loop_stream_add(int device, hipStream_t s, int d) {
  std::lock_guard<std::mutex> lk(g_loop_mu);
  auto &m = g_loop_streams[device];
  if ((m[s] += d) <= 0) m.erase(s);
}
static bool loop_stream_alone(int device) {
  std::lock_guard<std::mutex> lk(g_loop_mu);
  return g_loop_streams[device].size() <= 1;
}

struct Traversal {
  mk_graph *g = nullptr;
  uint32_t source = 0;
  bool weights = false;  // SSSP
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool defer = false;    // incremental/decremental return once queued (settle() reads them)
  bool pending = false;  // a deferred call's counters are on their way to `host`
  uint32_t pub_seq = 0;  // sequence number of the last counter copy
  uint32_t pend_seq = 0; // the pending call's
  uint64_t *dist = nullptr;
  uint32_t *stamp = nullptr;
  uint32_t *invbits = nullptr;
  uint32_t *inv_list = nullptr;
  uint32_t *qs[2] = {nullptr, nullptr};
  uint2 *ql[2] = {nullptr, nullptr};
  Counts *cnt = nullptr;  // device; BarLines follow at bar_offset
  BarLines *bar = nullptr;
  HostBuf host;           // pinned Counts
  HostBuf snap[2];        // pinned Counts snapshots of successive step groups
  hipEvent_t snap_ev[2] = {nullptr, nullptr};
  uint32_t epoch = 0;
  uint32_t last_iters = 0;
  uint64_t last_slabs = 0;
  uint64_t last_edges = 0;
  int max_waves = 0;
  bool repair = false;  // orphan repair before invalidation (with the in-edge mirror)
  bool persistent = true;        // frontier steps in one launch (k_expand_loop)
  uint32_t tail_enter = 128;     // tail-mode entry threshold (knob "tail_enter")
  uint32_t *vbits = nullptr;     // two frontier bitmaps (static BFS with the mirror)
  uint32_t pull_enter = 0, pull_exit = 1;  // knobs "pull_enter" / "pull_exit"
  bool pull_now = false;         // the launch being queued may take bottom-up steps
  uint2 *far[3] = {nullptr, nullptr, nullptr};  // near-far lists (static SSSP)
  uint32_t far_cap = 0;
  uint32_t delta = 0;            // knob "sssp_delta" (0: no near-far)
  bool last_far_over = false;    // the last call's far list overflowed on a hub
  uint32_t *ep_base = nullptr;   // device word added to stamp epochs (captured steps)
  uint32_t *capture_sticky = nullptr;  // recording a step graph: its outcome word
  uint32_t steps = 0;            // captured steps (mk_step) built on this handle
  bool delta_now = false;        // the launch being queued is a static SSSP's
  bool coop = false;             // loops always launched cooperatively (knob "coop_launch")
  unsigned loop_blocks = 0;      // resident-grid size of both loop kernels
  uint64_t barrier_timeout = 0;  // wall-clock ticks a grid-barrier wait may take
  int clock_khz = 100000;        // device wall-clock rate

  TravView view() const {
    TravView v;
    v.dist = dist;
    v.stamp = stamp;
    v.invbits = invbits;
    v.inv_list = inv_list;
    v.q_small[0] = qs[0];
    v.q_small[1] = qs[1];
    v.q_large[0] = ql[0];
    v.q_large[1] = ql[1];
    v.cnt = cnt;
    v.bar = bar;
    v.tail_enter = tail_enter;
    v.trace = profiling_on() ? 1u : 0u;
    v.tail_stay = std::min<uint32_t>(tail_enter + tail_enter / 2, kTailCap);
    const bool pull = pull_now && vbits && g->rev;
    v.R = pull ? g->rev->view() : GraphView{};
    v.vbits[0] = pull ? vbits : nullptr;
    v.vbits[1] = pull ? vbits + bit_words() : nullptr;
    v.pull_enter = pull_enter;
    v.pull_exit = pull_exit;
    for (int i = 0; i < 3; ++i) v.far[i] = far[i];
    v.far_cap = far_cap;
    v.delta = (delta_now && far[0]) ? delta : 0u;
    v.ep_base = ep_base;
    return v;
  }
  uint32_t bit_words() const { return (g->n + 31) / 32; }

  int init(mk_graph *graph, uint32_t src, bool w) {
    g = graph;
    source = src;
    weights = w;
    const size_t n = g->n;
    MK_HIP(hipSetDevice(g->device));
    // the graph's own stream: the graph's mutations given no stream run
    // there too, so a call that follows one is ordered after it as is
    stream = g->own_stream;
    MK_HIP(hipEventCreate(&ev0));
    MK_HIP(hipEventCreate(&ev1));
    MK_HIP(hipMalloc(&dist, n * 8));
    MK_HIP(hipMalloc(&stamp, n * 4));
    MK_HIP(hipMalloc(&invbits, ((n + 31) / 32) * 4));
    MK_HIP(hipMalloc(&inv_list, n * 4));
    for (int i = 0; i < 2; ++i) {
      MK_HIP(hipMalloc(&qs[i], n * 4));
      // one entry per kHubChunk buckets of every listed hub (a vertex is
      // listed once per step; the pull re-seed lists in-buckets)
      const size_t buckets = std::max<size_t>(g->total_buckets, g->rev ? g->rev->total_buckets : 0);
      MK_HIP(hipMalloc(&ql[i], (n + buckets / kHubChunk + 64) * 8));
    }
    MK_HIP(hipMalloc(&cnt, bar_offset() + sizeof(BarLines)));
    MK_HIP(hipMalloc(&ep_base, 64));
    fill_u32(ep_base, 0u, 16, stream);
    bar = reinterpret_cast<BarLines *>(reinterpret_cast<char *>(cnt) + bar_offset());
    if (int r = host.ensure(sizeof(Counts) + 64)) return r;  // + the sequence word
    for (int i = 0; i < 2; ++i) {
      if (int r = snap[i].ensure(sizeof(Counts))) return r;
      MK_HIP(hipEventCreateWithFlags(&snap_ev[i], hipEventDisableTiming));
    }
    fill_u64(dist, kInfInf, n, stream);
    fill_u32(stamp, 0u, n, stream);
    fill_u32(invbits, 0u, (n + 31) / 32, stream);
    clear_counts();
    k_set_source<<<1, 1, 0, stream>>>(g->view(), dist, source, qs[0], ql[0], cnt, nullptr);
    MK_HIP(hipStreamSynchronize(stream));
    max_waves = grid_waves(g->device);
    repair = tune_get("orphan_repair", 0) != 0;
    persistent = tune_get("persistent_steps", 1) != 0;
    tail_enter = uint32_t(tune_get("tail_enter", 128));
    pull_enter = uint32_t(tune_get("pull_enter", 14));
    delta = uint32_t(tune_get("sssp_delta", 0));
    if (weights && delta) {
      // three far lists of up to n entries each (near-far static SSSP)
      // knob "sssp_far_cap" (0: n entries) bounds each list
      const int64_t cap = tune_get("sssp_far_cap", 0);
      far_cap = uint32_t(std::min<size_t>(cap > 0 ? size_t(cap) : n, 0xFFFFFFF0ull));
      for (int i = 0; i < 3; ++i) MK_HIP(hipMalloc(&far[i], size_t(far_cap) * 8));
    }
    pull_exit = uint32_t(std::max<int64_t>(tune_get("pull_exit", 24), 1));
    if (!weights && g->rev && pull_enter) {
      MK_HIP(hipMalloc(&vbits, size_t(bit_words()) * 8));
      if (int r = mirror_flags_enable(g->rev)) return r;
    }
    coop = tune_get("coop_launch", 0) != 0;
    loop_stream_add(g->device, stream, 1);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess ||
        khz <= 0)
      khz = 100000;
    clock_khz = khz;
    // knob "barrier_timeout_ms" (default five seconds)
    barrier_timeout = uint64_t(khz) * uint64_t(tune_get("barrier_timeout_ms", 5000));
    if (persistent) {
      // one grid for both loop kernels: the barrier lines count arrivals
      // cumulatively across the loops of one call (cnt->bar_done), so every
      // loop between two clears must arrive with the same number of blocks
      unsigned b0 = 0, b1 = 0;
      if (int r = loop_grid<0>(b0)) return r;
      if (int r = loop_grid<1>(b1)) return r;
      loop_blocks = std::min(b0, b1);
    }
    return MK_OK;
  }

  void destroy() {
    (void)hipSetDevice(g->device);
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(dist);
    (void)hipFree(stamp);
    (void)hipFree(invbits);
    (void)hipFree(inv_list);
    (void)hipFree(vbits);
    for (int i = 0; i < 3; ++i) (void)hipFree(far[i]);
    for (int i = 0; i < 2; ++i) {
      (void)hipFree(qs[i]);
      (void)hipFree(ql[i]);
    }
    (void)hipFree(cnt);
    (void)hipFree(ep_base);
    host.release();
    for (int i = 0; i < 2; ++i) {
      snap[i].release();
      if (snap_ev[i]) (void)hipEventDestroy(snap_ev[i]);
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) loop_stream_add(g->device, stream, -1);
    stream = nullptr;  // the graph's
  }

  // Epoch base of the next launch: kEpochStride values are reserved per
  // launch (a tail-mode launch uses one per inner step).  Stamps start at 0,
  // so epochs are never 0; on wrap-around the stamps are cleared.
  uint32_t next_epoch() { return reserve_epochs(1); }

  // Base of `steps` consecutive epoch ranges of kEpochStride values each.
  uint32_t reserve_epochs(uint32_t steps) {
    const uint64_t need = uint64_t(steps + 1) * kEpochStride;
    if (uint64_t(epoch) + need > 0xFFFFFFFFull) {
      fill_u32(stamp, 0u, g->n, stream);
      epoch = 0;
    }
    const uint32_t base = epoch + kEpochStride;
    epoch = base + (steps - 1) * kEpochStride;
    return base;
  }

  template <bool MAP, bool WEIGHTS, int MODE> static const void *loop_kernel() {
    return reinterpret_cast<const void *>(&k_expand_loop<MAP, WEIGHTS, MODE>);
  }
  template <int MODE> const void *loop_kernel_for() const {
    return g->weighted ? (weights ? loop_kernel<true, true, MODE>()
                                  : loop_kernel<true, false, MODE>())
                       : loop_kernel<false, false, MODE>();
  }

  // Blocks of a k_expand_loop grid: max_waves in blocks of kWideBlock
  // threads, capped at what the device keeps resident at once (a grid
  // barrier needs every block running).  One block slot per CU is kept in
  // reserve where there are several: the occupancy figure of 7 blocks of
  // 256 threads per CU given for the unweighted relax loop (67 VGPRs) was
  // measured not to be resident under an ordinary launch (its first barrier
  // timed out), while 6 per CU ran for every variant.
  template <int MODE> int loop_grid(unsigned &blocks) {
    int per_cu = 0;
    MK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, loop_kernel_for<MODE>(),
                                                        kWideBlock, 0));
    if (per_cu > 1) --per_cu;
    const uint64_t resident = uint64_t(per_cu > 0 ? per_cu : 0) * uint64_t(device_cus(g->device));
    if (resident == 0) return fail(MK_ERR_HIP, "frontier loop kernel cannot be resident");
    blocks = unsigned(std::min<uint64_t>(
        std::min<uint64_t>((uint64_t(max_waves) + kWideBlock / 64 - 1) / (kWideBlock / 64),
                           resident),
        uint64_t(kMaxBarGroups) * kBarGroup));  // the barrier's group lines
    return MK_OK;
  }

  // One cooperative launch of k_expand_loop over steps [j0, j1).
  template <int MODE>
  int launch_loop(uint32_t j0, uint32_t j1, uint32_t ep, uint32_t *zero = nullptr,
                  uint32_t nzero = 0, uint32_t clear_exit = 0) {
    GraphView G = g->view();
    TravView S = view();
    uint64_t to = barrier_timeout;
    void *args[] = {&G, &S, &j0, &j1, &ep, &to, &zero, &nzero, &clear_exit};
    // timed (while profiling) by the kernel's own clock readings: a
    // cooperative launch takes no dispatch events, and events around it
    // would include the runtime's extra dispatch work
    if (coop || !loop_stream_alone(g->device))
      MK_HIP(hipLaunchCooperativeKernel(loop_kernel_for<MODE>(), dim3(loop_blocks),
                                        dim3(kWideBlock), args, 0u, stream));
    else  // the only stream running loops; the grid is sized to be resident
      MK_HIP(hipLaunchKernel(loop_kernel_for<MODE>(), dim3(loop_blocks), dim3(kWideBlock), args,
                             0u, stream));
    return MK_OK;
  }

  // Zero every counter (statistics and grid-barrier lines included).
  void clear_counts() {
    fill_u32(reinterpret_cast<uint32_t *>(cnt), 0u, (bar_offset() + sizeof(BarLines)) / 4, stream);
  }
  static constexpr size_t bar_offset() { return (sizeof(Counts) + 127) / 128 * 128; }
  // Zero the frontier slots only.
  void clear_slots() {
    fill_u32(cnt->small, 0u, 3, stream);
    fill_u64(reinterpret_cast<uint64_t *>(cnt->large), 0ull, 3, stream);
  }

  // Queue the counters' copy into the pinned host copy (k_publish_counts);
  // returns the sequence number it ends with.
  uint32_t publish_counts() {
    if (++pub_seq == 0) pub_seq = 1;  // the word starts at 0
    k_publish_counts<<<1, 256, 0, stream>>>(cnt, static_cast<Counts *>(host.dev),
                                            profiling_on() ? 1u : 0u, pub_seq);
    return pub_seq;
  }
  const volatile uint32_t *pub_word() const {
    return reinterpret_cast<const volatile uint32_t *>(host.as<Counts>() + 1);
  }
  int read_counts(Totals &h) {
    publish_counts();
    MK_HIP(hipStreamSynchronize(stream));
    parse(*host.as<Counts>(), h);
    return MK_OK;
  }
  // The counter copy, then the stop event of a timed call.  The caller
  // synchronises the stream before parsing `host`.
  int copy_counts_then_stop() {
    publish_counts();
    MK_HIP(hipEventRecord(ev1, stream));
    return MK_OK;
  }

  static void parse(const Counts &c, Totals &h) {
    for (int i = 0; i < 3; ++i) {
      h.small[i] = c.small[i];
      h.large[i] = c.large[i];
    }
    h.steps = c.steps;
    h.abort = c.abort;
    h.far_over = c.far_over;
    h.nf_split = c.nf_split;
    h.loop_end[0] = c.loop_end[0];
    h.loop_end[1] = c.loop_end[1];
    for (uint32_t s = 0; s < kSets; ++s)
      for (uint32_t w = 0; w < 4; ++w) {
        unsigned long long x = 0;
        for (uint32_t k = 0; k < kStatSlots; ++k) x += c.stats[k][s * 4 + w];
        h.set[s][w] = x;
      }
    h.slabs = h.edges = h.visited = h.pushed = 0;
    for (uint32_t s = 0; s < kSets; ++s) {
      h.slabs += h.set[s][kStatSlabs];
      h.edges += h.set[s][kStatEdges];
      h.visited += h.set[s][kStatVisited];
      h.pushed += h.set[s][kStatPushed];
    }
  }

  template <int MODE> void launch_expand(uint32_t side, uint32_t slot, uint32_t ep) {
    const unsigned grid = blocks_for_waves(uint64_t(max_waves));
    const GraphView G = g->view();
    const TravView S = view();
    const char *name = MODE == 0 ? "k_expand_relax" : "k_expand_invalidate";
    auto k = g->weighted ? (weights ? k_expand<true, true, MODE> : k_expand<true, false, MODE>)
                         : k_expand<false, false, MODE>;
    prof_launch(name, 0.0, k, dim3(grid), dim3(kBlock), 0u, stream, G, S, side, slot, ep);
  }

  // Persistent frontier loops (k_expand_loop) are queued without waiting
  // for the device; end() checks them with the call's one host wait.  An
  // invalidation loop runs until its frontier drains (it uses no epochs); a
  // relax loop gets kLoopSteps steps' epochs and end() queues a further
  // launch in the rare case that it used up its range.
  static constexpr uint32_t kLoopSteps = 256;
  uint32_t relax_end = 0;     // range end of the queued relax loop (0: none)
  bool inval_queued = false;  // an invalidation loop was queued
  bool dirty = true;          // counters may be non-zero (cleared at the next begin)
  bool stale = false;         // results computed on a batch that failed (until Static)
  int reseed_kind = 0;        // queued re-seed: 0 none, 1 pull (mirror), 2 bucket scan
  uint64_t deferred_settled = 0;    // deferred calls settled since create
  uint64_t deferred_recomputed = 0; // of those, recomputed from scratch (ran out of steps)

  uint32_t inval_limit() const {
    return uint32_t(std::min<uint64_t>(uint64_t(g->n) + 8, 0xFFFFFFF0ull));
  }
  // `zero`: words the launch clears first (the invalidation bitmap after a
  // re-seed).  The invalidation loop leaves the frontier counters and ninv
  // cleared for the re-seed.
  int queue_relax(uint32_t *zero = nullptr, uint32_t nzero = 0) {
    relax_end = kLoopSteps;
    return launch_loop<0>(0, kLoopSteps, reserve_epochs(kLoopSteps), zero, nzero);
  }
  int queue_invalidate() {
    inval_queued = true;
    return launch_loop<1>(0, inval_limit(), 0u, nullptr, 0u, 1u);
  }

  // Step trace of the loops of this call (the host copy of the counters).
  // Duration of the last k_expand_loop<mode> launch (host copy of the
  // counters), added to the profiler's totals.
  void record_loop_time(int mode) {
    const Counts &c = *host.as<Counts>();
    const double ms = double(c.loop_t[mode][1] - c.loop_t[mode][0]) / double(clock_khz);
    prof_time(mode == 0 ? "k_expand_relax" : "k_expand_invalidate", ms);
  }

  void record_steps() {
    const Counts &c = *host.as<Counts>();
    const double us_per_tick = 1000.0 / double(clock_khz);
    for (int mode = 0; mode < 2; ++mode) {
      if (mode == 0 ? !relax_end : !inval_queued) continue;
      record_loop_time(mode);
      const uint32_t k = std::min(c.trace_n[mode], kTraceSteps);
      for (uint32_t i = 0; i < k; ++i)
        prof_step(mode == 0 ? "k_expand_relax" : "k_expand_invalidate", i, c.trace_small[mode][i],
                  c.trace_items[mode][i],
                  double(c.trace_t[mode][i + 1] - c.trace_t[mode][i]) * us_per_tick);
    }
  }

  // end()'s part for queued loops: counters `h` were just read.  Algorithmic
  // bytes (profiling) per phase from its statistics set, as in iterate and
  // run_decremental.
  int finish_loops(Totals &h) {
    if (h.abort) return fail(MK_ERR_STATE, "frontier loop: a grid barrier wait timed out");
    if (inval_queued && h.loop_end[1] >= inval_limit())
      return fail(MK_ERR_STATE, "frontier did not drain within n steps");
    while (relax_end && h.loop_end[0] == relax_end) {
      if (uint64_t(relax_end) > uint64_t(g->n) + 8)
        return fail(MK_ERR_STATE, "frontier did not drain within n steps");
      if (profiling_on()) record_loop_time(0);
      const uint32_t j0 = relax_end;
      relax_end += kLoopSteps;
      if (int r = launch_loop<0>(j0, relax_end, reserve_epochs(kLoopSteps))) return r;
      if (int r = copy_counts_then_stop()) return r;
      MK_HIP(hipStreamSynchronize(stream));
      parse(*host.as<Counts>(), h);
      if (h.abort) return fail(MK_ERR_STATE, "frontier loop: a grid barrier wait timed out");
    }
    if (profiling_on()) {
      auto steps_bytes = [&](int set) {
        return 256.0 * double(h.set[set][kStatSlabs]) + 8.0 * double(h.set[set][kStatEdges]) +
               24.0 * double(h.set[set][kStatVisited]) + 16.0 * double(h.set[set][kStatPushed]);
      };
      if (relax_end) prof_add_bytes("k_expand_relax", steps_bytes(kSetRelax));
      if (inval_queued) prof_add_bytes("k_expand_invalidate", steps_bytes(kSetInvalidate));
      const unsigned long long *r = h.set[kSetReseed];
      if (reseed_kind == 1) {
        // 12 B per listed vertex (entry, in-bucket range), 256 B per
        // in-slab, 4 + 8 B per in-neighbour (invalid bit, distance), 16 B
        // per re-seeded vertex; the hubs' in-buckets likewise, 12 B per
        // bucket work item
        prof_add_bytes("k_reseed_pull", 12.0 * double(r[kStatVisited]) +
                                            256.0 * double(r[kStatSlabs]) +
                                            12.0 * double(r[kStatEdges]) +
                                            16.0 * double(r[kStatPushed]));
        const unsigned long long *hb = h.set[kSetHubs];
        prof_add_bytes("k_pull_hubs", 12.0 * double(hb[kStatVisited]) +
                                          256.0 * double(hb[kStatSlabs]) +
                                          12.0 * double(hb[kStatEdges]) +
                                          16.0 * double(hb[kStatPushed]));
      }
      else if (reseed_kind == 2 && h.set[kSetInvalidate][kStatVisited])
        // 20 B per bucket (owner, degree, distance, invalid bit of the
        // owner), 256 B per slab read, 4 B per key (invalid bit), 16 B per
        // re-seeded vertex
        prof_add_bytes("k_reseed", 20.0 * double(g->total_buckets) + 256.0 * double(r[kStatSlabs]) +
                                       4.0 * double(r[kStatEdges]) + 16.0 * double(r[kStatPushed]));
    }
    return MK_OK;
  }

  // Run frontier steps from queue side 0 / slot 0 until the frontier is
  // empty.  Steps are launched in groups without waiting; after each group
  // the counters are copied to a pinned snapshot.  The host queues the next
  // group before it waits for the previous group's snapshot, so the device
  // keeps working through the host's round trip; groups launched after the
  // frontier emptied are no-op steps.  Algorithmic bytes (profiling) per
  // group: 256 B per slab, 8 B per key (neighbour distance), 24 B per
  // expanded vertex (entry, bucket range, distance), 16 B per push (atomic,
  // stamp, entry).
  template <int MODE> int iterate(Totals &h) {
    uint32_t side = 0, slot = 0;
    uint64_t launched = 0;
    const uint64_t limit = uint64_t(g->n) + 8;
    Totals prev;
    if (profiling_all()) {
      if (int r = read_counts(prev)) return r;
    }
    uint32_t slot_after[2] = {0, 0};
    auto group = [&](uint32_t k, int b) -> int {
      for (uint32_t i = 0; i < k; ++i) {
        launch_expand<MODE>(side, slot, MODE == 0 ? next_epoch() : 0u);
        side ^= 1u;
        slot = slot == 2 ? 0 : slot + 1;
        ++launched;
      }
      MK_HIP(hipGetLastError());
      MK_HIP(hipMemcpyAsync(snap[b].ptr, cnt, sizeof(Counts), hipMemcpyDeviceToHost, stream));
      MK_HIP(hipEventRecord(snap_ev[b], stream));
      slot_after[b] = slot;
      return MK_OK;
    };
    if (int r = group(2, 0)) return r;
    uint32_t k = 2;
    for (int b = 0;; b ^= 1) {
      if (launched > limit) return fail(MK_ERR_STATE, "frontier did not drain within n steps");
      k = std::min<uint32_t>(k * 2, 4);
      if (int r = group(k, b ^ 1)) return r;  // queued before waiting
      MK_HIP(hipEventSynchronize(snap_ev[b]));
      parse(*snap[b].as<Counts>(), h);
      if (profiling_all()) {
        const double bytes = 256.0 * double(h.slabs - prev.slabs) +
                             8.0 * double(h.edges - prev.edges) +
                             24.0 * double(h.visited - prev.visited) +
                             16.0 * double(h.pushed - prev.pushed);
        prof_add_bytes(MODE == 0 ? "k_expand_relax" : "k_expand_invalidate", bytes);
        prev = h;
      }
      const uint32_t sl = slot_after[b];
      if (h.small[sl] == 0 && (h.large[sl] >> 32) == 0) break;
    }
    return MK_OK;
  }

  // Counters are zero between calls: end() queues their clearing after it
  // read them (off the next call's path); a call that stopped early leaves
  // them `dirty` for the next begin() to clear.
  // `timed`: the call reports its time (ev0 .. ev1); an event record between
  // kernels was measured to leave ~10 us of device idle.
  int begin(bool timed = true) {
    MK_HIP(hipSetDevice(g->device));
    if (capture_sticky) {
      // recorded into a captured update step: counters start clear
      relax_end = 0;
      inval_queued = false;
      reseed_kind = 0;
      clear_counts();
      dirty = true;
      return MK_OK;
    }
    if (pending) return fail(MK_ERR_STATE, "internal: a deferred call was not settled");
    // a deferred mutation of the graph is ordered before this call's
    // kernels on the device; its outcome is settled in end()
    if (g->pend.active && g->pend.s != stream) MK_HIP(hipStreamWaitEvent(stream, g->mut_ev, 0));
    relax_end = 0;
    inval_queued = false;
    reseed_kind = 0;
    if (dirty) clear_counts();
    dirty = true;
    if (timed) MK_HIP(hipEventRecord(ev0, stream));
    return MK_OK;
  }

  // End of a call: its counters (and outcome) now, or, for a deferred
  // incremental/decremental call (`may_defer`), once the copy is queued:
  // settle() reads them at the handle's next call, read or sync.
  int end(float *ms, bool may_defer = false) {
    if (capture_sticky) {
      // recorded into a captured update step: its outcome flags go to the
      // step's sticky word, the counters are cleared for the next call
      k_step_check<<<1, 1, 0, stream>>>(cnt, relax_end, inval_queued ? inval_limit() : 0u,
                                        capture_sticky);
      clear_counts();
      dirty = false;
      relax_end = 0;
      inval_queued = false;
      reseed_kind = 0;
      if (ms) *ms = 0.f;
      return MK_OK;
    }
    if (may_defer && defer) {
      pend_seq = publish_counts();  // no events: the host waits for pend_seq
      int lr = MK_OK;
      if (g->pend.active) {
        // the mutation this call ran on (queued before it) is settled now:
        // the host waits for it while the device runs this call
        lr = graph_settle(g);
        if (lr) stale = true;
      }
      clear_counts();  // after the copy on the stream
      dirty = false;
      pending = true;
      if (ms) *ms = 0.f;  // not measured: the call has not run yet
      return lr;
    }
    if (int r = copy_counts_then_stop()) return r;
    MK_HIP(hipStreamSynchronize(stream));
    Totals h;
    parse(*host.as<Counts>(), h);
    int lr = finish_loops(h);
    if (lr) stale = true;  // a loop gave up: the results are incomplete
    if (!lr && profiling_on()) record_steps();
    if (!lr && g->pend.active) {
      // the mutation this call ran on failed: the graph is restored, and
      // these results (computed on the failed batch) are not valid
      lr = graph_settle(g);
      if (lr) stale = true;
    }
    relax_end = 0;
    inval_queued = false;
    reseed_kind = 0;
    if (lr) return lr;
    clear_counts();
    dirty = false;
    float t = 0.f;
    MK_HIP(hipEventElapsedTime(&t, ev0, ev1));
    if (ms) *ms = t;
    last_slabs = h.slabs;
    last_edges = h.edges;
    last_iters = h.steps;
    last_far_over = h.far_over != 0;
    return MK_OK;
  }

  // Read a deferred call's counters (waiting for it) and report its
  // outcome once: a loop that gave up leaves the handle stale; a relax loop
  // that ran out of its steps cannot be continued in place (later work on
  // the graph may already have run), so the results are recomputed from
  // scratch, which is exact for the graph as it is now.
  int settle() {
    if (!pending) return MK_OK;
    pending = false;
    if (int r = host_wait_word(pub_word(), pend_seq, stream)) {
      stale = true;
      return r;
    }
    Totals h;
    parse(*host.as<Counts>(), h);
    const bool ran_out = relax_end && h.loop_end[0] == relax_end;
    const int lr = ran_out ? MK_OK : finish_loops(h);  // no relaunch unless ran_out
    if (!lr && !ran_out && profiling_on()) record_steps();
    relax_end = 0;
    inval_queued = false;
    reseed_kind = 0;
    if (lr) {
      stale = true;
      return lr;
    }
    last_slabs = h.slabs;
    last_edges = h.edges;
    last_iters = h.steps;
    ++deferred_settled;
    if (ran_out) {
      ++deferred_recomputed;
      return run_static(nullptr);
    }
    return MK_OK;
  }

  int set_deferred(bool on) {
    if (!on) {
      if (int r = settle()) {
        defer = false;
        return r;
      }
    }
    defer = on;
    return MK_OK;
  }

  int run_static(float *ms, bool near_far = true) {
    if (int r = settle()) return r;
    if (int r = begin()) return r;
    stale = false;
    fill_u64(dist, kInfInf, g->n, stream);
    // a BFS over a graph with the in-edge mirror may take bottom-up steps
    pull_now = persistent && vbits != nullptr;
    if (pull_now) fill_u32(vbits, 0u, uint64_t(bit_words()) * 2, stream);
    // an SSSP with the knob sssp_delta set runs near-far (a loop that runs
    // out of steps is continued near-far too: end() relaunches it)
    delta_now = near_far && persistent && far[0] != nullptr;
    k_set_source<<<1, 1, 0, stream>>>(g->view(), dist, source, qs[0], ql[0], cnt,
                                      pull_now ? vbits : nullptr);
    if (persistent) {
      const int r = queue_relax();
      pull_now = false;
      if (r) {
        delta_now = false;
        return r;
      }
    } else {
      Totals h;
      if (int r = iterate<0>(h)) return r;
    }
    const bool nf = delta_now;
    const int r = end(ms);
    delta_now = false;
    // a deferred hub that found its far list full left buckets unexpanded:
    // the call is repeated without near-far
    if (r == MK_OK && nf && last_far_over) return run_static(ms, false);
    return r;
  }

  int run_incremental(const uint32_t *src, const uint32_t *dst, const uint32_t *w, uint32_t n,
                      float *ms) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "incremental: NULL batch");
    if (int r = settle()) return r;
    if (int r = graph_failed_batch(g, src, dst, n)) {
      stale = true;
      return r;
    }
    if (stale) return fail(MK_ERR_STATE, "incremental: results are invalid after a failed batch; "
                                         "run Static() first");
    if (int r = begin(!defer)) return r;
    if (n) {
      const uint32_t ep = next_epoch();
      const GraphView G = g->view();
      // the insert's per-edge status applies when this is the batch it saw
      // (SSSP: with the same weights, which it then takes from the batch)
      const bool same = g->last_op == 1 && g->last_src == src && g->last_dst == dst &&
                        g->status_n == n && (!weights || g->last_w == w);
      const uint8_t *st = same ? g->status.as<uint8_t>() : nullptr;
      {
        const unsigned grid = unsigned((uint64_t(n) + kItemBlock - 1) / kItemBlock);
        auto k = weights ? k_seed_batch<true, true>
                         : (g->weighted ? k_seed_batch<true, false> : k_seed_batch<false, false>);
        prof_launch("k_seed_batch", 24.0 * n, k, dim3(grid), dim3(kItemBlock), 0u, stream, G,
                    view(), src, dst, weights ? w : nullptr, st, n, ep);
      }
      MK_HIP(hipGetLastError());
      if (persistent) {
        if (int r = queue_relax()) return r;
      } else {
        Totals h;
        if (int r = iterate<0>(h)) return r;
      }
    }
    return end(ms, true);
  }

  // Decremental after the batch seeds, persistent loops: the invalidation
  // loop, the re-seed and the relax loop are all queued, then end() waits
  // once.  The re-seed kernels return at once when nothing was invalidated
  // (no invalidation step ran), as the host-driven path skips them.
  // A weight update (`rw`: the batch and its new weights) also relaxes
  // every batch edge from its source before the relax loop, into the same
  // frontier as the re-seeded vertices.
  struct Reweight {
    const uint32_t *src, *dst;
    uint32_t n;
  };
  void seed_reweighted(const Reweight &rw, uint32_t ep) {
    // weights are read from the graph (status NULL), which holds the
    // update's outcome: a batch naming an edge twice keeps one weight
    const unsigned grid = unsigned((uint64_t(rw.n) + kItemBlock - 1) / kItemBlock);
    prof_launch("k_seed_batch", 24.0 * rw.n, k_seed_batch<true, true>, dim3(grid),
                dim3(kItemBlock), 0u, stream, g->view(), view(), rw.src, rw.dst,
                static_cast<const uint32_t *>(nullptr), static_cast<const uint8_t *>(nullptr),
                rw.n, ep);
  }

  int queue_decremental(float *ms, const Reweight *rw = nullptr) {
    const GraphView G = g->view();
    if (int r = queue_invalidate()) return r;
    const uint32_t ep = next_epoch();
    const uint32_t nwords = uint32_t((uint64_t(g->n) + 31) / 32);
    if (g->rev) {
      const GraphView R = g->rev->view();
      k_inv_compact<<<blocks_for_threads((uint64_t(nwords) + kInvWords - 1) / kInvWords), kBlock,
                      0, stream>>>(view(), nwords);
      const unsigned grid = blocks_for_waves(uint64_t(max_waves));
      auto kp = g->weighted ? (weights ? k_reseed_pull<true, true> : k_reseed_pull<true, false>)
                            : k_reseed_pull<false, false>;
      auto kh = g->weighted ? (weights ? k_pull_hubs<true, true> : k_pull_hubs<true, false>)
                            : k_pull_hubs<false, false>;
      prof_launch("k_reseed_pull", 0.0, kp, dim3(grid), dim3(kBlock), 0u, stream, G, R, view(), ep);
      prof_launch("k_pull_hubs", 0.0, kh, dim3(grid), dim3(kBlock), 0u, stream, G, R, view(), ep);
      reseed_kind = 1;
    } else {
      const unsigned grid = blocks_for_waves(std::min<uint64_t>(
          (uint64_t(g->total_buckets) + 63) / 64, uint64_t(max_waves) * 2));
      auto k = g->weighted ? (weights ? k_reseed<true, true> : k_reseed<true, false>)
                           : k_reseed<false, false>;
      prof_launch("k_reseed", 0.0, k, dim3(grid), dim3(kBlock), 0u, stream, G, view(), ep);
      reseed_kind = 2;
    }
    if (rw) seed_reweighted(*rw, ep);
    MK_HIP(hipGetLastError());
    if (int r = queue_relax(invbits, nwords)) return r;
    return end(ms, true);
  }

  // Decremental (after DeleteEdges with the batch), or with `wts` a weight
  // update (after mk_graph_update_weights with the batch): tree edges that
  // became heavier invalidate their subtrees, which are re-seeded, and
  // every batch edge is relaxed with its weight in the graph.
  int run_decremental(const uint32_t *src, const uint32_t *dst, uint32_t n, float *ms,
                      const uint32_t *wts = nullptr) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "decremental: NULL batch");
    if (n && wts && !weights) return fail(MK_ERR_BAD_ARG, "weight update: not an SSSP handle");
    if (int r = settle()) return r;
    if (int r = graph_failed_batch(g, src, dst, n)) {
      stale = true;
      return r;
    }
    if (stale) return fail(MK_ERR_STATE, "decremental: results are invalid after a failed batch; "
                                         "run Static() first");
    if (int r = begin(!defer)) return r;
    if (n) {
      const GraphView G = g->view();
      {
        const unsigned grid = unsigned((uint64_t(n) + kItemBlock - 1) / kItemBlock);
        // with the in-edge mirror, orphans are first offered another parent
        const uint32_t collect = g->rev && repair ? 1u : 0u;
        prof_launch("k_invalidate_batch", (wts ? 36.0 : 24.0) * n, k_invalidate_batch,
                    dim3(grid), dim3(kItemBlock), 0u, stream, G, view(), src, dst, wts, n, source,
                    collect);
        if (collect) {
          auto k = g->weighted ? (weights ? k_repair<true, true> : k_repair<true, false>)
                               : k_repair<false, false>;
          prof_launch("k_repair", 0.0, k, dim3(blocks_for_waves(uint64_t(max_waves))),
                      dim3(kBlock), 0u, stream, G, g->rev->view(), view());
        }
      }
      MK_HIP(hipGetLastError());
      const Reweight rw{src, dst, n};
      if (persistent) return queue_decremental(ms, wts ? &rw : nullptr);
      // propagate invalidation down the tree; any step ran = something was
      // invalidated
      Totals h;
      if (int r = iterate<1>(h)) return r;
      if (h.steps == 0) {
        if (wts) {  // nothing invalidated: only the reweighted edges to relax
          clear_slots();
          seed_reweighted(rw, next_epoch());
          MK_HIP(hipGetLastError());
          if (int r = iterate<0>(h)) return r;
        }
      } else {
        // re-seed from valid in-neighbours; the re-seeded vertices start a
        // fresh frontier at side 0 / slot 0
        const uint32_t ep = next_epoch();
        clear_slots();
        const uint64_t slabs0 = h.slabs, edges0 = h.edges, pushed0 = h.pushed;
        if (g->rev) {
          // pull from in-neighbours of the invalidated vertices
          const GraphView R = g->rev->view();
          const uint32_t nwords = uint32_t((uint64_t(g->n) + 31) / 32);
          fill_u32(&cnt->ninv, 0u, 1, stream);
          k_inv_compact<<<blocks_for_threads((uint64_t(nwords) + kInvWords - 1) / kInvWords),
                          kBlock, 0, stream>>>(view(), nwords);
          const unsigned grid = blocks_for_waves(uint64_t(max_waves));
          const unsigned hgrid = blocks_for_waves(uint64_t(max_waves));
          auto kp = g->weighted ? (weights ? k_reseed_pull<true, true> : k_reseed_pull<true, false>)
                                : k_reseed_pull<false, false>;
          auto kh = g->weighted ? (weights ? k_pull_hubs<true, true> : k_pull_hubs<true, false>)
                                : k_pull_hubs<false, false>;
          prof_launch("k_reseed_pull", 0.0, kp, dim3(grid), dim3(kBlock), 0u, stream, G, R, view(),
                      ep);
          // hub half of the pull
          prof_launch("k_pull_hubs", 0.0, kh, dim3(hgrid), dim3(kBlock), 0u, stream, G, R, view(),
                      ep);
        } else {
          const unsigned grid = blocks_for_waves(std::min<uint64_t>(
              (uint64_t(g->total_buckets) + 63) / 64, uint64_t(max_waves) * 2));
          ProfScope prof(stream, "k_reseed", 0.0);
          if (g->weighted) {
            if (weights)
              k_reseed<true, true><<<grid, kBlock, 0, stream>>>(G, view(), ep);
            else
              k_reseed<true, false><<<grid, kBlock, 0, stream>>>(G, view(), ep);
          } else {
            k_reseed<false, false><<<grid, kBlock, 0, stream>>>(G, view(), ep);
          }
        }
        MK_HIP(hipGetLastError());
        if (profiling_all() && g->rev) {
          // 4 B per bitmap word, 8 B per invalidated vertex (in-bucket
          // range), 256 B per in-slab, 4 + 8 B per in-neighbour (invalid bit,
          // distance), 16 B per re-seeded vertex
          Totals a;
          if (int r = read_counts(a)) return r;
          prof_add_bytes("k_reseed_pull", 4.0 * double((g->n + 31) / 32) +
                                              8.0 * double(a.visited - h.visited) +
                                              256.0 * double(a.slabs - slabs0) +
                                              12.0 * double(a.edges - edges0) +
                                              16.0 * double(a.pushed - pushed0));
        } else if (profiling_all()) {
          // 20 B per bucket (owner, degree, distance, invalid bit of the
          // owner), 256 B per slab read, 4 B per key (invalid bit), 16 B per
          // re-seeded vertex
          Totals a;
          if (int r = read_counts(a)) return r;
          prof_add_bytes("k_reseed", 20.0 * double(g->total_buckets) +
                                         256.0 * double(a.slabs - slabs0) +
                                         4.0 * double(a.edges - edges0) +
                                         16.0 * double(a.pushed - pushed0));
        }
        fill_u32(invbits, 0u, (uint64_t(g->n) + 31) / 32, stream);
        if (wts) seed_reweighted(rw, ep);
        MK_HIP(hipGetLastError());
        if (int r = iterate<0>(h)) return r;
      }
    }
    return end(ms, true);
  }

  int read(uint32_t *d, uint32_t *p) {
    MK_HIP(hipSetDevice(g->device));
    if (int r = settle()) return r;
    DevBuf b;
    if (int r = b.ensure(size_t(g->n) * 8)) return r;
    const int r = unpack_to_host(b.as<uint32_t>(), d, p);
    b.release();
    return r;
  }

  // Distances and parents unpacked into `staging` (2n words of device
  // memory) on the handle's stream, then copied to the host.
  int unpack_to_host(uint32_t *staging, uint32_t *d, uint32_t *p) {
    const size_t n = g->n;
    uint32_t *dd = staging, *pp = staging + n;
    k_unpack<<<blocks_for_threads(n), kBlock, 0, stream>>>(dist, g->n, dd, pp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && d) e = hipMemcpyAsync(d, dd, n * 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess && p) e = hipMemcpyAsync(p, pp, n * 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "read distances");
    return MK_OK;
  }
};


}  // namespace mk

struct mk_bfs {
  mk::Traversal t;
};
struct mk_sssp {
  mk::Traversal t;
};

using namespace mk;

template <typename H> static int trav_create(mk_graph *g, uint32_t source, bool w, H **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "create: out is NULL");
  *out = nullptr;
  if (!g) return fail(MK_ERR_BAD_ARG, "create: NULL graph");
  if (source >= g->n) return fail(MK_ERR_RANGE, "create: source vertex >= n");
  if (w && !g->weighted) return fail(MK_ERR_BAD_ARG, "SSSP needs a weighted (map) graph");
  H *h = new H();
  int r = h->t.init(g, source, w);
  if (r) {
    h->t.destroy();
    delete h;
    return r;
  }
  g->users++;
  *out = h;
  return MK_OK;
}

template <typename H> static int trav_destroy(H *h) {
  if (!h) return fail(MK_ERR_BAD_ARG, "destroy: NULL handle");
  if (h->t.steps) return fail(MK_ERR_STATE, "destroy: a captured step still uses the handle");
  h->t.g->users--;
  h->t.destroy();
  delete h;
  return MK_OK;
}

extern "C" {

int mk_bfs_create(mk_graph *g, uint32_t source, mk_bfs **out) {
  return trav_create(g, source, false, out);
}
int mk_bfs_destroy(mk_bfs *b) { return trav_destroy(b); }
int mk_bfs_static(mk_bfs *b, float *ms) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_static: NULL handle");
  return b->t.run_static(ms);
}
int mk_bfs_incremental(mk_bfs *b, const uint32_t *src, const uint32_t *dst, uint32_t n,
                       float *ms) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_incremental: NULL handle");
  return b->t.run_incremental(src, dst, nullptr, n, ms);
}
int mk_bfs_decremental(mk_bfs *b, const uint32_t *src, const uint32_t *dst, uint32_t n,
                       float *ms) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_decremental: NULL handle");
  return b->t.run_decremental(src, dst, n, ms);
}
int mk_bfs_set_source(mk_bfs *b, uint32_t source) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_set_source: NULL handle");
  if (source >= b->t.g->n) return fail(MK_ERR_RANGE, "mk_bfs_set_source: source >= n");
  if (int r = b->t.settle()) return r;
  b->t.source = source;
  return MK_OK;
}
int mk_sssp_set_source(mk_sssp *s, uint32_t source) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_set_source: NULL handle");
  if (source >= s->t.g->n) return fail(MK_ERR_RANGE, "mk_sssp_set_source: source >= n");
  if (int r = s->t.settle()) return r;
  s->t.source = source;
  return MK_OK;
}
int mk_bfs_read(mk_bfs *b, uint32_t *d, uint32_t *p) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_read: NULL handle");
  return b->t.read(d, p);
}
int mk_bfs_last_stats(mk_bfs *b, uint32_t *iters, uint64_t *slabs, uint64_t *edges) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_last_stats: NULL handle");
  if (int r = b->t.settle()) return r;
  if (iters) *iters = b->t.last_iters;
  if (slabs) *slabs = b->t.last_slabs;
  if (edges) *edges = b->t.last_edges;
  return MK_OK;
}

int mk_sssp_create(mk_graph *g, uint32_t source, mk_sssp **out) {
  return trav_create(g, source, true, out);
}
int mk_sssp_destroy(mk_sssp *s) { return trav_destroy(s); }
int mk_sssp_static(mk_sssp *s, float *ms) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_static: NULL handle");
  return s->t.run_static(ms);
}
int mk_sssp_incremental(mk_sssp *s, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                        uint32_t n, float *ms) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_incremental: NULL handle");
  return s->t.run_incremental(src, dst, w, n, ms);
}
int mk_sssp_decremental(mk_sssp *s, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                        uint32_t n, float *ms) {
  (void)w;
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_decremental: NULL handle");
  return s->t.run_decremental(src, dst, n, ms);
}
int mk_sssp_reweight(mk_sssp *s, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                     uint32_t n, float *ms) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_reweight: NULL handle");
  if (n && !w) return fail(MK_ERR_BAD_ARG, "mk_sssp_reweight: NULL weights");
  return s->t.run_decremental(src, dst, n, ms, w);
}
int mk_sssp_read(mk_sssp *s, uint32_t *d, uint32_t *p) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_read: NULL handle");
  return s->t.read(d, p);
}
int mk_sssp_last_stats(mk_sssp *s, uint32_t *iters, uint64_t *slabs, uint64_t *edges) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_last_stats: NULL handle");
  if (int r = s->t.settle()) return r;
  if (iters) *iters = s->t.last_iters;
  if (slabs) *slabs = s->t.last_slabs;
  if (edges) *edges = s->t.last_edges;
  return MK_OK;
}

int mk_bfs_set_deferred(mk_bfs *b, int on) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_set_deferred: NULL handle");
  return b->t.set_deferred(on != 0);
}
int mk_bfs_sync(mk_bfs *b) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_sync: NULL handle");
  return b->t.settle();
}
int mk_bfs_deferred_stats(mk_bfs *b, uint64_t *settled, uint64_t *recomputed) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_deferred_stats: NULL handle");
  if (settled) *settled = b->t.deferred_settled;
  if (recomputed) *recomputed = b->t.deferred_recomputed;
  return MK_OK;
}
int mk_sssp_deferred_stats(mk_sssp *s, uint64_t *settled, uint64_t *recomputed) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_deferred_stats: NULL handle");
  if (settled) *settled = s->t.deferred_settled;
  if (recomputed) *recomputed = s->t.deferred_recomputed;
  return MK_OK;
}
int mk_sssp_set_deferred(mk_sssp *s, int on) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_set_deferred: NULL handle");
  return s->t.set_deferred(on != 0);
}
int mk_sssp_sync(mk_sssp *s) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_sync: NULL handle");
  return s->t.settle();
}

}  // extern "C"

// ------------------------------------------------- captured update steps
// One update step of a BFS/SSSP handle, DeleteEdges -> Decremental ->
// InsertEdges -> Incremental on fixed batch buffers, captured once into a
// hipGraph and replayed (no reference counterpart: the reference launches
// every kernel from the host, src/SSSP.cu:100-125).  The replays' kernels
// take their stamp and mutation epochs relative to two device words (the
// handle's ep_base and the graph's counters[1]); the host sets them before
// the first replay of a run of replays and the graph's last node advances
// them by one replay's epochs, so replays queued back to back never reuse an
// epoch and need no host work between them.  Outcomes are collected on the
// device: the mutation counters keep the epoch of the last failed batch, and
// each call ORs its loop flags into a sticky word (k_step_check).
struct mk_step {
  mk::Traversal *t = nullptr;
  const uint32_t *ds = nullptr, *dd = nullptr, *is = nullptr, *id = nullptr, *iw = nullptr;
  uint32_t nd = 0, ni = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t *sticky = nullptr;     // device: OR of the calls' outcome flags
  uint32_t mut0 = 0;              // the delete's captured mutation epoch (the insert's is +1)
  uint32_t e0 = 0, span = 0;      // captured stamp epochs lie in (e0, e0 + span + stride)
  uint32_t replays = 0;           // queued since the bases were last set
  uint32_t mut_first = 0;         // effective mutation epoch of the first of them
  bool bases_set = false;
};

namespace mk {
static int step_capture(mk_step *st) {
  Traversal &t = *st->t;
  mk_graph *g = t.g;
  hipStream_t s = t.stream;
  MK_HIP(hipSetDevice(g->device));
  // room for one capture's epochs, outside the capture (nothing wraps in it)
  const uint64_t room = uint64_t(4) * (Traversal::kLoopSteps + 4) * kEpochStride;
  if (uint64_t(t.epoch) + room > 0xFFFFFFFFull) {
    fill_u32(t.stamp, 0u, g->n, s);
    t.epoch = 0;
  }
  if (g->mut_epoch > 0xFFFFFF00u) {
    fill_u32(g->counters, 0u, 4, s);
    g->mut_epoch = 0;
  }
  if (int r = g->status.ensure(std::max<size_t>(std::max(st->nd, st->ni), 1))) return r;
  if (g->rev)
    if (int r = g->rev->status.ensure(std::max<size_t>(std::max(st->nd, st->ni), 1))) return r;
  if (int r = g->host_scratch.ensure(64)) return r;
  MK_HIP(hipMalloc(&st->sticky, 64));
  fill_u32(st->sticky, 0u, 16, s);
  MK_HIP(hipStreamSynchronize(s));
  st->e0 = t.epoch;
  MK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  t.capture_sticky = st->sticky;
  uint32_t mi = 0;
  int r = graph_queue_mutation(g, 2, st->ds, st->dd, nullptr, st->nd, nullptr, &st->mut0);
  if (!r) r = t.run_decremental(st->ds, st->dd, st->nd, nullptr);
  if (!r) r = graph_queue_mutation(g, 1, st->is, st->id, st->iw, st->ni, nullptr, &mi);
  if (!r) r = t.run_incremental(st->is, st->id, st->iw, st->ni, nullptr);
  t.capture_sticky = nullptr;
  st->span = t.epoch - st->e0;
  // the last node: the next replay's bases
  k_step_bases<<<1, 1, 0, s>>>(g->counters + 1, t.ep_base, 2u, st->span, 1u);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (ec != hipSuccess) return hip_fail(ec, "mk_*_capture_step: stream capture");
  if (mi != st->mut0 + 1) return fail(MK_ERR_STATE, "mk_*_capture_step: mutation epochs");
  st->graph = graph;
  MK_HIP(hipGraphInstantiate(&st->exec, graph, nullptr, nullptr, 0));
  // nothing ran: the handle and graph hold what they held before capture
  t.dirty = true;
  g->last_op = 0;
  return MK_OK;
}

// Queue one replay; the bases are set first when the previous replays were
// settled (or none ran), else the last replay's final node has advanced them.
static int step_launch(mk_step *st) {
  Traversal &t = *st->t;
  mk_graph *g = t.g;
  hipStream_t s = t.stream;
  MK_HIP(hipSetDevice(g->device));
  if (int r = t.settle()) return r;
  if (int r = graph_settle(g)) return r;
  if (t.stale) return fail(MK_ERR_STATE, "mk_step_launch: results are invalid; run Static() first");
  const uint64_t need = uint64_t(st->span) + 2 * kEpochStride;
  const bool wrap_t = uint64_t(t.epoch) + need > 0xFFFFFFFFull;
  const bool wrap_m = g->mut_epoch > 0xFFFFFF00u;
  if (st->bases_set && (wrap_t || wrap_m)) {
    // an epoch counter would wrap: settle the queued replays first
    if (int r = mk_step_sync(st)) return r;
  }
  if (!st->bases_set) {
    if (uint64_t(t.epoch) + need > 0xFFFFFFFFull) {
      fill_u32(t.stamp, 0u, g->n, s);
      t.epoch = 0;
    }
    if (g->mut_epoch > 0xFFFFFF00u) {
      fill_u32(g->counters, 0u, 4, s);
      g->mut_epoch = 0;
    }
    // replay k's epochs: captured + base (unsigned arithmetic wraps)
    k_step_bases<<<1, 1, 0, s>>>(g->counters + 1, t.ep_base, g->mut_epoch + 1 - st->mut0,
                                 t.epoch - st->e0, 0u);
    MK_HIP(hipGetLastError());
    st->bases_set = true;
    st->replays = 0;
    st->mut_first = g->mut_epoch + 1;
  }
  MK_HIP(hipGraphLaunch(st->exec, s));
  t.epoch += st->span;
  g->mut_epoch += 2;
  ++st->replays;
  t.dirty = false;  // each call of the step clears its counters
  g->last_op = 0;
  return MK_OK;
}
}  // namespace mk

extern "C" {

// Wait for the queued replays and report their outcome: MK_ERR_RANGE /
// MK_ERR_POOL_FULL if a batch failed (the handle then needs Static()),
// MK_ERR_STATE if a loop gave up; a relax loop that ran out of its steps is
// recomputed from scratch (exact for the graph as it is).  The bases go back
// to zero for the calls made through the API.
int mk_step_sync(mk_step *st) {
  if (!st) return fail(MK_ERR_BAD_ARG, "mk_step_sync: NULL step");
  Traversal &t = *st->t;
  mk_graph *g = t.g;
  hipStream_t s = t.stream;
  MK_HIP(hipSetDevice(g->device));
  if (!st->bases_set) return MK_OK;
  uint32_t words[8] = {0};
  MK_HIP(hipMemcpyAsync(words, g->counters, 16, hipMemcpyDeviceToHost, s));
  MK_HIP(hipMemcpyAsync(words + 4, st->sticky, 4, hipMemcpyDeviceToHost, s));
  MK_HIP(hipStreamSynchronize(s));
  k_step_bases<<<1, 1, 0, s>>>(g->counters + 1, t.ep_base, 0u, 0u, 0u);
  fill_u32(st->sticky, 0u, 1, s);
  MK_HIP(hipGetLastError());
  const uint32_t first = st->mut_first, count = 2 * st->replays;
  st->bases_set = false;
  st->replays = 0;
  auto in_run = [&](uint32_t e) { return uint32_t(e - first) < count; };
  // a failed batch: its epoch is the last one recorded in the counters
  if (in_run(words[0])) {
    t.stale = true;
    const bool last = words[0] == first + count - 1 || words[0] == first + count - 2;
    return fail(MK_ERR_RANGE, std::string("captured step: a batch had a vertex id >= n (") +
                                  (last ? "the last replay's" : "an earlier replay's") +
                                  "); that batch changed nothing; run Static()");
  }
  if (in_run(words[3])) {
    t.stale = true;
    std::string msg = "captured step: the slab pool ran out inserting a batch";
    if (words[3] == first + count - 1) {
      // the last replay's insert: its statuses are the graph's last batch's
      const uint32_t w4[4] = {0u, words[1], words[2], words[3]};
      const int r = graph_replay_outcome(g, 1, st->is, st->id, st->ni, w4, words[3]);
      return r ? r : fail(MK_ERR_POOL_FULL, msg + "; rolled back; run Static()");
    }
    return fail(MK_ERR_POOL_FULL, msg + " in an earlier replay, which stays partly applied");
  }
  const uint32_t flags = words[4];
  if (flags & 5u) {
    t.stale = true;
    return fail(MK_ERR_STATE, (flags & 1u) ? "captured step: a grid-barrier wait timed out"
                                           : "captured step: the frontier did not drain");
  }
  if (flags & 2u) return t.run_static(nullptr);  // a relax loop ran out of its steps
  return MK_OK;
}

static int step_create(mk::Traversal *t, const uint32_t *ds, const uint32_t *dd, uint32_t nd,
                       const uint32_t *is, const uint32_t *id, const uint32_t *iw, uint32_t ni,
                       mk_step **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "capture_step: out is NULL");
  *out = nullptr;
  if (!ds || !dd || !is || !id || nd == 0 || ni == 0)
    return fail(MK_ERR_BAD_ARG, "capture_step: both batches must be non-empty device buffers");
  if (!t->persistent) return fail(MK_ERR_STATE, "capture_step: needs persistent_steps = 1");
  if (profiling_on()) return fail(MK_ERR_STATE, "capture_step: profiling must be off");
  if (int r = t->settle()) return r;
  if (int r = graph_settle(t->g)) return r;
  if (t->stale) return fail(MK_ERR_STATE, "capture_step: results are invalid; run Static() first");
  if (t->g->steps) return fail(MK_ERR_STATE, "capture_step: the graph has a captured step already");
  mk_step *st = new mk_step();
  st->t = t;
  st->ds = ds;
  st->dd = dd;
  st->nd = nd;
  st->is = is;
  st->id = id;
  st->iw = iw;
  st->ni = ni;
  if (int r = step_capture(st)) {
    (void)hipFree(st->sticky);
    delete st;
    return r;
  }
  t->g->steps++;
  t->steps++;
  *out = st;
  return MK_OK;
}

int mk_sssp_capture_step(mk_sssp *s, const uint32_t *del_src, const uint32_t *del_dst,
                         uint32_t n_del, const uint32_t *ins_src, const uint32_t *ins_dst,
                         const uint32_t *ins_w, uint32_t n_ins, mk_step **out) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_capture_step: NULL handle");
  return step_create(&s->t, del_src, del_dst, n_del, ins_src, ins_dst, ins_w, n_ins, out);
}

int mk_bfs_capture_step(mk_bfs *b, const uint32_t *del_src, const uint32_t *del_dst,
                        uint32_t n_del, const uint32_t *ins_src, const uint32_t *ins_dst,
                        uint32_t n_ins, mk_step **out) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_capture_step: NULL handle");
  return step_create(&b->t, del_src, del_dst, n_del, ins_src, ins_dst, nullptr, n_ins, out);
}

int mk_step_launch(mk_step *st) {
  if (!st) return fail(MK_ERR_BAD_ARG, "mk_step_launch: NULL step");
  return step_launch(st);
}

int mk_step_load(mk_step *st, const uint32_t *del_src, const uint32_t *del_dst,
                 const uint32_t *ins_src, const uint32_t *ins_dst, const uint32_t *ins_w) {
  if (!st || !del_src || !del_dst || !ins_src || !ins_dst)
    return fail(MK_ERR_BAD_ARG, "mk_step_load: NULL argument");
  if (ins_w && !st->iw) return fail(MK_ERR_BAD_ARG, "mk_step_load: the step has no weights");
  hipStream_t s = st->t->stream;
  MK_HIP(hipSetDevice(st->t->g->device));
  const size_t bd = size_t(st->nd) * 4, bi = size_t(st->ni) * 4;
  MK_HIP(hipMemcpyAsync(const_cast<uint32_t *>(st->ds), del_src, bd, hipMemcpyDeviceToDevice, s));
  MK_HIP(hipMemcpyAsync(const_cast<uint32_t *>(st->dd), del_dst, bd, hipMemcpyDeviceToDevice, s));
  MK_HIP(hipMemcpyAsync(const_cast<uint32_t *>(st->is), ins_src, bi, hipMemcpyDeviceToDevice, s));
  MK_HIP(hipMemcpyAsync(const_cast<uint32_t *>(st->id), ins_dst, bi, hipMemcpyDeviceToDevice, s));
  if (ins_w)
    MK_HIP(hipMemcpyAsync(const_cast<uint32_t *>(st->iw), ins_w, bi, hipMemcpyDeviceToDevice, s));
  return MK_OK;
}

int mk_step_destroy(mk_step *st) {
  if (!st) return fail(MK_ERR_BAD_ARG, "mk_step_destroy: NULL step");
  const int r = mk_step_sync(st);
  (void)hipSetDevice(st->t->g->device);
  if (st->exec) (void)hipGraphExecDestroy(st->exec);
  if (st->graph) (void)hipGraphDestroy(st->graph);
  (void)hipFree(st->sticky);
  st->t->g->steps--;
  st->t->steps--;
  delete st;
  return r;
}

}  // extern "C"
