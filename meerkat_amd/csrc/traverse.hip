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
//  - The frontier holds vertices, not edges.  Vertices with at most
//    kSmallBuckets buckets are expanded by one wave each; larger ones are
//    queued with a reserved range of "bucket work items" (one 64-bit
//    atomicAdd returns the entry index and the range start together, so
//    entries are sorted by range start) and expanded bucket-by-bucket by all
//    waves, which balances hub vertices.
//  - Duplicate pushes within an iteration are dropped with a per-vertex
//    epoch stamp instead of the reference's Queued flags, so nothing is
//    reset between iterations.
//  - Decremental invalidation walks the shortest-path tree downward from the
//    deleted tree edges (children of x are out-neighbours whose packed parent
//    is x), O(edges of invalidated vertices), instead of every vertex walking
//    its ancestor chain.  Re-seeding scans every bucket once for edges from a
//    valid vertex into an invalidated one, as the reference's frontier build
//    does (BFSImpl.cuh:406-440), filtered through an L2-resident bitmap.
//  - Every frontier buffer is sized for n vertices once, at create time
//    (the reference sizes them from the batch and can overflow).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "profile.hpp"
#include "slabhash.hpp"

namespace mk {

constexpr uint32_t kSmallBuckets = 2;
constexpr uint32_t kLargeChunk = 8;  // bucket work items per wave step

struct TravView {
  uint64_t *dist;
  uint32_t *stamp;
  uint32_t *invbits;
  uint32_t *q_small[2];
  uint2 *q_large[2];
  uint32_t *cnt_small;             // [2]
  unsigned long long *cnt_large;   // [2]: (entries << 32) | bucket items
  unsigned long long *slabs_read;  // statistics: slabs and live keys read
  unsigned long long *edges_read;
};

// Append v to the next frontier `nx` (lanes with `push`).
MK_DEV void enqueue(const GraphView &G, const TravView &S, uint32_t nx, bool push, uint32_t v) {
  uint32_t nb = 0;
  if (push) nb = G.bucket_off[v + 1] - G.bucket_off[v];
  const bool small = push && nb <= kSmallBuckets;
  wave_push(small, v, S.q_small[nx], S.cnt_small + nx);
  if (push && !small) {
    const unsigned long long r = atomicAdd(S.cnt_large + nx, (1ull << 32) | nb);
    S.q_large[nx][uint32_t(r >> 32)] = make_uint2(v, uint32_t(r));
  }
}

// Relax every edge of one bucket chain of u (distance du).
template <bool MAP, bool WEIGHTS>
MK_DEV uint32_t relax_bucket(const GraphView &G, const TravView &S, uint32_t nx, uint32_t epoch,
                             uint32_t u, uint32_t du, uint32_t b, uint32_t &edges) {
  const uint32_t lane = lane_id();
  const uint64_t km = key_mask<MAP>();
  const bool key_lane = (km >> lane) & 1ull;
  const uint32_t *slab = G.heads + size_t(b) * kSlabWords;
  uint32_t slabs = 0;
  for (;;) {
    const uint32_t v = slab[lane];
    ++slabs;
    edges += uint32_t(__popcll(ballot(key_lane && v < kTomb)));
    const uint32_t wv = MAP ? __shfl_down(v, 1, 64) : 1u;
    bool push = false;
    if (key_lane && v < kTomb) {
      const uint32_t D = du + (WEIGHTS ? wv : 1u);
      if (D >= du && D < dist_of(S.dist[v])) {
        const uint64_t old = atomicMin((unsigned long long *)(S.dist + v),
                                       (unsigned long long)pack(D, u));
        if (D < dist_of(old)) push = atomicExch(S.stamp + v, epoch) != epoch;
      }
    }
    enqueue(G, S, nx, push, v);
    const uint32_t nxt = read_lane(v, kNextLane);
    if (nxt == kEmpty) break;
    slab = slab_ptr(G.pool, nxt);
  }
  return slabs;
}

// Invalidate the children of invalidated vertex x found in one bucket chain:
// out-neighbours whose packed parent is x.
template <bool MAP>
MK_DEV uint32_t invalidate_bucket(const GraphView &G, const TravView &S, uint32_t nx, uint32_t x,
                                  uint32_t b, uint32_t &edges) {
  const uint32_t lane = lane_id();
  const uint64_t km = key_mask<MAP>();
  const bool key_lane = (km >> lane) & 1ull;
  const uint32_t *slab = G.heads + size_t(b) * kSlabWords;
  uint32_t slabs = 0;
  for (;;) {
    const uint32_t v = slab[lane];
    ++slabs;
    edges += uint32_t(__popcll(ballot(key_lane && v < kTomb)));
    bool push = false;
    if (key_lane && v < kTomb) {
      const uint64_t cur = coherent_load64(S.dist + v);
      if (cur != kInfInf && parent_of(cur) == x) {
        if (atomicCAS((unsigned long long *)(S.dist + v), (unsigned long long)cur,
                      (unsigned long long)kInfInf) == cur) {
          atomicOr(S.invbits + (v >> 5), 1u << (v & 31u));
          push = true;
        }
      }
    }
    enqueue(G, S, nx, push, v);
    const uint32_t nxt = read_lane(v, kNextLane);
    if (nxt == kEmpty) break;
    slab = slab_ptr(G.pool, nxt);
  }
  return slabs;
}

// One frontier step.  MODE 0 relaxes out-edges of every frontier vertex;
// MODE 1 propagates invalidation to tree children.
template <bool MAP, bool WEIGHTS, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_expand(GraphView G, TravView S, uint32_t c, uint32_t epoch, uint32_t n_small,
             uint32_t n_large, uint32_t large_items) {
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t nx = c ^ 1u;
  // side c's counters were read by the host before launch; clear them now so
  // they are ready for the step after this one
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.cnt_small[c] = 0;
    S.cnt_large[c] = 0ull;
  }
  uint32_t slabs = 0, edges = 0;
  const uint2 *ql = S.q_large[c];
  for (uint32_t t0 = gw * kLargeChunk; t0 < large_items; t0 += W * kLargeChunk) {
    uint32_t lo = 0, hi = n_large;  // last entry with start <= t0
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ql[mid].y <= t0) lo = mid;
      else hi = mid;
    }
    uint32_t e = lo;
    const uint32_t t1 = min(t0 + kLargeChunk, large_items);
    for (uint32_t t = t0; t < t1; ++t) {
      while (e + 1 < n_large && ql[e + 1].y <= t) ++e;
      const uint2 ent = ql[e];
      const uint32_t u = ent.x;
      const uint32_t b = G.bucket_off[u] + (t - ent.y);
      if (MODE == 0) {
        const uint32_t du = dist_of(coherent_load64(S.dist + u));
        if (du != kInf) slabs += relax_bucket<MAP, WEIGHTS>(G, S, nx, epoch, u, du, b, edges);
      } else {
        slabs += invalidate_bucket<MAP>(G, S, nx, u, b, edges);
      }
    }
  }
  const uint32_t *qs = S.q_small[c];
  for (uint32_t i = gw; i < n_small; i += W) {
    const uint32_t u = qs[i];
    const uint32_t b0 = G.bucket_off[u], b1 = G.bucket_off[u + 1];
    if (MODE == 0) {
      const uint32_t du = dist_of(coherent_load64(S.dist + u));
      if (du == kInf) continue;
      for (uint32_t b = b0; b < b1; ++b)
        slabs += relax_bucket<MAP, WEIGHTS>(G, S, nx, epoch, u, du, b, edges);
    } else {
      for (uint32_t b = b0; b < b1; ++b) slabs += invalidate_bucket<MAP>(G, S, nx, u, b, edges);
    }
  }
  if (lane_id() == 0 && slabs) {
    atomicAdd(S.slabs_read, (unsigned long long)slabs);
    atomicAdd(S.edges_read, (unsigned long long)edges);
  }
}

// Relax the batch edges (u, v): thread per edge for BFS.
__global__ void __launch_bounds__(kBlock)
    k_seed_batch_unit(GraphView G, TravView S, const uint32_t *src, const uint32_t *dst,
                      uint32_t n, uint32_t nx, uint32_t epoch) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool push = false;
  uint32_t v = 0;
  if (i < n) {
    const uint32_t u = src[i];
    v = dst[i];
    if (u != v && u < G.n && v < G.n) {
      const uint32_t du = dist_of(S.dist[u]);
      const uint32_t D = du + 1u;
      if (du != kInf && D < dist_of(S.dist[v])) {
        const uint64_t old = atomicMin((unsigned long long *)(S.dist + v),
                                       (unsigned long long)pack(D, u));
        if (D < dist_of(old)) push = atomicExch(S.stamp + v, epoch) != epoch;
      }
    }
  }
  enqueue(G, S, nx, push, v);
}

// Relax the batch edges with the weight stored in the graph (the graph is
// authoritative if the batch repeated an existing edge): wave-serial search.
__global__ void __launch_bounds__(kBlock)
    k_seed_batch_weighted(GraphView G, TravView S, const uint32_t *src, const uint32_t *dst,
                          uint32_t n, uint32_t epw, uint32_t nx, uint32_t epoch) {
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t base = uint64_t(wave) * epw;
  if (base >= n) return;
  const uint32_t cnt = uint32_t(min<uint64_t>(epw, n - base));
  uint32_t u = 0, v = 0, wgt = 0;
  bool act = false;
  if (lane < cnt) {
    u = src[base + lane];
    v = dst[base + lane];
    act = u != v && u < G.n && v < G.n && dist_of(S.dist[u]) != kInf;
  }
  bool have = false;
  uint64_t pending = ballot(act);
  while (pending) {
    const uint32_t j = ffs64(pending);
    pending &= pending - 1;
    uint32_t wj = 0;
    const int r = slab_search<true>(G, read_lane(u, j), read_lane(v, j), &wj);
    if (lane == j) {
      have = r == 1;
      wgt = wj;
    }
  }
  bool push = false;
  if (have) {
    const uint32_t du = dist_of(S.dist[u]);
    const uint32_t D = du + wgt;
    if (D >= du && D < dist_of(S.dist[v])) {
      const uint64_t old = atomicMin((unsigned long long *)(S.dist + v),
                                     (unsigned long long)pack(D, u));
      if (D < dist_of(old)) push = atomicExch(S.stamp + v, epoch) != epoch;
    }
  }
  enqueue(G, S, nx, push, v);
}

// Decremental seed: batch edge (u, v) was deleted; if it was v's tree edge,
// v becomes invalid (Invalidate, BFSImpl.cuh:3-15).
__global__ void __launch_bounds__(kBlock)
    k_invalidate_batch(GraphView G, TravView S, const uint32_t *src, const uint32_t *dst,
                       uint32_t n, uint32_t nx, uint32_t source) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool push = false;
  uint32_t v = 0;
  if (i < n) {
    const uint32_t u = src[i];
    v = dst[i];
    if (u != v && u < G.n && v < G.n && v != source) {
      const uint64_t cur = coherent_load64(S.dist + v);
      if (cur != kInfInf && parent_of(cur) == u) {
        if (atomicCAS((unsigned long long *)(S.dist + v), (unsigned long long)cur,
                      (unsigned long long)kInfInf) == cur) {
          atomicOr(S.invbits + (v >> 5), 1u << (v & 31u));
          push = true;
        }
      }
    }
  }
  enqueue(G, S, nx, push, v);
}

MK_DEV bool inv_bit(const uint32_t *bits, uint32_t v) { return (bits[v >> 5] >> (v & 31u)) & 1u; }

// Re-seed invalidated vertices from valid in-neighbours: one wave per bucket
// over the whole graph, edges kept only if the target is invalidated.
template <bool MAP, bool WEIGHTS>
__global__ void __launch_bounds__(kBlock)
    k_reseed(GraphView G, TravView S, uint32_t nx, uint32_t epoch) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint64_t km = key_mask<MAP>();
  const bool key_lane = (km >> lane) & 1ull;
  uint32_t slabs = 0, edges = 0;
  for (uint32_t b = gw; b < G.total_buckets; b += W) {
    const uint32_t u = G.bucket_owner[b];
    if (inv_bit(S.invbits, u)) continue;
    const uint32_t du = dist_of(S.dist[u]);
    if (du == kInf) continue;
    const uint32_t *slab = G.heads + size_t(b) * kSlabWords;
    for (;;) {
      const uint32_t v = slab[lane];
      ++slabs;
      edges += uint32_t(__popcll(ballot(key_lane && v < kTomb)));
      const uint32_t wv = MAP ? __shfl_down(v, 1, 64) : 1u;
      bool push = false;
      if (key_lane && v < kTomb && inv_bit(S.invbits, v)) {
        const uint32_t D = du + (WEIGHTS ? wv : 1u);
        if (D >= du) {
          const uint64_t old = atomicMin((unsigned long long *)(S.dist + v),
                                         (unsigned long long)pack(D, u));
          if (D < dist_of(old)) push = atomicExch(S.stamp + v, epoch) != epoch;
        }
      }
      enqueue(G, S, nx, push, v);
      const uint32_t nxt = read_lane(v, kNextLane);
      if (nxt == kEmpty) break;
      slab = slab_ptr(G.pool, nxt);
    }
  }
  if (lane == 0 && slabs) {
    atomicAdd(S.slabs_read, (unsigned long long)slabs);
    atomicAdd(S.edges_read, (unsigned long long)edges);
  }
}

__global__ void k_set_source(uint64_t *dist, uint32_t source, uint32_t *q_small,
                             uint32_t *cnt_small) {
  dist[source] = pack(0, source);
  q_small[0] = source;
  cnt_small[0] = 1;
}

__global__ void k_unpack(const uint64_t *dist, uint32_t n, uint32_t *d, uint32_t *p) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t x = dist[i];
  if (d) d[i] = dist_of(x);
  if (p) p[i] = parent_of(x);
}

// ---------------------------------------------------------------- host
struct Counts {
  uint32_t small[2];
  uint32_t pad[2];
  unsigned long long large[2];
  unsigned long long slabs;
  unsigned long long edges;
};

struct Traversal {
  mk_graph *g = nullptr;
  uint32_t source = 0;
  bool weights = false;  // SSSP
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t *dist = nullptr;
  uint32_t *stamp = nullptr;
  uint32_t *invbits = nullptr;
  uint32_t *qs[2] = {nullptr, nullptr};
  uint2 *ql[2] = {nullptr, nullptr};
  Counts *cnt = nullptr;  // device
  HostBuf host;           // pinned Counts
  uint32_t epoch = 0;
  uint32_t last_iters = 0;
  uint64_t last_slabs = 0;
  uint64_t last_edges = 0;
  int max_waves = 0;

  TravView view() const {
    TravView v;
    v.dist = dist;
    v.stamp = stamp;
    v.invbits = invbits;
    v.q_small[0] = qs[0];
    v.q_small[1] = qs[1];
    v.q_large[0] = ql[0];
    v.q_large[1] = ql[1];
    v.cnt_small = cnt->small;
    v.cnt_large = cnt->large;
    v.slabs_read = &cnt->slabs;
    v.edges_read = &cnt->edges;
    return v;
  }

  int init(mk_graph *graph, uint32_t src, bool w) {
    g = graph;
    source = src;
    weights = w;
    const size_t n = g->n;
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipStreamCreate(&stream));
    MK_HIP(hipEventCreate(&ev0));
    MK_HIP(hipEventCreate(&ev1));
    MK_HIP(hipMalloc(&dist, n * 8));
    MK_HIP(hipMalloc(&stamp, n * 4));
    MK_HIP(hipMalloc(&invbits, ((n + 31) / 32) * 4));
    for (int i = 0; i < 2; ++i) {
      MK_HIP(hipMalloc(&qs[i], n * 4));
      MK_HIP(hipMalloc(&ql[i], n * 8));
    }
    MK_HIP(hipMalloc(&cnt, sizeof(Counts)));
    if (int r = host.ensure(sizeof(Counts))) return r;
    fill_u64(dist, kInfInf, n, stream);
    fill_u32(stamp, 0u, n, stream);
    fill_u32(invbits, 0u, (n + 31) / 32, stream);
    fill_u32(reinterpret_cast<uint32_t *>(cnt), 0u, sizeof(Counts) / 4, stream);
    k_set_source<<<1, 1, 0, stream>>>(dist, source, qs[0], cnt->small);
    MK_HIP(hipStreamSynchronize(stream));
    max_waves = device_cus(g->device) * 16;
    return MK_OK;
  }

  void destroy() {
    (void)hipSetDevice(g->device);
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(dist);
    (void)hipFree(stamp);
    (void)hipFree(invbits);
    for (int i = 0; i < 2; ++i) {
      (void)hipFree(qs[i]);
      (void)hipFree(ql[i]);
    }
    (void)hipFree(cnt);
    host.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  uint32_t next_epoch() {
    if (++epoch == 0xFFFFFFFFu) {  // wrap: clear stamps once per 4G iterations
      fill_u32(stamp, 0u, g->n, stream);
      epoch = 1;
    }
    return epoch;
  }

  // Clear frontier side `c` counters.
  void clear_side(uint32_t c) {
    fill_u32(&cnt->small[c], 0u, 1, stream);
    fill_u64(reinterpret_cast<uint64_t *>(&cnt->large[c]), 0ull, 1, stream);
  }

  int read_counts() {
    MK_HIP(hipMemcpyAsync(host.ptr, cnt, sizeof(Counts), hipMemcpyDeviceToHost, stream));
    MK_HIP(hipStreamSynchronize(stream));
    return MK_OK;
  }

  template <int MODE> void launch_expand(uint32_t c, uint32_t ep, const Counts &h) {
    const uint32_t ns = h.small[c];
    const uint32_t nl = uint32_t(h.large[c] >> 32);
    const uint32_t items = uint32_t(h.large[c]);
    uint64_t want = std::max<uint64_t>(ns, (uint64_t(items) + kLargeChunk - 1) / kLargeChunk);
    want = std::min<uint64_t>(std::max<uint64_t>(want, 1), uint64_t(max_waves));
    const unsigned grid = blocks_for_waves(want);
    const GraphView G = g->view();
    const TravView S = view();
    ProfScope prof(stream, MODE == 0 ? "k_expand_relax" : "k_expand_invalidate", 0.0);
    if (g->weighted) {
      if (weights)
        k_expand<true, true, MODE><<<grid, kBlock, 0, stream>>>(G, S, c, ep, ns, nl, items);
      else
        k_expand<true, false, MODE><<<grid, kBlock, 0, stream>>>(G, S, c, ep, ns, nl, items);
    } else {
      k_expand<false, false, MODE><<<grid, kBlock, 0, stream>>>(G, S, c, ep, ns, nl, items);
    }
  }

  // Algorithmic bytes of one expand step, from the counters before (p) and
  // after (h) it; `c` is the side it wrote.  256 B per slab read, 8 B per
  // neighbour's packed distance, 24 B per frontier vertex (entry, bucket
  // offsets, its distance), 16 B per vertex pushed (atomic, stamp, entry).
  template <int MODE> void account(const Counts &p, const Counts &h, uint32_t c) {
    const uint32_t src_side = c ^ 1u;
    const double frontier = double(p.small[src_side]) + double(p.large[src_side] >> 32);
    const double pushed = double(h.small[c]) + double(h.large[c] >> 32);
    const double bytes = 256.0 * double(h.slabs - p.slabs) + 8.0 * double(h.edges - p.edges) +
                         24.0 * frontier + 16.0 * pushed;
    prof_add_bytes(MODE == 0 ? "k_expand_relax" : "k_expand_invalidate", bytes);
  }

  // Iterate frontier steps starting from side `c` until the frontier is
  // empty.  Returns the final side.
  template <int MODE> int iterate(uint32_t &c, uint32_t &iters) {
    // a label-correcting run settles within n frontier steps (Bellman-Ford
    // bound); more means corrupted state, reported instead of spinning
    const uint64_t limit = uint64_t(iters) + uint64_t(g->n) + 2;
    bool launched = false;
    Counts prev{};
    for (;;) {
      if (iters > limit) return fail(MK_ERR_STATE, "frontier did not drain within n steps");
      if (int r = read_counts()) return r;
      const Counts h = *host.as<Counts>();
      if (launched && profiling_on()) account<MODE>(prev, h, c);
      prev = h;
      launched = true;
      if (h.small[c] == 0 && (h.large[c] >> 32) == 0) break;
      const uint32_t nx = c ^ 1u;
      launch_expand<MODE>(c, MODE == 0 ? next_epoch() : 0u, h);
      MK_HIP(hipGetLastError());
      c = nx;
      ++iters;
    }
    return MK_OK;
  }

  int begin(float *ms) {
    (void)ms;
    MK_HIP(hipSetDevice(g->device));
    // graph mutations were issued on the caller's stream and completed
    fill_u64(reinterpret_cast<uint64_t *>(&cnt->slabs), 0ull, 2, stream);
    MK_HIP(hipEventRecord(ev0, stream));
    return MK_OK;
  }

  int end(float *ms, uint32_t iters) {
    MK_HIP(hipEventRecord(ev1, stream));
    MK_HIP(hipEventSynchronize(ev1));
    float t = 0.f;
    MK_HIP(hipEventElapsedTime(&t, ev0, ev1));
    if (ms) *ms = t;
    if (int r = read_counts()) return r;
    last_slabs = host.as<Counts>()->slabs;
    last_edges = host.as<Counts>()->edges;
    last_iters = iters;
    return MK_OK;
  }

  int run_static(float *ms) {
    if (int r = begin(ms)) return r;
    fill_u64(dist, kInfInf, g->n, stream);
    clear_side(0);
    clear_side(1);
    k_set_source<<<1, 1, 0, stream>>>(dist, source, qs[0], &cnt->small[0]);
    uint32_t c = 0, iters = 0;
    if (int r = iterate<0>(c, iters)) return r;
    return end(ms, iters);
  }

  int run_incremental(const uint32_t *src, const uint32_t *dst, uint32_t n, float *ms) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "incremental: NULL batch");
    if (int r = begin(ms)) return r;
    uint32_t c = 0, iters = 0;
    clear_side(0);
    clear_side(1);
    if (n) {
      const uint32_t ep = next_epoch();
      const GraphView G = g->view();
      ProfScope prof(stream, "k_seed_batch", 40.0 * n);
      if (weights) {
        const uint64_t target = uint64_t(max_waves);
        uint32_t epw = uint32_t(std::min<uint64_t>(64, std::max<uint64_t>(1, (n + target - 1) / target)));
        const uint64_t waves = (uint64_t(n) + epw - 1) / epw;
        k_seed_batch_weighted<<<blocks_for_waves(waves), kBlock, 0, stream>>>(G, view(), src, dst,
                                                                            n, epw, 0, ep);
      } else {
        k_seed_batch_unit<<<blocks_for_threads(n), kBlock, 0, stream>>>(G, view(), src, dst, n, 0,
                                                                       ep);
      }
      MK_HIP(hipGetLastError());
      if (int r = iterate<0>(c, iters)) return r;
    }
    return end(ms, iters);
  }

  int run_decremental(const uint32_t *src, const uint32_t *dst, uint32_t n, float *ms) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "decremental: NULL batch");
    if (int r = begin(ms)) return r;
    uint32_t c = 0, iters = 0;
    clear_side(0);
    clear_side(1);
    if (n) {
      const GraphView G = g->view();
      {
        ProfScope prof(stream, "k_invalidate_batch", 24.0 * n);
        k_invalidate_batch<<<blocks_for_threads(n), kBlock, 0, stream>>>(G, view(), src, dst, n,
                                                                        0, source);
      }
      MK_HIP(hipGetLastError());
      // propagate invalidation down the tree; count what was invalidated
      uint32_t inv_steps = 0;
      if (int r = read_counts()) return r;
      const Counts h0 = *host.as<Counts>();
      const bool any = h0.small[0] != 0 || (h0.large[0] >> 32) != 0;
      if (any) {
        if (int r = iterate<1>(c, inv_steps)) return r;
        // re-seed from valid in-neighbours over the whole graph
        const uint32_t ep = next_epoch();
        clear_side(0);
        clear_side(1);
        c = 0;
        const unsigned grid = blocks_for_waves(
            std::min<uint64_t>(g->total_buckets, uint64_t(max_waves) * 4));
        Counts before{};
        if (profiling_on()) {
          if (int r = read_counts()) return r;
          before = *host.as<Counts>();
        }
        {
          ProfScope prof(stream, "k_reseed", 0.0);
          if (g->weighted) {
            if (weights)
              k_reseed<true, true><<<grid, kBlock, 0, stream>>>(G, view(), 0, ep);
            else
              k_reseed<true, false><<<grid, kBlock, 0, stream>>>(G, view(), 0, ep);
          } else {
            k_reseed<false, false><<<grid, kBlock, 0, stream>>>(G, view(), 0, ep);
          }
        }
        MK_HIP(hipGetLastError());
        if (profiling_on()) {
          // 256 B per slab, 16 B per bucket (owner, invalid bit, distance),
          // 4 B per key (invalid bit), 16 B per re-seeded vertex
          if (int r = read_counts()) return r;
          const Counts h = *host.as<Counts>();
          prof_add_bytes("k_reseed", 256.0 * double(h.slabs - before.slabs) +
                                         16.0 * double(g->total_buckets) +
                                         4.0 * double(h.edges - before.edges) +
                                         16.0 * (double(h.small[0]) + double(h.large[0] >> 32)));
        }
        fill_u32(invbits, 0u, (uint64_t(g->n) + 31) / 32, stream);
        // the re-seeded vertices are on side 0
        iters = inv_steps + 1;
        if (int r = iterate<0>(c, iters)) return r;
      }
    }
    return end(ms, iters);
  }

  int read(uint32_t *d, uint32_t *p) {
    MK_HIP(hipSetDevice(g->device));
    const size_t n = g->n;
    DevBuf b;
    if (int r = b.ensure(n * 8)) return r;
    uint32_t *dd = b.as<uint32_t>(), *pp = dd + n;
    k_unpack<<<blocks_for_threads(n), kBlock, 0, stream>>>(dist, g->n, dd, pp);
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess && d) e = hipMemcpy(d, dd, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && p) e = hipMemcpy(p, pp, n * 4, hipMemcpyDeviceToHost);
    b.release();
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
  return b->t.run_incremental(src, dst, n, ms);
}
int mk_bfs_decremental(mk_bfs *b, const uint32_t *src, const uint32_t *dst, uint32_t n,
                       float *ms) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_decremental: NULL handle");
  return b->t.run_decremental(src, dst, n, ms);
}
int mk_bfs_read(mk_bfs *b, uint32_t *d, uint32_t *p) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_read: NULL handle");
  return b->t.read(d, p);
}
int mk_bfs_last_stats(mk_bfs *b, uint32_t *iters, uint64_t *slabs, uint64_t *edges) {
  if (!b) return fail(MK_ERR_BAD_ARG, "mk_bfs_last_stats: NULL handle");
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
  (void)w;
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_incremental: NULL handle");
  return s->t.run_incremental(src, dst, n, ms);
}
int mk_sssp_decremental(mk_sssp *s, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                        uint32_t n, float *ms) {
  (void)w;
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_decremental: NULL handle");
  return s->t.run_decremental(src, dst, n, ms);
}
int mk_sssp_read(mk_sssp *s, uint32_t *d, uint32_t *p) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_read: NULL handle");
  return s->t.read(d, p);
}
int mk_sssp_last_stats(mk_sssp *s, uint32_t *iters, uint64_t *slabs, uint64_t *edges) {
  if (!s) return fail(MK_ERR_BAD_ARG, "mk_sssp_last_stats: NULL handle");
  if (iters) *iters = s->t.last_iters;
  if (slabs) *slabs = s->t.last_slabs;
  if (edges) *edges = s->t.last_edges;
  return MK_OK;
}

}  // extern "C"
