// Triangle counting over an undirected graph stored in both directions.
//
// Reference: include/Algorithms/TC.cuh:9-41 and impl/TCImpl.cuh.  Static
// (:288-380) counts, for every directed edge (u, v) oriented from the lower
// degree end, the neighbours w of u found in adj[v] by a warp SearchEdge
// (TriangleCount_1, :149-214), and divides by 6 (:368).  Incremental
// (:431-497) and Decremental (:499-565) combine three counts over the batch
// graph in closed form.
//
// Method here (same counts):
//  - Static: vertices are ranked by (degree, id) and N+(x) holds the
//    neighbours of x ranked above x (|N+(x)| <= sqrt(2m)).  Each triangle
//    a < b < c is counted once, at a, as an element of N+(a) and N+(b) for b
//    in N+(a).  The N+ lists are built compactly in HBM (count pass,
//    exclusive scan, fill pass over the slabs); then one wave per vertex a
//    puts N+(a) (1024 keys at a time) into an LDS hash table and streams
//    every N+(b), b in N+(a), through it with coalesced loads.
//  - Incremental / Decremental: the batch's changed undirected edges (the
//    entries the last insert / delete applied) go into a device hash set.
//    For each changed edge {u, v} the wave enumerates the neighbours w of the
//    cheaper end and tests the other end; a triangle found with j changed
//    edges among its three adds 6 / j, so each changed triangle sums to 6
//    over its j changed edges and the total / 6 is exact (no closed form over
//    separate counts).  For deletions the neighbourhoods are those before
//    the delete: current slabs plus the deleted edges (a CSR of the deleted
//    set built per call).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "profile.hpp"
#include "slabhash.hpp"

namespace mk {

constexpr uint64_t kNoKey = ~0ull;
constexpr uint32_t kTcSlots = 64;

// rank(a) < rank(b) for (degree, id) ordering.
MK_DEV bool rank_lt(uint32_t da, uint32_t a, uint32_t db, uint32_t b) {
  return da < db || (da == db && a < b);
}

// Lanes with `act` hold key c: returns whether c is a live key of vertex y.
// Lanes targeting the same bucket of y share one walk of its chain.
// Wave-collective.
template <bool MAP>
MK_DEV bool probe_member(const GraphView &G, uint32_t y, bool act, uint32_t c, uint32_t &slabs) {
  const uint32_t lane = lane_id();
  const uint32_t off = G.bucket_off[y], nb = G.bucket_off[y + 1] - off;
  const uint64_t km = key_mask<MAP>();
  const uint32_t bk = act ? off + bucket_in(c, nb) : 0u;
  bool hit = false;
  uint64_t pending = ballot(act);
  while (pending) {
    const uint32_t bb = read_lane(bk, ffs64(pending));
    const bool mine = act && bk == bb;
    pending &= ~ballot(mine);
    const uint64_t mm = ballot(mine);
    const uint32_t *slab = G.heads + size_t(bb) * kSlabWords;
    for (;;) {
      const uint32_t x = slab[lane];
      ++slabs;
      uint64_t live = ballot(x < kTomb) & km;
      if (__popcll(mm) < __popcll(live)) {
        // few candidates: broadcast each and compare with the whole slab
        uint64_t t = mm;
        while (t) {
          const uint32_t j = ffs64(t);
          t &= t - 1;
          const uint32_t cj = read_lane(c, j);
          if (ballot(x == cj) & live) hit |= lane == j;
        }
      } else {
        while (live) {
          const uint32_t k = ffs64(live);
          live &= live - 1;
          hit |= mine && c == read_lane(x, k);
        }
      }
      const uint32_t nx = read_lane(x, kNextLane);
      if (nx == kEmpty) break;
      slab = slab_ptr(G.pool, nx);
    }
  }
  return hit;
}

MK_DEV uint64_t edge_key(uint32_t a, uint32_t b) {
  return a < b ? (uint64_t(a) << 32) | b : (uint64_t(b) << 32) | a;
}
MK_DEV uint32_t key_slot(uint64_t k, uint32_t mask) {
  uint64_t h = k * 0x9E3779B97F4A7C15ull;
  return uint32_t(h >> 32) & mask;
}
// Insert into the open-addressing set; true if the key was not present.
MK_DEV bool set_insert(unsigned long long *tab, uint32_t mask, uint64_t k) {
  uint32_t h = key_slot(k, mask);
  for (;;) {
    const unsigned long long old = atomicCAS(tab + h, kNoKey, (unsigned long long)k);
    if (old == kNoKey) return true;
    if (old == k) return false;
    h = (h + 1) & mask;
  }
}
MK_DEV bool set_contains(const unsigned long long *tab, uint32_t mask, uint64_t k) {
  uint32_t h = key_slot(k, mask);
  for (;;) {
    const unsigned long long x = tab[h];
    if (x == k) return true;
    if (x == kNoKey) return false;
    h = (h + 1) & mask;
  }
}

MK_DEV void add_count(unsigned long long *slots, uint64_t v) {
  if (v) atomicAdd(&slots[blockIdx.x % kTcSlots], (unsigned long long)v);
}

// Static count over higher-ranked neighbour lists.  With vertices ranked by
// (degree, id) and N+(x) the neighbours of x ranked above x, every triangle
// a < b < c is counted once, at a, as an element of N+(a) \cap N+(b) for
// b in N+(a).  Pass 1 counts |N+(a)|, an exclusive scan places the lists,
// pass 2 writes them (compact, in HBM); the intersection kernel puts N+(a)
// (in chunks of kUpChunk) in an LDS hash table and streams each N+(b)
// through it with coalesced loads.

// Wave-collective walk of all slabs of vertex a (buckets loaded four at a
// time); for every live key x ranked above a calls f(up-ballot, x).
template <bool MAP, typename F>
MK_DEV void walk_up(const GraphView &G, uint32_t a, uint32_t da, uint32_t &slabs, F f) {
  const uint32_t lane = lane_id();
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  const uint32_t b0 = G.bucket_off[a], b1 = G.bucket_off[a + 1];
  for (uint32_t bk = b0; bk < b1; bk += 4) {
    uint32_t x[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      x[k] = bk + k < b1 ? G.heads[size_t(bk + k) * kSlabWords + lane] : kEmpty;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      if (bk + k >= b1) break;
      uint32_t v = x[k];
      for (;;) {
        ++slabs;
        const bool live = key_lane && v < kTomb;
        const uint32_t dv = live ? G.vdeg[v] : 0u;
        const bool up = live && rank_lt(da, a, dv, v);
        f(ballot(up), up, v);
        const uint32_t nx = read_lane(v, kNextLane);
        if (nx == kEmpty) break;
        v = slab_ptr(G.pool, nx)[lane];
      }
    }
  }
}

// Vertices are handed to waves kGrab at a time from a device counter: RMAT
// puts its heavy vertices at ids with many low zero bits, so a fixed
// grid stride (a power of two) would give them all to the same few waves.
constexpr uint32_t kGrab = 16;

// 64-bit so the counter cannot wrap past n while waves are still taking.
MK_DEV uint64_t grab(unsigned long long *next) {
  unsigned long long a0 = 0;
  if (lane_id() == 0) a0 = atomicAdd(next, (unsigned long long)kGrab);
  const uint32_t lo = read_lane(uint32_t(a0), 0), hi = read_lane(uint32_t(a0 >> 32), 0);
  return (uint64_t(hi) << 32) | lo;
}

template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_tc_count_up(GraphView G, unsigned long long *K, unsigned long long *slab_count,
                  unsigned long long *next) {
  const uint32_t lane = lane_id();
  uint32_t slabs = 0;
  for (uint64_t a0 = grab(next); a0 < G.n; a0 = grab(next)) {
    const uint32_t a1 = uint32_t(min<uint64_t>(a0 + kGrab, G.n));
    for (uint32_t a = uint32_t(a0); a < a1; ++a) {
      const uint32_t da = G.vdeg[a];
      uint64_t k = 0;
      if (da) walk_up<MAP>(G, a, da, slabs, [&](uint64_t m, bool, uint32_t) { k += __popcll(m); });
      if (lane == 0) K[a] = k;
    }
  }
  if (lane == 0 && slabs) atomicAdd(slab_count, (unsigned long long)slabs);
}

template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_tc_fill_up(GraphView G, const unsigned long long *off, uint32_t *up,
                 unsigned long long *slab_count, unsigned long long *next) {
  uint32_t slabs = 0;
  for (uint64_t a0 = grab(next); a0 < G.n; a0 = grab(next))
  for (uint32_t a = uint32_t(a0); a < uint32_t(min<uint64_t>(a0 + kGrab, G.n)); ++a) {
    const uint32_t da = G.vdeg[a];
    if (!da) continue;
    uint64_t pos = off[a];
    walk_up<MAP>(G, a, da, slabs, [&](uint64_t m, bool mine, uint32_t v) {
      if (mine) up[pos + lanes_below(m)] = v;
      pos += __popcll(m);
    });
  }
  if (lane_id() == 0 && slabs) atomicAdd(slab_count, (unsigned long long)slabs);
}

constexpr uint32_t kUpChunk = 1024;     // N+(a) keys per LDS table fill
constexpr uint32_t kHashSlots = 2048;   // LDS hash slots per wave (load <= 1/2)

MK_DEV uint32_t lds_slot(uint32_t x) { return (x * 2654435761u) >> (32 - 11); }

__global__ void __launch_bounds__(kBlock)
    k_tc_intersect(const unsigned long long *off, const uint32_t *up, uint32_t n,
                   unsigned long long *slots, unsigned long long *streamed,
                   unsigned long long *next) {
  static_assert(kHashSlots == 1u << 11, "lds_slot assumes 2048 slots");
  const uint32_t lane = lane_id();
  __shared__ uint32_t table[kBlock / 64][kHashSlots];
  uint32_t *h = table[threadIdx.x >> 6];
  uint64_t count = 0, reads = 0;
  for (uint64_t a0 = grab(next); a0 < n; a0 = grab(next))
  for (uint32_t a = uint32_t(a0); a < uint32_t(min<uint64_t>(a0 + kGrab, n)); ++a) {
    const uint64_t oa = off[a];
    const uint32_t Ka = uint32_t(off[a + 1] - oa);
    if (Ka < 2) continue;
    for (uint32_t c0 = 0; c0 < Ka; c0 += kUpChunk) {
      const uint32_t cn = min(kUpChunk, Ka - c0);
      for (uint32_t i = lane; i < kHashSlots; i += 64) h[i] = kEmpty;
      __builtin_amdgcn_wave_barrier();
      for (uint32_t i = lane; i < cn; i += 64) {
        const uint32_t x = up[oa + c0 + i];
        uint32_t sl = lds_slot(x);
        while (atomicCAS(h + sl, kEmpty, x) != kEmpty) sl = (sl + 1) & (kHashSlots - 1);
      }
      __builtin_amdgcn_wave_barrier();
      // every b in N+(a): count N+(b) elements in this chunk of N+(a); the
      // next b's list bounds load while the current list streams, four
      // 64-element rows at a time
      uint32_t bn = up[oa];
      uint64_t obn = off[bn], oen = off[bn + 1];
      for (uint32_t i = 0; i < Ka; ++i) {
        const uint64_t ob = obn;
        const uint32_t Kb = uint32_t(oen - obn);
        if (i + 1 < Ka) {
          bn = up[oa + i + 1];
          obn = off[bn];
          oen = off[bn + 1];
        }
        reads += Kb;
        for (uint32_t j0 = 0; j0 < Kb; j0 += 256) {
          uint32_t c[4];
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t j = j0 + k * 64 + lane;
            c[k] = j < Kb ? up[ob + j] : kEmpty;  // ids are < kEmpty
          }
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) {
            if (c[k] == kEmpty) continue;
            uint32_t sl = lds_slot(c[k]);
            for (;;) {
              const uint32_t e = h[sl];
              if (e == c[k]) {
                ++count;
                break;
              }
              if (e == kEmpty) break;
              sl = (sl + 1) & (kHashSlots - 1);
            }
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  count = wave_sum_u64(count);
  if (lane == 0) {
    add_count(slots, count);
    if (reads) atomicAdd(streamed, (unsigned long long)reads);
  }
}

// Canonical changed edges of a batch: entries with status == APPLIED (or all
// non-loop entries when status is NULL) enter the set; the first entry of
// each undirected edge appends it to `owned`.
__global__ void __launch_bounds__(kBlock)
    k_tc_collect(const uint32_t *src, const uint32_t *dst, const uint8_t *status, uint32_t n,
                 uint32_t nv, unsigned long long *tab, uint32_t mask, uint2 *owned,
                 uint32_t *n_owned) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool mine = false;
  uint32_t u = 0, v = 0;
  if (i < n) {
    u = src[i];
    v = dst[i];
    const bool ok = u != v && u < nv && v < nv && (!status || status[i] == MK_EDGE_APPLIED);
    if (ok) mine = set_insert(tab, mask, edge_key(u, v));
  }
  const uint64_t m = ballot(mine);
  if (m) {
    uint32_t base = 0;
    const uint32_t leader = ffs64(m);
    if (lane_id() == leader) base = atomicAdd(n_owned, uint32_t(__popcll(m)));
    base = read_lane(base, leader);
    if (mine) owned[base + lanes_below(m)] = make_uint2(min(u, v), max(u, v));
  }
}

// Weight of a triangle with j changed edges: 6 / j.
MK_DEV uint32_t six_over(uint32_t j) { return j == 1 ? 6u : (j == 2 ? 3u : 2u); }

// Triangles through each inserted edge {u, v} (graph already updated):
// w over adj[x] of the lower-degree end x, tested in adj[y].
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_tc_inserted(GraphView G, const uint2 *owned, uint32_t n_owned,
                  const unsigned long long *tab, uint32_t mask, unsigned long long *slots,
                  unsigned long long *slab_count) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  uint64_t count = 0;
  uint32_t slabs = 0;
  for (uint32_t e = gw; e < n_owned; e += W) {
    const uint2 uv = owned[e];
    uint32_t x = uv.x, y = uv.y;
    if (G.vdeg[x] > G.vdeg[y]) {
      const uint32_t t = x;
      x = y;
      y = t;
    }
    for (uint32_t bk = G.bucket_off[x]; bk < G.bucket_off[x + 1]; ++bk) {
      const uint32_t *slab = G.heads + size_t(bk) * kSlabWords;
      for (;;) {
        const uint32_t w = slab[lane];
        ++slabs;
        const bool act = key_lane && w < kTomb && w != y;
        if (ballot(act)) {
          const bool hit = probe_member<MAP>(G, y, act, w, slabs);
          uint32_t add = 0;
          if (hit) {
            const uint32_t j = 1u + (set_contains(tab, mask, edge_key(x, w)) ? 1u : 0u) +
                               (set_contains(tab, mask, edge_key(y, w)) ? 1u : 0u);
            add = six_over(j);
          }
          count += wave_sum_u32(add);
        }
        const uint32_t nx = read_lane(w, kNextLane);
        if (nx == kEmpty) break;
        slab = slab_ptr(G.pool, nx);
      }
    }
  }
  if (lane == 0) {
    add_count(slots, count);
    if (slabs) atomicAdd(slab_count, (unsigned long long)slabs);
  }
}

// Degree of every vertex in the deleted set (both ends of each edge).
__global__ void k_tc_dcount(const uint2 *owned, uint32_t n_owned, uint32_t *ddeg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_owned) return;
  atomicAdd(ddeg + owned[i].x, 1u);
  atomicAdd(ddeg + owned[i].y, 1u);
}

__global__ void k_tc_dfill(const uint2 *owned, uint32_t n_owned, const uint32_t *doff,
                           uint32_t *cursor, uint32_t *dlist) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_owned) return;
  const uint2 e = owned[i];
  dlist[doff[e.x] + atomicAdd(cursor + e.x, 1u)] = e.y;
  dlist[doff[e.y] + atomicAdd(cursor + e.y, 1u)] = e.x;
}

// Triangles of the graph before the delete through each deleted edge
// {u, v}: w over (current adj[x]) + (deleted neighbours of x), tested
// against (current adj[y]) + (deleted neighbours of y, via the set).
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_tc_deleted(GraphView G, const uint2 *owned, uint32_t n_owned,
                 const unsigned long long *tab, uint32_t mask, const uint32_t *doff,
                 const uint32_t *dlist, unsigned long long *slots, unsigned long long *slab_count) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  uint64_t count = 0;
  uint32_t slabs = 0;
  for (uint32_t e = gw; e < n_owned; e += W) {
    const uint2 uv = owned[e];
    uint32_t x = uv.x, y = uv.y;
    const uint32_t cx = G.vdeg[x] + (doff[x + 1] - doff[x]);
    const uint32_t cy = G.vdeg[y] + (doff[y + 1] - doff[y]);
    if (cx > cy) {
      const uint32_t t = x;
      x = y;
      y = t;
    }
    // (a) w in the current adjacency of x: {x, w} is not deleted
    for (uint32_t bk = G.bucket_off[x]; bk < G.bucket_off[x + 1]; ++bk) {
      const uint32_t *slab = G.heads + size_t(bk) * kSlabWords;
      for (;;) {
        const uint32_t w = slab[lane];
        ++slabs;
        const bool act = key_lane && w < kTomb && w != y;
        if (ballot(act)) {
          const bool in_now = probe_member<MAP>(G, y, act, w, slabs);
          uint32_t add = 0;
          if (act) {
            const bool yw_del = set_contains(tab, mask, edge_key(y, w));
            if (in_now || yw_del) add = six_over(1u + (yw_del ? 1u : 0u));
          }
          count += wave_sum_u32(add);
        }
        const uint32_t nx = read_lane(w, kNextLane);
        if (nx == kEmpty) break;
        slab = slab_ptr(G.pool, nx);
      }
    }
    // (b) w a deleted neighbour of x: {x, w} is deleted
    const uint32_t d0 = doff[x], d1 = doff[x + 1];
    for (uint32_t j0 = d0; j0 < d1; j0 += 64) {
      const uint32_t j = j0 + lane;
      uint32_t w = 0;
      bool act = false;
      if (j < d1) {
        w = dlist[j];
        act = w != y;
      }
      if (ballot(act) == 0) continue;
      const bool in_now = probe_member<MAP>(G, y, act, w, slabs);
      uint32_t add = 0;
      if (act) {
        const bool yw_del = set_contains(tab, mask, edge_key(y, w));
        if (in_now || yw_del) add = six_over(2u + (yw_del ? 1u : 0u));
      }
      count += wave_sum_u32(add);
    }
  }
  if (lane == 0) {
    add_count(slots, count);
    if (slab_count && slabs) atomicAdd(slab_count, (unsigned long long)slabs);
  }
}

struct TC {
  mk_graph *g = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // kTcSlots counts, then the slab counter and the streamed-elements counter
  unsigned long long *slots = nullptr;
  uint32_t *n_owned = nullptr;
  DevBuf tab, owned, ddeg, doff, cursor, dlist, scan_tmp, kup, koff, uplist, grabs;
  HostBuf host;
  unsigned long long count = 0;
  uint64_t last_slabs = 0, last_streamed = 0;
  int max_waves = 0;

  int init(mk_graph *graph) {
    g = graph;
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipStreamCreate(&stream));
    loop_stream_add(g->device, stream, 1);
    MK_HIP(hipEventCreate(&ev0));
    MK_HIP(hipEventCreate(&ev1));
    MK_HIP(hipMalloc(&slots, (kTcSlots + 2) * 8));
    MK_HIP(hipMalloc(&n_owned, 4));
    if (int r = host.ensure((kTcSlots + 3) * 8)) return r;
    max_waves = grid_waves(g->device);
    return MK_OK;
  }

  void destroy() {
    (void)hipSetDevice(g->device);
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(slots);
    (void)hipFree(n_owned);
    for (DevBuf *b :
         {&tab, &owned, &ddeg, &doff, &cursor, &dlist, &scan_tmp, &kup, &koff, &uplist, &grabs})
      b->release();
    host.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) {
      loop_stream_add(g->device, stream, -1);
      (void)hipStreamDestroy(stream);
    }
  }

  int start() {
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipEventRecord(ev0, stream));
    fill_u64(reinterpret_cast<uint64_t *>(slots), 0ull, kTcSlots + 2, stream);
    return MK_OK;
  }

  // Sum of the count slots; also ends timing.
  int finish(uint64_t &total, float *ms) {
    MK_HIP(hipGetLastError());
    MK_HIP(hipMemcpyAsync(host.ptr, slots, (kTcSlots + 2) * 8, hipMemcpyDeviceToHost, stream));
    MK_HIP(hipEventRecord(ev1, stream));
    MK_HIP(hipEventSynchronize(ev1));
    const unsigned long long *h = host.as<unsigned long long>();
    total = 0;
    for (uint32_t k = 0; k < kTcSlots; ++k) total += h[k];
    last_slabs = h[kTcSlots];
    last_streamed = h[kTcSlots + 1];
    float t = 0.f;
    MK_HIP(hipEventElapsedTime(&t, ev0, ev1));
    if (ms) *ms = t;
    return MK_OK;
  }

  int run_static(float *ms) {
    if (int r = start()) return r;
    const size_t n = g->n;
    if (int r = kup.ensure((n + 1) * 8)) return r;
    if (int r = koff.ensure((n + 1) * 8)) return r;
    if (int r = grabs.ensure(3 * 8)) return r;
    unsigned long long *K = kup.as<unsigned long long>(), *off = koff.as<unsigned long long>();
    // one vertex counter per pass
    unsigned long long *next = grabs.as<unsigned long long>();
    fill_u64(reinterpret_cast<uint64_t *>(next), 0ull, 3, stream);
    const unsigned grid = blocks_for_waves(
        std::min<uint64_t>((n + kGrab - 1) / kGrab, uint64_t(max_waves) * 2));
    {
      // higher-ranked neighbour lists: count, scan, fill
      ProfScope prof(stream, "k_tc_lists", 0.0);
      fill_u64(reinterpret_cast<uint64_t *>(K) + n, 0ull, 1, stream);
      if (g->weighted)
        k_tc_count_up<true><<<grid, kBlock, 0, stream>>>(g->view(), K, slots + kTcSlots, next);
      else
        k_tc_count_up<false><<<grid, kBlock, 0, stream>>>(g->view(), K, slots + kTcSlots, next);
      size_t tmp = 0;
      MK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, K, off, int(n + 1), stream));
      if (int r = scan_tmp.ensure(tmp)) return r;
      MK_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.ptr, tmp, K, off, int(n + 1), stream));
      MK_HIP(hipMemcpyAsync(host.ptr, off + n, 8, hipMemcpyDeviceToHost, stream));
      MK_HIP(hipStreamSynchronize(stream));
      const uint64_t total = *host.as<uint64_t>();
      if (int r = uplist.ensure(std::max<uint64_t>(total, 1) * 4)) return r;
      if (g->weighted)
        k_tc_fill_up<true><<<grid, kBlock, 0, stream>>>(g->view(), off, uplist.as<uint32_t>(),
                                                       slots + kTcSlots, next + 1);
      else
        k_tc_fill_up<false><<<grid, kBlock, 0, stream>>>(g->view(), off, uplist.as<uint32_t>(),
                                                        slots + kTcSlots, next + 1);
    }
    {
      ProfScope prof(stream, "k_tc_intersect", 0.0);
      k_tc_intersect<<<grid, kBlock, 0, stream>>>(off, uplist.as<uint32_t>(), g->n, slots,
                                                  slots + kTcSlots + 1, next + 2);
    }
    uint64_t t = 0;
    if (int r = finish(t, ms)) return r;
    if (profiling_on()) {
      // lists: 256 B per slab read (twice), 4 B per neighbour degree, 4 B
      // per list entry written; intersection: 4 B per streamed element and
      // per staged entry, 16 B per vertex (offsets)
      prof_add_bytes("k_tc_lists", 256.0 * double(last_slabs));
      prof_add_bytes("k_tc_intersect", 4.0 * double(last_streamed) + 16.0 * double(n));
    }
    count = t;
    return MK_OK;
  }

  // Changed undirected edges of the batch into the set; returns their count.
  int collect(const uint32_t *src, const uint32_t *dst, uint32_t n, int op, uint32_t &mask,
              uint32_t &m) {
    uint32_t cap = 1024;
    while (cap < 2ull * n) cap <<= 1;
    mask = cap - 1;
    if (int r = tab.ensure(size_t(cap) * 8)) return r;
    if (int r = owned.ensure(size_t(std::max(n, 1u)) * 8)) return r;
    fill_u64(tab.as<uint64_t>(), ~0ull, cap, stream);
    fill_u32(n_owned, 0u, 1, stream);
    const bool same = g->last_op == op && g->last_src == src && g->last_dst == dst &&
                      g->status_n == n;
    const uint8_t *st = same ? g->status.as<uint8_t>() : nullptr;
    k_tc_collect<<<blocks_for_threads(n), kBlock, 0, stream>>>(
        src, dst, st, n, g->n, tab.as<unsigned long long>(), mask, owned.as<uint2>(), n_owned);
    MK_HIP(hipGetLastError());
    MK_HIP(hipMemcpyAsync(host.ptr, n_owned, 4, hipMemcpyDeviceToHost, stream));
    MK_HIP(hipStreamSynchronize(stream));
    m = *host.as<uint32_t>();
    return MK_OK;
  }

  int run_incremental(const uint32_t *src, const uint32_t *dst, uint32_t n, float *ms) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "tc incremental: NULL batch");
    if (int r = start()) return r;
    uint64_t t = 0;
    if (n) {
      uint32_t mask = 0, m = 0;
      if (int r = collect(src, dst, n, 1, mask, m)) return r;
      if (m) {
        const unsigned grid = blocks_for_waves(std::min<uint64_t>(m, uint64_t(max_waves) * 2));
        ProfScope prof(stream, "k_tc_inserted", 0.0);
        if (g->weighted)
          k_tc_inserted<true><<<grid, kBlock, 0, stream>>>(g->view(), owned.as<uint2>(), m,
                                                          tab.as<unsigned long long>(), mask,
                                                          slots, slots + kTcSlots);
        else
          k_tc_inserted<false><<<grid, kBlock, 0, stream>>>(g->view(), owned.as<uint2>(), m,
                                                           tab.as<unsigned long long>(), mask,
                                                           slots, slots + kTcSlots);
      }
    }
    if (int r = finish(t, ms)) return r;
    if (profiling_on()) prof_add_bytes("k_tc_inserted", 256.0 * double(last_slabs));
    count += t / 6;
    return MK_OK;
  }

  int run_decremental(const uint32_t *src, const uint32_t *dst, uint32_t n, float *ms) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "tc decremental: NULL batch");
    if (int r = start()) return r;
    uint64_t t = 0;
    if (n) {
      uint32_t mask = 0, m = 0;
      if (int r = collect(src, dst, n, 2, mask, m)) return r;
      if (m) {
        // CSR of the deleted edges (both directions)
        const size_t nv = g->n;
        if (int r = ddeg.ensure((nv + 1) * 4)) return r;
        if (int r = doff.ensure((nv + 1) * 4)) return r;
        if (int r = cursor.ensure(nv * 4)) return r;
        if (int r = dlist.ensure(size_t(m) * 8)) return r;
        fill_u32(ddeg.as<uint32_t>(), 0u, nv + 1, stream);
        fill_u32(cursor.as<uint32_t>(), 0u, nv, stream);
        k_tc_dcount<<<blocks_for_threads(m), kBlock, 0, stream>>>(owned.as<uint2>(), m,
                                                                 ddeg.as<uint32_t>());
        size_t tmp = 0;
        MK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, ddeg.as<uint32_t>(),
                                                doff.as<uint32_t>(), int(nv + 1), stream));
        if (int r = scan_tmp.ensure(tmp)) return r;
        MK_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.ptr, tmp, ddeg.as<uint32_t>(),
                                                doff.as<uint32_t>(), int(nv + 1), stream));
        k_tc_dfill<<<blocks_for_threads(m), kBlock, 0, stream>>>(
            owned.as<uint2>(), m, doff.as<uint32_t>(), cursor.as<uint32_t>(),
            dlist.as<uint32_t>());
        const unsigned grid = blocks_for_waves(std::min<uint64_t>(m, uint64_t(max_waves) * 2));
        ProfScope prof(stream, "k_tc_deleted", 0.0);
        if (g->weighted)
          k_tc_deleted<true><<<grid, kBlock, 0, stream>>>(
              g->view(), owned.as<uint2>(), m, tab.as<unsigned long long>(), mask,
              doff.as<uint32_t>(), dlist.as<uint32_t>(), slots, slots + kTcSlots);
        else
          k_tc_deleted<false><<<grid, kBlock, 0, stream>>>(
              g->view(), owned.as<uint2>(), m, tab.as<unsigned long long>(), mask,
              doff.as<uint32_t>(), dlist.as<uint32_t>(), slots, slots + kTcSlots);
      }
    }
    if (int r = finish(t, ms)) return r;
    if (profiling_on()) prof_add_bytes("k_tc_deleted", 256.0 * double(last_slabs));
    count -= t / 6;
    return MK_OK;
  }
};

}  // namespace mk

struct mk_tc {
  mk::TC t;
};

using namespace mk;

extern "C" {

int mk_tc_create(mk_graph *g, mk_tc **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "mk_tc_create: out is NULL");
  *out = nullptr;
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_tc_create: NULL graph");
  mk_tc *h = new mk_tc();
  int r = h->t.init(g);
  if (r) {
    h->t.destroy();
    delete h;
    return r;
  }
  g->users++;
  *out = h;
  return MK_OK;
}

int mk_tc_destroy(mk_tc *t) {
  if (!t) return fail(MK_ERR_BAD_ARG, "mk_tc_destroy: NULL handle");
  t->t.g->users--;
  t->t.destroy();
  delete t;
  return MK_OK;
}

int mk_tc_static(mk_tc *t, float *ms) {
  if (!t) return fail(MK_ERR_BAD_ARG, "mk_tc_static: NULL handle");
  if (int r = graph_settle(t->t.g)) return r;
  return t->t.run_static(ms);
}

int mk_tc_incremental(mk_tc *t, const uint32_t *src, const uint32_t *dst, uint32_t n,
                      float *ms) {
  if (!t) return fail(MK_ERR_BAD_ARG, "mk_tc_incremental: NULL handle");
  if (int r = graph_settle(t->t.g)) return r;
  if (int r = graph_failed_batch(t->t.g, src, dst, n)) return r;
  return t->t.run_incremental(src, dst, n, ms);
}

int mk_tc_decremental(mk_tc *t, const uint32_t *src, const uint32_t *dst, uint32_t n,
                      float *ms) {
  if (!t) return fail(MK_ERR_BAD_ARG, "mk_tc_decremental: NULL handle");
  if (int r = graph_settle(t->t.g)) return r;
  if (int r = graph_failed_batch(t->t.g, src, dst, n)) return r;
  return t->t.run_decremental(src, dst, n, ms);
}

int mk_tc_count(mk_tc *t, uint64_t *count) {
  if (!t || !count) return fail(MK_ERR_BAD_ARG, "mk_tc_count: NULL argument");
  *count = t->t.count;
  return MK_OK;
}

}  // extern "C"
