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
//  - Static: vertices are ranked by (degree, id); each triangle a < b < c
//    (by rank) is counted once at a, for the pair (b, c) of a's
//    higher-ranked neighbours, by testing c in adj[b].  A wave stages a's
//    higher-ranked neighbours (with degrees) in LDS and tests 64 candidates c
//    at a time: lanes that target the same bucket of b share one walk of its
//    slab chain, every slab key being compared against every lane's c.
//    Vertices with more than kLdsCap higher-ranked neighbours re-read a's
//    slabs instead of LDS.
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

constexpr uint32_t kLdsCap = 512;     // higher-ranked neighbours staged per wave
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
    const uint32_t *slab = G.heads + size_t(bb) * kSlabWords;
    for (;;) {
      const uint32_t x = slab[lane];
      ++slabs;
      uint64_t live = ballot(x < kTomb) & km;
      while (live) {
        const uint32_t k = ffs64(live);
        live &= live - 1;
        hit |= mine && c == read_lane(x, k);
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

// Static count.  One wave per vertex a (grid-stride).
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_tc_static(GraphView G, unsigned long long *slots, unsigned long long *slab_count) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint64_t km = key_mask<MAP>();
  const bool key_lane = (km >> lane) & 1ull;
  __shared__ uint32_t lk[kBlock / 64][kLdsCap];
  __shared__ uint32_t ld[kBlock / 64][kLdsCap];
  uint32_t *nk = lk[threadIdx.x >> 6], *nd = ld[threadIdx.x >> 6];
  uint64_t count = 0;
  uint32_t slabs = 0;
  for (uint32_t a = gw; a < G.n; a += W) {
    const uint32_t da = G.vdeg[a];
    if (da < 2) continue;
    const uint32_t b0 = G.bucket_off[a], b1 = G.bucket_off[a + 1];
    // stage a's higher-ranked neighbours
    uint32_t K = 0;
    bool overflow = false;
    for (uint32_t bk = b0; bk < b1; ++bk) {
      const uint32_t *slab = G.heads + size_t(bk) * kSlabWords;
      for (;;) {
        const uint32_t x = slab[lane];
        ++slabs;
        const bool live = key_lane && x < kTomb;
        const uint32_t dx = live ? G.vdeg[x] : 0u;
        const bool up = live && rank_lt(da, a, dx, x);
        const uint64_t m = ballot(up);
        const uint32_t c = uint32_t(__popcll(m));
        if (K + c <= kLdsCap) {
          if (up) {
            nk[K + lanes_below(m)] = x;
            nd[K + lanes_below(m)] = dx;
          }
        } else {
          overflow = true;
        }
        K += c;
        const uint32_t nx = read_lane(x, kNextLane);
        if (nx == kEmpty) break;
        slab = slab_ptr(G.pool, nx);
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (K < 2) continue;
    if (!overflow) {
      // pairs (b, c) of the staged list with rank(b) < rank(c)
      for (uint32_t i = 0; i < K; ++i) {
        const uint32_t b = nk[i], db = nd[i];
        for (uint32_t j0 = 0; j0 < K; j0 += 64) {
          const uint32_t j = j0 + lane;
          uint32_t c = 0;
          bool act = false;
          if (j < K) {
            c = nk[j];
            act = rank_lt(db, b, nd[j], c);
          }
          if (ballot(act) == 0) continue;
          const bool hit = probe_member<MAP>(G, b, act, c, slabs);
          count += uint64_t(__popcll(ballot(hit)));
        }
      }
    } else {
      // too many to stage: take b and c from a's slabs directly
      for (uint32_t bk = b0; bk < b1; ++bk) {
        const uint32_t *slab = G.heads + size_t(bk) * kSlabWords;
        for (;;) {
          const uint32_t x = slab[lane];
          const bool live = key_lane && x < kTomb;
          const uint32_t dx = live ? G.vdeg[x] : 0u;
          uint64_t mb = ballot(live && rank_lt(da, a, dx, x));
          while (mb) {
            const uint32_t jl = ffs64(mb);
            mb &= mb - 1;
            const uint32_t b = read_lane(x, jl), db = read_lane(dx, jl);
            for (uint32_t ck = b0; ck < b1; ++ck) {
              const uint32_t *cs = G.heads + size_t(ck) * kSlabWords;
              for (;;) {
                const uint32_t y = cs[lane];
                const bool ly = key_lane && y < kTomb;
                const uint32_t dy = ly ? G.vdeg[y] : 0u;
                const bool act = ly && rank_lt(db, b, dy, y);
                if (ballot(act)) {
                  const bool hit = probe_member<MAP>(G, b, act, y, slabs);
                  count += uint64_t(__popcll(ballot(hit)));
                }
                const uint32_t ny = read_lane(y, kNextLane);
                if (ny == kEmpty) break;
                cs = slab_ptr(G.pool, ny);
              }
            }
          }
          const uint32_t nx = read_lane(x, kNextLane);
          if (nx == kEmpty) break;
          slab = slab_ptr(G.pool, nx);
        }
      }
    }
  }
  if (lane == 0) {
    add_count(slots, count);
    if (slabs) atomicAdd(slab_count, (unsigned long long)slabs);
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
  unsigned long long *slots = nullptr;  // kTcSlots counts + 1 slab counter
  uint32_t *n_owned = nullptr;
  DevBuf tab, owned, ddeg, doff, cursor, dlist, scan_tmp;
  HostBuf host;
  unsigned long long count = 0;
  uint64_t last_slabs = 0;
  int max_waves = 0;

  int init(mk_graph *graph) {
    g = graph;
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipStreamCreate(&stream));
    MK_HIP(hipEventCreate(&ev0));
    MK_HIP(hipEventCreate(&ev1));
    MK_HIP(hipMalloc(&slots, (kTcSlots + 1) * 8));
    MK_HIP(hipMalloc(&n_owned, 4));
    if (int r = host.ensure((kTcSlots + 2) * 8)) return r;
    max_waves = grid_waves(g->device);
    return MK_OK;
  }

  void destroy() {
    (void)hipSetDevice(g->device);
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(slots);
    (void)hipFree(n_owned);
    for (DevBuf *b : {&tab, &owned, &ddeg, &doff, &cursor, &dlist, &scan_tmp}) b->release();
    host.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int start() {
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipEventRecord(ev0, stream));
    fill_u64(reinterpret_cast<uint64_t *>(slots), 0ull, kTcSlots + 1, stream);
    return MK_OK;
  }

  // Sum of the count slots; also ends timing.
  int finish(uint64_t &total, float *ms) {
    MK_HIP(hipGetLastError());
    MK_HIP(hipMemcpyAsync(host.ptr, slots, (kTcSlots + 1) * 8, hipMemcpyDeviceToHost, stream));
    MK_HIP(hipEventRecord(ev1, stream));
    MK_HIP(hipEventSynchronize(ev1));
    const unsigned long long *h = host.as<unsigned long long>();
    total = 0;
    for (uint32_t k = 0; k < kTcSlots; ++k) total += h[k];
    last_slabs = h[kTcSlots];
    float t = 0.f;
    MK_HIP(hipEventElapsedTime(&t, ev0, ev1));
    if (ms) *ms = t;
    return MK_OK;
  }

  int run_static(float *ms) {
    if (int r = start()) return r;
    const unsigned grid = blocks_for_waves(std::min<uint64_t>(g->n, uint64_t(max_waves) * 2));
    {
      ProfScope prof(stream, "k_tc_static", 0.0);
      if (g->weighted)
        k_tc_static<true><<<grid, kBlock, 0, stream>>>(g->view(), slots, slots + kTcSlots);
      else
        k_tc_static<false><<<grid, kBlock, 0, stream>>>(g->view(), slots, slots + kTcSlots);
    }
    uint64_t t = 0;
    if (int r = finish(t, ms)) return r;
    if (profiling_on()) prof_add_bytes("k_tc_static", 256.0 * double(last_slabs));
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
  return t->t.run_static(ms);
}

int mk_tc_incremental(mk_tc *t, const uint32_t *src, const uint32_t *dst, uint32_t n,
                      float *ms) {
  if (!t) return fail(MK_ERR_BAD_ARG, "mk_tc_incremental: NULL handle");
  return t->t.run_incremental(src, dst, n, ms);
}

int mk_tc_decremental(mk_tc *t, const uint32_t *src, const uint32_t *dst, uint32_t n,
                      float *ms) {
  if (!t) return fail(MK_ERR_BAD_ARG, "mk_tc_decremental: NULL handle");
  return t->t.run_decremental(src, dst, n, ms);
}

int mk_tc_count(mk_tc *t, uint64_t *count) {
  if (!t || !count) return fail(MK_ERR_BAD_ARG, "mk_tc_count: NULL argument");
  *count = t->t.count;
  return MK_OK;
}

}  // extern "C"
