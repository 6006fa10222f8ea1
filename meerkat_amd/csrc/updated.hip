// Update-aware iteration: the edges inserted since the last
// UpdateSlabPointers, visited without a separate batch graph.
//
// Reference: DynamicGraph::UpdateSlabPointers (include/Graph/SlabGraphImpl.
// cuh:64-82) -> UpdateSlabListPointers (include/Graph/device/UpdateSlablist-
// Pointers.cuh:4-104), with the per-bucket markers FirstUpdatedSlab /
// FirstUpdatedLaneId / IsSlablistUpdated (include/Graph/SlabGraph.cuh:
// 377-379, 425-439).  There, insertion is append-only, so "the keys at or
// after a bucket's marker" are the keys inserted since the marker was last
// moved to the chain's end.
//
// Here a deleted key's slot is reused by later inserts, so a position
// marker cannot tell new keys from old ones.  Instead every slab carries a
// 64-bit mask of its words holding keys inserted since the last
// UpdateSlabPointers (head slabs: one mask per bucket, owned by the graph;
// pool slabs: one per pool slab, owned by the pool), and each bucket a flag
// "inserted into since" (the reference's IsSlablistUpdated).  An insert
// batch marks the slots of the keys it applied (k_mark_new, one chain walk
// per applied edge, run only on graphs created with MK_GRAPH_TRACK_UPDATES);
// UpdateSlabPointers clears the masks along the flagged buckets' chains and
// the flags (k_clear_marks), as the reference walks its flagged buckets to
// move their markers; the iterator (k_emit_new) walks the flagged buckets
// and reports each live key whose mask bit is set.
//
// Why stale bits are harmless: a bit is never cleared by a delete, so a
// tombstone (or, in a freshly linked pool slab, an EMPTY word) may carry
// one; the iterator reports live keys only, and a slot that later receives a
// key receives a new key, whose bit is set anyway.  A live key with its bit
// set was inserted since the bucket was last cleared: every slab that joins
// a chain does so through an insert, which flags the bucket, so the next
// UpdateSlabPointers clears that slab's mask too.
#include <hip/hip_runtime.h>

#include <string>

#include "slabhash.hpp"

namespace mk {

// Mark the slot of every applied edge of an insert batch (status ==
// MK_EDGE_APPLIED) as new, and flag its bucket.  One 8-lane group per edge,
// as k_search.
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_mark_new(GraphView G, const uint32_t *src, const uint32_t *dst, const uint8_t *status,
               uint32_t n, uint64_t *head_new, uint64_t *slab_new, uint8_t *bucket_upd) {
  const uint32_t gl = group_lane();
  const uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x);
  bool act = false;
  uint32_t s = 0, d = 0;
  if (i < n) {
    act = status[i] == MK_EDGE_APPLIED;
    if (act) {
      s = src[i];
      d = dst[i];
    }
  }
  const uint32_t bk = act ? bucket_of(G, s, d) : 0u;
  uint32_t pending = group_bits(ballot(act));
  while (pending) {
    const uint32_t j = ffs32(pending);
    pending &= pending - 1;
    const uint32_t b = group_read(bk, j), key = group_read(d, j);
    const uint32_t *slab = G.heads + size_t(b) * kSlabWords;
    uint64_t *mask = head_new + b;
    for (;;) {
      uint32_t v[kWordsPerLane];
      load_words(slab + kWordsPerLane * gl, v);
      uint32_t hj = kWordsPerLane;
#pragma unroll
      for (int k = kWordsPerLane - 1; k >= 0; --k)
        if (key_word<MAP>(kWordsPerLane * gl + k) && v[k] == key) hj = uint32_t(k);
      const uint32_t hb = group_bits(ballot(hj < kWordsPerLane));
      if (hb) {
        const uint32_t hl = ffs32(hb);
        const uint32_t word = kWordsPerLane * hl + group_read(hj, hl);
        if (gl == 0) {
          atomicOr(reinterpret_cast<unsigned long long *>(mask), 1ull << word);
          bucket_upd[b] = 1;
        }
        break;
      }
      const uint32_t nx = group_read(v[kWordsPerLane - 1], kGroupLanes - 1);
      if (nx == kEmpty) break;  // not found (the insert just applied it: unreachable)
      slab = slab_ptr(G.pool, nx);
      mask = slab_new + nx;
    }
  }
}

// UpdateSlabPointers: for every flagged bucket, clear the masks of its
// chain's slabs and the flag.  One thread per bucket.
__global__ void __launch_bounds__(kBlock)
    k_clear_marks(GraphView G, uint64_t *head_new, uint64_t *slab_new, uint8_t *bucket_upd) {
  const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= G.total_buckets || !bucket_upd[b]) return;
  head_new[b] = 0ull;
  uint32_t nx = G.heads[b * kSlabWords + kNextLane];
  while (nx != kEmpty) {
    slab_new[nx] = 0ull;
    nx = slab_ptr(G.pool, nx)[kNextLane];
  }
  bucket_upd[b] = 0;
}

// The live keys marked new along bucket b's chain, for this group lane (it
// reads words 8gl .. 8gl+7 of each slab; a map key's value is the next word,
// in the same lane): their number, and with `write` each is stored as
// (owner, key[, value]) at out[pos++] while pos < cap.  Group-collective.
template <bool MAP>
MK_DEV uint32_t chain_new_keys(const GraphView &G, const uint64_t *head_new,
                               const uint64_t *slab_new, uint64_t b, uint32_t gl, bool write,
                               uint64_t pos, uint64_t cap, uint32_t *osrc, uint32_t *odst,
                               uint32_t *ow) {
  const uint32_t owner = G.bucket_owner[b];
  const uint32_t *slab = G.heads + b * kSlabWords;
  uint64_t mask = head_new[b];
  uint32_t c = 0;
  for (;;) {
    uint32_t v[kWordsPerLane];
    load_words(slab + kWordsPerLane * gl, v);
#pragma unroll
    for (uint32_t k = 0; k < kWordsPerLane; ++k) {
      const uint32_t wi = kWordsPerLane * gl + k;
      if (key_word<MAP>(wi) && v[k] < kTomb && ((mask >> wi) & 1ull)) {
        if (write && pos < cap) {
          osrc[pos] = owner;
          odst[pos] = v[k];
          if (MAP && ow) ow[pos] = v[k + 1 < kWordsPerLane ? k + 1 : k];
        }
        ++pos;
        ++c;
      }
    }
    const uint32_t nx = group_read(v[kWordsPerLane - 1], kGroupLanes - 1);
    if (nx == kEmpty) break;
    slab = slab_ptr(G.pool, nx);
    mask = slab_new[nx];
  }
  return c;
}

// Exclusive prefix sum of x over the wave (every lane active); *total gets
// the sum.
MK_DEV uint32_t wave_excl_scan(uint32_t x, uint32_t *total) {
  const uint32_t lane = lane_id();
  uint32_t inc = x;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off, 64);
    if (lane >= off) inc += y;
  }
  *total = read_lane(inc, 63);
  return inc - x;
}

// The group's bucket in a round over the flagged buckets of a 64-bucket
// chunk (m: their bits, wave-uniform): group g takes the g-th lowest set bit;
// the round's eight are then taken off m.  False when the group has none.
MK_DEV bool round_bucket(uint64_t &m, uint32_t grp, uint32_t &j) {
  uint64_t mm = m;
  for (uint32_t t = 0; t < grp && mm; ++t) mm &= mm - 1;
  const bool has = mm != 0;
  j = has ? ffs64(mm) : 0u;
  for (uint32_t t = 0; t < 64 / kGroupLanes && m; ++t) m &= m - 1;
  return has;
}

// The update iterator: every live key of a flagged bucket whose slot is
// marked new, as (owner, key[, value]) at out[position] while below `cap`
// (the count includes the ones that did not fit).  Each wave takes a
// contiguous range of buckets in chunks of 64 (one flag per lane, the next
// chunk's flags loaded ahead), its 8-lane groups walking the flagged buckets'
// chains eight at a time.  The first pass counts the range's keys and takes
// the range's positions with one atomic on the shared counter (one per key
// serialised on it: 10 ms for a million keys); the second, unless
// `count_only`, writes them, each group's chain at its offset in the wave's
// running total.
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_emit_new(GraphView G, const uint64_t *head_new, const uint64_t *slab_new,
               const uint8_t *bucket_upd, uint32_t *osrc, uint32_t *odst, uint32_t *ow,
               uint64_t cap, unsigned long long *count, uint32_t count_only) {
  const uint32_t lane = lane_id(), gl = group_lane(), grp = lane / kGroupLanes;
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t W = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  const uint64_t nb = G.total_buckets;
  const uint64_t per = ((nb + W - 1) / W + 63) / 64 * 64;  // whole chunks per wave
  const uint64_t b0 = gw * per, b1 = min(nb, b0 + per);
  auto flag = [&](uint64_t cb) { return cb + lane < b1 && bucket_upd[cb + lane] != 0; };
  uint32_t mine = 0;
  bool fnext = b0 < b1 && flag(b0);
  for (uint64_t cb = b0; cb < b1; cb += 64) {
    uint64_t m = ballot(fnext);
    fnext = cb + 64 < b1 && flag(cb + 64);
    while (m) {
      uint32_t j;
      if (round_bucket(m, grp, j))
        mine += chain_new_keys<MAP>(G, head_new, slab_new, cb + j, gl, false, 0, 0, nullptr,
                                    nullptr, nullptr);
    }
  }
  uint32_t total = 0;
  (void)wave_excl_scan(mine, &total);
  if (total == 0) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(count, (unsigned long long)total);
  if (count_only) return;
  uint64_t run = (uint64_t(read_lane(uint32_t(base >> 32), 0)) << 32) |
                 read_lane(uint32_t(base), 0);
  fnext = b0 < b1 && flag(b0);
  for (uint64_t cb = b0; cb < b1; cb += 64) {
    uint64_t m = ballot(fnext);
    fnext = cb + 64 < b1 && flag(cb + 64);
    while (m) {
      uint32_t j;
      const bool act = round_bucket(m, grp, j);
      const uint32_t c = act ? chain_new_keys<MAP>(G, head_new, slab_new, cb + j, gl, false, 0,
                                                   0, nullptr, nullptr, nullptr)
                             : 0u;
      uint32_t step = 0;
      const uint32_t ex = wave_excl_scan(c, &step);
      if (act)  // the whole group walks (the next-slab word is read across it)
        chain_new_keys<MAP>(G, head_new, slab_new, cb + j, gl, true, run + ex, cap, osrc, odst,
                            ow);
      run += step;
    }
  }
}

int track_enable(mk_graph *g) {
  mk_pool *p = g->pool;
  if (!p->slab_new) {
    MK_HIP(hipMalloc(&p->slab_new, p->total_slabs * 8));
    fill_u64(p->slab_new, 0ull, p->total_slabs, nullptr);
  }
  MK_HIP(hipMalloc(&g->head_new, size_t(g->total_buckets) * 8));
  MK_HIP(hipMalloc(&g->bucket_upd, size_t(g->total_buckets)));
  fill_u64(g->head_new, 0ull, g->total_buckets, nullptr);
  fill_u8(g->bucket_upd, 0, g->total_buckets, nullptr);
  MK_HIP(hipDeviceSynchronize());
  g->track = true;
  return MK_OK;
}

void track_release(mk_graph *g) {
  (void)hipFree(g->head_new);
  (void)hipFree(g->bucket_upd);
  g->head_new = nullptr;
  g->bucket_upd = nullptr;
  g->track = false;
}

void mark_new(mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n, hipStream_t s) {
  if (!g->track || n == 0) return;
  // one thread per edge; each 8-lane group walks its lanes' edges in turn
  const unsigned grid = blocks_for_threads(uint64_t(n));
  auto k = g->weighted ? k_mark_new<true> : k_mark_new<false>;
  k<<<grid, kBlock, 0, s>>>(g->view(), src, dst, g->status.as<uint8_t>(), n, g->head_new,
                            g->pool->slab_new, g->bucket_upd);
}

}  // namespace mk

using namespace mk;

extern "C" {

int mk_graph_update_slab_pointers(mk_graph *g) {
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_graph_update_slab_pointers: NULL graph");
  if (g->mirror_of) return fail(MK_ERR_STATE, "mk_graph_update_slab_pointers: mirror");
  if (int r = graph_settle(g)) return r;
  if (!g->track) return MK_OK;  // nothing tracked: nothing to move (as the reference's no-op)
  MK_HIP(hipSetDevice(g->device));
  k_clear_marks<<<blocks_for_threads(g->total_buckets), kBlock, 0, g->own_stream>>>(
      g->view(), g->head_new, g->pool->slab_new, g->bucket_upd);
  MK_HIP(hipGetLastError());
  MK_HIP(hipStreamSynchronize(g->own_stream));
  return MK_OK;
}

int mk_graph_updated_edges(mk_graph *g, uint32_t *src_dev, uint32_t *dst_dev, uint32_t *w_dev,
                           uint64_t capacity, uint64_t *count) {
  if (!g || !count) return fail(MK_ERR_BAD_ARG, "mk_graph_updated_edges: NULL argument");
  if (capacity && (!src_dev || !dst_dev))
    return fail(MK_ERR_BAD_ARG, "mk_graph_updated_edges: NULL output with capacity > 0");
  if (!g->track)
    return fail(MK_ERR_STATE, "mk_graph_updated_edges: the graph was not created with "
                              "MK_GRAPH_TRACK_UPDATES");
  if (int r = graph_settle(g)) return r;
  MK_HIP(hipSetDevice(g->device));
  DevBuf c;
  if (int r = c.ensure(8)) return r;
  unsigned long long *cnt = c.as<unsigned long long>();
  fill_u64(reinterpret_cast<uint64_t *>(cnt), 0ull, 1, g->own_stream);
  // 32 waves per CU: the chain walks are latency-bound
  const unsigned grid = blocks_for_waves(std::min<uint64_t>(
      uint64_t(device_cus(g->device)) * 32, (uint64_t(g->total_buckets) + 63) / 64));
  auto k = g->weighted ? k_emit_new<true> : k_emit_new<false>;
  k<<<grid, kBlock, 0, g->own_stream>>>(g->view(), g->head_new, g->pool->slab_new,
                                         g->bucket_upd, src_dev, dst_dev,
                                         g->weighted ? w_dev : nullptr, capacity, cnt,
                                         capacity == 0 ? 1u : 0u);
  hipError_t e = hipGetLastError();
  uint64_t h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, g->own_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->own_stream);
  c.release();
  if (e != hipSuccess) return hip_fail(e, "mk_graph_updated_edges");
  *count = h;
  return MK_OK;
}

}  // extern "C"
