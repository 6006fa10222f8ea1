// Wave-cooperative slab-hash operations (device side).
//
// Each vertex owns `nb` buckets (contiguous head slabs); a key lives in bucket
// bucket_in(key, nb), in a chain of 256-byte slabs linked through lane 63.
// All functions here are called by a whole wave64 with uniform arguments:
// every lane reads one word of the current slab, and ballots pick the match or
// the first free slot.  This replaces the reference's warp-cooperative
// SlabHash calls (insertKey / insertPair / deleteKey / searchKey, used at
// include/Graph/warp/InsertEdge.cuh:46-48,86-88, DeleteEdge.cuh:23-25,
// SearchEdge.cuh:23-24,53-55) and the SlabAlloc warp allocator.
//
// Concurrency contract: one kernel either inserts or deletes, never both.
// Inserts claim the first free slot (EMPTY or TOMB) in chain order after
// walking the whole chain, and re-walk after a lost CAS, so two inserts of
// the same key cannot both succeed: slots only go free -> key during an
// insert kernel, hence a later claimant has observed every earlier slot
// occupied, including the one holding the key.
#pragma once

#include "mk_internal.hpp"

namespace mk {

// ---------------------------------------------------------------- pool
// Claim one slab for the calling wave.  Blocks are scanned starting from a
// hash of `seed`; within a block, lane l owns bitmap word l and the wave
// claims the lowest clear bit of the first non-full word with atomicOr.
// Returns the slab index or kEmpty when every block is full.
MK_DEV uint32_t pool_alloc(const PoolView &P, uint32_t seed) {
  const uint32_t lane = lane_id();
  if (coherent_load(P.counters) >= P.total_slabs) {
    if (lane == 0) atomicAdd(P.counters + 1, 1u);
    return kEmpty;
  }
  uint32_t blk = key_hash(seed * 0x9E3779B9u + 0x632BE5ABu) % P.nblocks;
  for (uint32_t tries = 0; tries < P.nblocks; ++tries) {
    uint32_t *wp = P.bitmap + size_t(blk) * kBlockWords;
    uint32_t word = coherent_load(wp + lane);
    uint64_t cand = ballot(word != 0xFFFFFFFFu);
    // rotate the starting lane so waves that share a block spread out
    const uint32_t rot = (seed >> 7) & 63u;
    while (cand) {
      uint64_t rotated = (cand >> rot) | (rot ? (cand << (64 - rot)) : 0);
      uint32_t l = (ffs64(rotated) + rot) & 63u;
      uint32_t wl = read_lane(word, l);
      uint32_t bit = uint32_t(__ffs(~wl) - 1);
      uint32_t old = 0;
      if (lane == l) old = atomicOr(wp + l, 1u << bit);
      old = read_lane(old, l);
      if ((old & (1u << bit)) == 0) {
        if (lane == 0) atomicAdd(P.counters, 1u);
        return blk * kBlockSlabs + l * 32u + bit;
      }
      if (lane == l) word = old | (1u << bit);
      cand = ballot(word != 0xFFFFFFFFu);
    }
    blk = (blk + 1 == P.nblocks) ? 0 : blk + 1;
  }
  if (lane == 0) atomicAdd(P.counters + 1, 1u);
  return kEmpty;
}

// Return a slab to the pool (single lane).
MK_DEV void pool_free_lane(const PoolView &P, uint32_t idx) {
  atomicAnd(P.bitmap + (idx >> 5), ~(1u << (idx & 31u)));
  atomicSub(P.counters, 1u);
}

// ---------------------------------------------------------------- lookup
MK_DEV uint32_t *head_of(const GraphView &G, uint32_t src, uint32_t key) {
  const uint32_t off = G.bucket_off[src];
  const uint32_t nb = G.bucket_off[src + 1] - off;
  return G.heads + size_t(off + bucket_in(key, nb)) * kSlabWords;
}

template <bool MAP> MK_DEV uint64_t key_mask() { return MAP ? kMapKeyMask : kSetKeyMask; }

// Insert (src -> dst [, w]).  Returns 1 if inserted, 0 if already present,
// -1 if the pool is exhausted.
template <bool MAP>
MK_DEV int slab_insert(const GraphView &G, const PoolView &P, uint32_t src, uint32_t dst,
                       uint32_t w, uint32_t seed) {
  const uint32_t lane = lane_id();
  const uint64_t km = key_mask<MAP>();
  uint32_t *slab = head_of(G, src, dst);
  uint32_t *free_slab = nullptr;
  uint32_t free_lane = 0, free_old = 0;
  for (;;) {
    const uint32_t v = coherent_load(slab + lane);
    if (ballot(v == dst) & km) return 0;
    if (free_slab == nullptr) {
      const uint64_t fr = ballot(v >= kTomb) & km;
      if (fr) {
        free_lane = ffs64(fr);
        free_slab = slab;
        free_old = read_lane(v, free_lane);
      }
    }
    const uint32_t nx = read_lane(v, kNextLane);
    if (nx != kEmpty) {
      slab = slab_ptr(G.pool, nx);
      continue;
    }
    if (free_slab != nullptr) {
      uint32_t old = 0;
      if (lane == 0) {
        old = atomicCAS(free_slab + free_lane, free_old, dst);
        if (MAP && old == free_old) free_slab[free_lane + 1] = w;
      }
      old = read_lane(old, 0);
      if (old == free_old) return 1;
      if (old == dst) return 0;
      slab = free_slab;  // lost the slot: re-walk from it
      free_slab = nullptr;
      continue;
    }
    // chain full: link a fresh slab holding the key at its first slot
    const uint32_t ns = pool_alloc(P, seed);
    if (ns == kEmpty) return -1;
    uint32_t *nsp = slab_ptr(G.pool, ns);
    uint32_t init = kEmpty;
    if (lane == 0) init = dst;
    if (MAP && lane == 1) init = w;
    coherent_store(nsp + lane, init);
    __atomic_thread_fence(__ATOMIC_RELEASE);  // slab contents before the link
    uint32_t old = 0;
    if (lane == 0) old = atomicCAS(slab + kNextLane, kEmpty, ns);
    old = read_lane(old, 0);
    if (old == kEmpty) return 1;
    if (lane == 0) pool_free_lane(P, ns);
    slab = slab_ptr(G.pool, old);
  }
}

// Delete (src -> dst): the key becomes a tombstone.  Returns 1 if deleted.
template <bool MAP>
MK_DEV int slab_delete(const GraphView &G, uint32_t src, uint32_t dst) {
  const uint32_t lane = lane_id();
  const uint64_t km = key_mask<MAP>();
  uint32_t *slab = head_of(G, src, dst);
  for (;;) {
    const uint32_t v = coherent_load(slab + lane);
    const uint64_t hit = ballot(v == dst) & km;
    if (hit) {
      const uint32_t l = ffs64(hit);
      uint32_t old = 0;
      if (lane == 0) old = atomicCAS(slab + l, dst, kTomb);
      old = read_lane(old, 0);
      return old == dst ? 1 : 0;
    }
    const uint32_t nx = read_lane(v, kNextLane);
    if (nx == kEmpty) return 0;
    slab = slab_ptr(G.pool, nx);
  }
}

// Search (src -> dst).  Returns 1 if present and sets *w (map value, or 0).
template <bool MAP>
MK_DEV int slab_search(const GraphView &G, uint32_t src, uint32_t dst, uint32_t *w) {
  const uint32_t lane = lane_id();
  const uint64_t km = key_mask<MAP>();
  const uint32_t *slab = head_of(G, src, dst);
  for (;;) {
    const uint32_t v = slab[lane];
    const uint64_t hit = ballot(v == dst) & km;
    if (hit) {
      const uint32_t l = ffs64(hit);
      *w = MAP ? read_lane(v, l + 1) : 0u;
      return 1;
    }
    const uint32_t nx = read_lane(v, kNextLane);
    if (nx == kEmpty) return 0;
    slab = slab_ptr(G.pool, nx);
  }
}

}  // namespace mk
