// Cooperative slab-hash operations (device side).
//
// Each vertex owns `nb` buckets (contiguous head slabs); a key lives in bucket
// bucket_in(key, nb), in a chain of 256-byte slabs linked through word 63.
// slab_search is called by a whole wave64 (one word per lane); the mutation
// operations by 8-lane groups (below), eight independent operations per
// wave.  Ballots pick the match or the first free slot.  This replaces the
// reference's warp-cooperative SlabHash calls (insertKey / insertPair /
// deleteKey / searchKey, used at include/Graph/warp/InsertEdge.cuh:46-48,86-88,
// DeleteEdge.cuh:23-25, SearchEdge.cuh:23-24,53-55) and the SlabAlloc warp
// allocator.
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
// Return a slab to the pool (single lane).
MK_DEV void pool_free_lane(const PoolView &P, uint32_t idx) {
  atomicAnd(P.bitmap + (idx >> 5), ~(1u << (idx & 31u)));
  atomicSub(P.counters, 1u);
}

// ---------------------------------------------------------------- lookup
// Global bucket index of `key` in src's adjacency.
MK_DEV uint32_t bucket_of(const GraphView &G, uint32_t src, uint32_t key) {
  const uint32_t off = G.bucket_off[src];
  const uint32_t nb = G.bucket_off[src + 1] - off;
  return off + bucket_in(key, nb);
}
MK_DEV uint32_t *head_of(const GraphView &G, uint32_t src, uint32_t key) {
  return G.heads + size_t(bucket_of(G, src, key)) * kSlabWords;
}

template <bool MAP> MK_DEV uint64_t key_mask() { return MAP ? kMapKeyMask : kSetKeyMask; }

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

// ================================================================ groups
// Group-wide variants: a wave runs kGroupsPerWave independent slab
// operations, one per group of kGroupLanes lanes.  Each group lane holds
// kWordsPerLane words of the current slab (the group reads all 256 bytes at
// once): words kWordsPerLane*gl .. kWordsPerLane*(gl+1)-1 after load_words,
// words coh_word(gl, 0..7) after the agent-scope load_slab_coherent used
// where the slab may change during the kernel; a claim takes the first free
// word in the order of that layout.  So several chain walks / claims are in
// flight per wave instead of one.  Every value the
// group branches on is group-uniform (broadcast with a group-width shuffle),
// so the lanes of a group stay active together and ballots restricted to
// the group's bits are exact.  The concurrency contract is the one at the
// top of this file.
constexpr uint32_t kGroupLanes = 8;
constexpr uint32_t kGroupsPerWave = 64 / kGroupLanes;
constexpr uint32_t kWordsPerLane = 64 / kGroupLanes;
constexpr uint32_t kGroupMask = (1u << kGroupLanes) - 1u;

MK_DEV uint32_t group_lane() { return lane_id() & (kGroupLanes - 1); }
MK_DEV uint32_t group_first_lane() { return lane_id() & ~(kGroupLanes - 1); }
// The calling group's bits of a wave ballot, bit i = group lane i.
MK_DEV uint32_t group_bits(uint64_t wave_mask) {
  return uint32_t(wave_mask >> group_first_lane()) & kGroupMask;
}
// Value of `v` held by group lane `src` of the calling group.
MK_DEV uint32_t group_read(uint32_t v, uint32_t src) {
  return uint32_t(__shfl(int(v), int(src), int(kGroupLanes)));
}
MK_DEV uint32_t ffs32(uint32_t m) { return uint32_t(__ffs(m) - 1); }

// Slab word held in v[j] of group lane gl after load_slab_coherent: the
// group's k-th 8-byte load covers words 16k..16k+15, 64 contiguous bytes.
// With each lane's eight words side by side instead, each of the four loads
// touches both 128-byte lines of the slab: 2M random slabs read that way
// took 164 us on MI355X, against 115 us in this layout.  Word 63, the next
// pointer, is v[7] of lane 7 in both; a word's parity is j's.
MK_DEV uint32_t coh_word(uint32_t gl, uint32_t j) {
  return 16u * (j >> 1) + 2u * gl + (j & 1u);
}
// The calling lane's kWordsPerLane words of the 64-word row at `row`
// (a slab or a pool bitmap block, 256-byte aligned), agent-scope loads;
// v[j] is word coh_word(group_lane(), j).
MK_DEV void load_slab_coherent(const uint32_t *row, uint32_t (&v)[kWordsPerLane]) {
  const uint32_t gl = group_lane();
#pragma unroll
  for (uint32_t k = 0; k < kWordsPerLane / 2; ++k) {
    const uint64_t a = coherent_load64(reinterpret_cast<const uint64_t *>(row + coh_word(gl, 2 * k)));
    v[2 * k] = uint32_t(a);
    v[2 * k + 1] = uint32_t(a >> 32);
  }
}
// The same words with plain loads (they may come from this CU's L1, which
// is clean at kernel start, so a view from earlier in the kernel).
MK_DEV void load_slab_plain(const uint32_t *row, uint32_t (&v)[kWordsPerLane]) {
  const uint32_t gl = group_lane();
#pragma unroll
  for (uint32_t k = 0; k < kWordsPerLane / 2; ++k) {
    const uint2 a = *reinterpret_cast<const uint2 *>(row + coh_word(gl, 2 * k));
    v[2 * k] = a.x;
    v[2 * k + 1] = a.y;
  }
}
MK_DEV void load_words(const uint32_t *p, uint32_t (&v)[kWordsPerLane]) {
  const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
  for (uint32_t k = 0; k < kWordsPerLane / 4; ++k) {
    const uint4 x = q[k];
    v[4 * k] = x.x;
    v[4 * k + 1] = x.y;
    v[4 * k + 2] = x.z;
    v[4 * k + 3] = x.w;
  }
}
// v[j] for a lane-varying j (a select chain, no scratch indexing).
MK_DEV uint32_t pick(const uint32_t (&v)[kWordsPerLane], uint32_t j) {
  uint32_t r = v[0];
#pragma unroll
  for (uint32_t k = 1; k < kWordsPerLane; ++k)
    if (j == k) r = v[k];
  return r;
}

// Whether slab word i holds a key (set: 0..62; map: even words 0..60).
template <bool MAP> MK_DEV bool key_word(uint32_t i) {
  return MAP ? ((i & 1u) == 0 && i <= 60u) : i < kNextLane;
}

// Claim one slab for the calling group.  Blocks of the pool bitmap are
// scanned starting from a hash of `seed`; lane gl scans its kWordsPerLane
// bitmap words of a block (as load_slab_coherent lays them out) and the group
// claims a clear bit with atomicOr.  Returns the slab index or kEmpty when
// every block is full.
MK_DEV uint32_t pool_alloc_group(const PoolView &P, uint32_t seed) {
  const uint32_t gl = group_lane();
  if (coherent_load(P.counters) >= P.total_slabs) {
    if (gl == 0) atomicAdd(P.counters + 1, 1u);
    return kEmpty;
  }
  uint32_t blk = key_hash(seed * 0x9E3779B9u + 0x632BE5ABu) % P.nblocks;
  const uint32_t rot = (seed >> 7) & (kGroupLanes - 1);
  for (uint32_t tries = 0; tries < P.nblocks; ++tries) {
    uint32_t *wp = P.bitmap + size_t(blk) * kBlockWords;
    uint32_t v[kWordsPerLane];
    load_slab_coherent(wp, v);
    for (;;) {
      uint32_t fj = kWordsPerLane;  // this lane's first word with a clear bit
#pragma unroll
      for (int j = kWordsPerLane - 1; j >= 0; --j)
        if (v[j] != 0xFFFFFFFFu) fj = uint32_t(j);
      const uint32_t cand = group_bits(ballot(fj < kWordsPerLane));
      if (!cand) break;
      // rotate the starting lane so groups sharing a block spread out
      const uint32_t rotated = ((cand >> rot) | (cand << (kGroupLanes - rot))) & kGroupMask;
      const uint32_t l = (ffs32(rotated) + rot) & (kGroupLanes - 1);
      const uint32_t wj = group_read(fj, l);
      const uint32_t wv = group_read(fj < kWordsPerLane ? pick(v, fj) : 0u, l);
      const uint32_t bit = ffs32(~wv);
      uint32_t old = 0;
      if (gl == l) old = atomicOr(wp + coh_word(l, wj), 1u << bit);
      old = group_read(old, l);
      if ((old & (1u << bit)) == 0) {
        if (gl == 0) atomicAdd(P.counters, 1u);
        return blk * kBlockSlabs + coh_word(l, wj) * 32u + bit;
      }
      if (gl == l) {
        const uint32_t nv = old | (1u << bit);
#pragma unroll
        for (uint32_t k = 0; k < kWordsPerLane; ++k)
          if (wj == k) v[k] = nv;
      }
    }
    blk = (blk + 1 == P.nblocks) ? 0 : blk + 1;
  }
  if (gl == 0) atomicAdd(P.counters + 1, 1u);
  return kEmpty;
}

// Insert (src -> dst [, w]) for one group (arguments group-uniform; `bucket` is
// bucket_of(G, src, dst), computed by the caller for all its edges at once
// so that this load leaves each operation's dependent chain).  `v` holds the
// bucket's head slab in load_slab_coherent's layout, loaded by the caller so
// that it can be in flight during the previous operation.  Plain loads will
// do: during an insert kernel words only go from free to key and next
// pointers from empty to a slab, so an older view can only show a taken
// slot as free or a key as absent, and the claim made from it then fails
// (the swap finds the slot taken) and the walk is redone from agent-scope
// loads.  `v` is overwritten.  Returns 1 if inserted, 0 if already present,
// -1 if the pool is exhausted.
template <bool MAP>
MK_DEV int slab_insert_group(const GraphView &G, const PoolView &P, uint32_t bucket, uint32_t dst,
                             uint32_t w, uint32_t seed, uint32_t (&v)[kWordsPerLane]) {
  const uint32_t gl = group_lane();
  uint32_t *slab = G.heads + size_t(bucket) * kSlabWords;
  uint32_t *free_slab = nullptr;
  uint32_t free_word = 0, free_old = 0;
  for (;;) {
    bool hit = false;
    uint32_t fj = kWordsPerLane;  // this lane's first free key word
#pragma unroll
    for (int j = kWordsPerLane - 1; j >= 0; --j) {
      if (key_word<MAP>(coh_word(gl, j))) {
        hit |= v[j] == dst;
        if (v[j] >= kTomb) fj = uint32_t(j);
      }
    }
    if (group_bits(ballot(hit))) return 0;
    if (free_slab == nullptr) {
      const uint32_t fb = group_bits(ballot(fj < kWordsPerLane));
      if (fb) {
        const uint32_t fl = ffs32(fb);
        free_word = coh_word(fl, group_read(fj, fl));
        free_old = group_read(fj < kWordsPerLane ? pick(v, fj) : 0u, fl);
        free_slab = slab;
      }
    }
    const uint32_t nx = group_read(v[kWordsPerLane - 1], kGroupLanes - 1);  // word 63
    if (nx != kEmpty) {
      slab = slab_ptr(G.pool, nx);
      load_slab_coherent(slab, v);
      continue;
    }
    if (free_slab != nullptr) {
      uint32_t old = 0;
      if (gl == 0) {
        old = atomicCAS(free_slab + free_word, free_old, dst);
        if (MAP && old == free_old) free_slab[free_word + 1] = w;
      }
      old = group_read(old, 0);
      if (old == free_old) return 1;
      if (old == dst) return 0;
      slab = free_slab;  // lost the slot: re-walk from it
      free_slab = nullptr;
      load_slab_coherent(slab, v);
      continue;
    }
    // chain full: link a fresh slab holding the key at its first slot
    const uint32_t ns = pool_alloc_group(P, seed);
    if (ns == kEmpty) return -1;
    uint32_t *nsp = slab_ptr(G.pool, ns) + kWordsPerLane * gl;
#pragma unroll
    for (uint32_t k = 0; k < kWordsPerLane; ++k) {
      uint32_t init = kEmpty;
      if (gl == 0 && k == 0) init = dst;
      if (MAP && gl == 0 && k == 1) init = w;
      coherent_store(nsp + k, init);
    }
    __atomic_thread_fence(__ATOMIC_RELEASE);  // slab contents before the link
    uint32_t old = 0;
    if (gl == 0) old = atomicCAS(slab + kNextLane, kEmpty, ns);
    old = group_read(old, 0);
    if (old == kEmpty) return 1;
    if (gl == 0) pool_free_lane(P, ns);
    slab = slab_ptr(G.pool, old);
    load_slab_coherent(slab, v);
  }
}

// Delete (src -> dst) for one group: the key becomes a tombstone.  `v` holds the
// head slab in load_slab_coherent's layout, loaded by the caller (plain
// loads will do: during a delete kernel words only go from key to
// tombstone and chains do not change, so an older view can only show a key
// already deleted, and the swap, which checks the key, then fails); it is
// overwritten.  Returns 1 if deleted.
template <bool MAP>
MK_DEV int slab_delete_group(const GraphView &G, uint32_t bucket, uint32_t dst,
                             uint32_t (&v)[kWordsPerLane]) {
  const uint32_t gl = group_lane();
  uint32_t *slab = G.heads + size_t(bucket) * kSlabWords;
  for (;;) {
    uint32_t hj = kWordsPerLane;
#pragma unroll
    for (int j = kWordsPerLane - 1; j >= 0; --j)
      if (key_word<MAP>(coh_word(gl, j)) && v[j] == dst) hj = uint32_t(j);
    const uint32_t hb = group_bits(ballot(hj < kWordsPerLane));
    if (hb) {
      const uint32_t hl = ffs32(hb);
      const uint32_t word = coh_word(hl, group_read(hj, hl));
      uint32_t old = 0;
      if (gl == 0) old = atomicCAS(slab + word, dst, kTomb);
      old = group_read(old, 0);
      return old == dst ? 1 : 0;
    }
    const uint32_t nx = group_read(v[kWordsPerLane - 1], kGroupLanes - 1);
    if (nx == kEmpty) return 0;
    slab = slab_ptr(G.pool, nx);
    load_slab_plain(slab, v);
  }
}

// Map value update for one group: key dst's value word becomes w (key
// absent: nothing is written).  Returns 1 if the key is present.  Updates
// of one batch run concurrently; a batch naming a key twice keeps one of
// its values.
MK_DEV int slab_update_group(const GraphView &G, uint32_t bucket, uint32_t dst, uint32_t w) {
  const uint32_t gl = group_lane();
  uint32_t *slab = G.heads + size_t(bucket) * kSlabWords;
  for (;;) {
    uint32_t v[kWordsPerLane];
    load_slab_coherent(slab, v);
    uint32_t hj = kWordsPerLane;
#pragma unroll
    for (int j = kWordsPerLane - 1; j >= 0; --j)
      if (key_word<true>(coh_word(gl, j)) && v[j] == dst) hj = uint32_t(j);
    const uint32_t hb = group_bits(ballot(hj < kWordsPerLane));
    if (hb) {
      const uint32_t hl = ffs32(hb);
      const uint32_t word = coh_word(hl, group_read(hj, hl));
      if (gl == 0) coherent_store(slab + word + 1, w);
      return 1;
    }
    const uint32_t nx = group_read(v[kWordsPerLane - 1], kGroupLanes - 1);
    if (nx == kEmpty) return 0;
    slab = slab_ptr(G.pool, nx);
  }
}

// slab_search for one group (no concurrent writers).  Returns 1 if present
// and sets *w (map value, or 0).
template <bool MAP>
MK_DEV int slab_search_group(const GraphView &G, uint32_t bucket, uint32_t dst, uint32_t *w) {
  const uint32_t gl = group_lane();
  const uint32_t *slab = G.heads + size_t(bucket) * kSlabWords;
  for (;;) {
    uint32_t v[kWordsPerLane];
    load_words(slab + kWordsPerLane * gl, v);
    uint32_t hj = kWordsPerLane;
#pragma unroll
    for (int j = kWordsPerLane - 1; j >= 0; --j)
      if (key_word<MAP>(kWordsPerLane * gl + j) && v[j] == dst) hj = uint32_t(j);
    const uint32_t hb = group_bits(ballot(hj < kWordsPerLane));
    if (hb) {
      const uint32_t hl = ffs32(hb);
      // a map key sits at an even word; its value is the next word, in the
      // same lane (kWordsPerLane is even)
      const uint32_t val = hj < kWordsPerLane ? pick(v, hj + 1 < kWordsPerLane ? hj + 1 : hj) : 0u;
      *w = MAP ? group_read(val, hl) : 0u;
      return 1;
    }
    const uint32_t nx = group_read(v[kWordsPerLane - 1], kGroupLanes - 1);
    if (nx == kEmpty) return 0;
    slab = slab_ptr(G.pool, nx);
  }
}

}  // namespace mk
