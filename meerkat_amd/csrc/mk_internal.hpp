// Internal definitions shared by the HIP translation units of
// libmeerkat_hip.so: slab layout, bucket hashing, wavefront helpers, the
// device views passed to kernels by value, and the host handle structs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "meerkat.h"

namespace mk {

// ---------------------------------------------------------------- slabs
// A slab is 64 x uint32 = 256 B: one dword per lane of a wave64, read with a
// single coalesced load.  Lane 63 holds the index of the next slab in the
// pool (EMPTY = end of chain).
//   set slab: lanes 0..62 hold keys (63 keys).
//   map slab: even lanes 0..60 hold keys, lane k+1 holds key k's value
//             (31 pairs); lane 62 is unused.
// Words equal to EMPTY are free; TOMB marks a deleted key (reusable by a
// later insert kernel, never by a concurrent one).
constexpr uint32_t kSlabWords = 64;
constexpr uint32_t kNextLane = 63;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kTomb = 0xFFFFFFFEu;
constexpr uint64_t kSetKeyMask = 0x7FFFFFFFFFFFFFFFull;  // lanes 0..62
constexpr uint64_t kMapKeyMask = 0x1555555555555555ull;  // lanes 0,2,..,60
constexpr uint32_t kSetKeysPerSlab = 63;
constexpr uint32_t kMapKeysPerSlab = 31;

// Pool geometry: a block is 2048 slabs whose 2048 bitmap bits are 64 words,
// one per lane, so a wave scans a whole block with one load.
constexpr uint32_t kBlockSlabs = 2048;
constexpr uint32_t kBlockWords = 64;

// 64-bit packed distance: (distance << 32) | parent; a 64-bit atomicMin is
// a lexicographic (distance, parent) minimum (include/Algorithms/
// definitions.cuh:12-17).
constexpr uint64_t kInfInf = 0xFFFFFFFFFFFFFFFFull;
constexpr uint32_t kInf = 0xFFFFFFFFu;

__host__ __device__ inline uint64_t pack(uint32_t d, uint32_t p) {
  return (uint64_t(d) << 32) | p;
}
__host__ __device__ inline uint32_t dist_of(uint64_t x) { return uint32_t(x >> 32); }
__host__ __device__ inline uint32_t parent_of(uint64_t x) { return uint32_t(x); }

// Bucket of `key` among the `nb` buckets of a vertex: a 32-bit finaliser
// (lowbias32 of C. Wellons' hash-prospector) and a multiply-shift range map.
__host__ __device__ inline uint32_t key_hash(uint32_t k) {
  k ^= k >> 16;
  k *= 0x7FEB352Du;
  k ^= k >> 15;
  k *= 0x846CA68Bu;
  k ^= k >> 16;
  return k;
}
__host__ __device__ inline uint32_t bucket_in(uint32_t key, uint32_t nb) {
  return uint32_t((uint64_t(key_hash(key)) * nb) >> 32);
}

// ---------------------------------------------------------------- views
struct PoolView {
  uint32_t *slabs;     // total_slabs * 64 words
  uint32_t *bitmap;    // total_slabs / 32 words, bit set = slab in use
  uint32_t *counters;  // [0] slabs in use, [1] failed allocations
  uint32_t nblocks;
  uint32_t total_slabs;
};

struct GraphView {
  uint32_t *heads;               // total_buckets * 64 words
  uint32_t *pool;                // pool slab base
  const uint32_t *bucket_off;    // n + 1
  const uint32_t *bucket_owner;  // total_buckets
  uint32_t *vdeg;                // live out-degree per vertex
  uint32_t n;
  uint32_t total_buckets;
  uint32_t weighted;
};

#ifdef __HIP_DEVICE_COMPILE__
#define MK_DEV __device__ __forceinline__
#else
#define MK_DEV __device__ inline
#endif

MK_DEV uint32_t lane_id() { return __lane_id(); }

MK_DEV uint64_t ballot(bool p) { return __ballot(p); }

MK_DEV uint32_t read_lane(uint32_t v, uint32_t l) {
  return uint32_t(__builtin_amdgcn_readlane(int(v), int(l)));
}
MK_DEV uint32_t first_lane_val(uint32_t v) { return uint32_t(__builtin_amdgcn_readfirstlane(int(v))); }

// Number of set bits of `mask` below this lane.
MK_DEV uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

MK_DEV uint32_t ffs64(uint64_t m) { return uint32_t(__ffsll((unsigned long long)m) - 1); }

// Items each of W waves takes per step of a grid-stride loop over `count`
// items: spread small counts over every wave (short dependent-load chains
// per wave) and give a full 64 (one per lane) once the grid is covered.
MK_DEV uint32_t per_wave(uint32_t count, uint32_t W) {
  const uint32_t p = (count + W - 1) / W;
  return p < 1u ? 1u : (p > 64u ? 64u : p);
}

// Loads of words that other waves may be writing in the same kernel: agent-
// scope relaxed atomics, so a retry never re-reads a stale non-coherent copy.
MK_DEV uint32_t coherent_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MK_DEV uint64_t coherent_load64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MK_DEV void coherent_store(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

MK_DEV uint32_t *slab_ptr(uint32_t *pool, uint32_t idx) {
  return pool + size_t(idx) * kSlabWords;
}

// Wave-aggregated append of the lanes with `pred` to queue[*count...].
// Returns nothing; lanes with pred write their value.
MK_DEV void wave_push(bool pred, uint32_t value, uint32_t *queue, uint32_t *count) {
  uint64_t m = ballot(pred);
  if (m == 0) return;
  uint32_t base = 0;
  uint32_t leader = ffs64(m);
  if (lane_id() == leader) base = atomicAdd(count, uint32_t(__popcll(m)));
  base = read_lane(base, leader);
  if (pred) queue[base + lanes_below(m)] = value;
}

// Hub lists of the frontier kernels: an entry (vertex, global index of its
// first bucket) per kHubChunk buckets of a listed vertex.
constexpr uint32_t kHubChunk = 8;

// Append the hub vertices v (lanes with `hub`, nb buckets of G each) to a
// hub list: one entry (v, global index of its first bucket) per kHubChunk
// buckets, so a consumer maps its work straight to entries.  One 64-bit
// atomic per call claims the entries of all of them and counts their
// buckets: *count += (entries << 32) | buckets (an exclusive prefix over
// the lanes places each vertex's entries).  Wave-collective.
MK_DEV void push_hubs(const GraphView &G, uint2 *queue, unsigned long long *count, bool hub,
                      uint32_t v, uint32_t nb) {
  const uint64_t hm = ballot(hub);
  if (hm == 0) return;
  const uint32_t lane = lane_id();
  const uint32_t nc = hub ? (nb + kHubChunk - 1) / kHubChunk : 0u;
  uint32_t xc = nc, xb = hub ? nb : 0u;  // inclusive prefixes over the wave
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t yc = __shfl_up(xc, off, 64), yb = __shfl_up(xb, off, 64);
    if (lane >= off) {
      xc += yc;
      xb += yb;
    }
  }
  const uint32_t leader = ffs64(hm);
  unsigned long long r = 0;
  if (lane == leader)
    r = atomicAdd(count, (uint64_t(read_lane(xc, 63)) << 32) | read_lane(xb, 63));
  const uint32_t first = read_lane(uint32_t(r >> 32), leader);
  const uint32_t total = read_lane(xc, 63);
  const uint32_t b0 = hub ? G.bucket_off[v] : 0u;
  // the wave writes the entries together, 64 at a time: entry e belongs to
  // the first lane whose inclusive prefix exceeds e (found by halving)
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t e = base + lane;
    uint32_t lo = 0;  // lanes below `lo` have prefixes <= e
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1)
      if (__shfl(xc, int(lo + step - 1), 64) <= e) lo += step;
    const uint32_t ov = __shfl(v, int(lo), 64), ob0 = __shfl(b0, int(lo), 64);
    const uint32_t oc = e - (__shfl(xc, int(lo), 64) - __shfl(nc, int(lo), 64));
    if (e < total)  // coherent: read by later steps of a persistent loop
      __hip_atomic_store(reinterpret_cast<uint64_t *>(queue + first + e),
                         (uint64_t(ob0 + oc * kHubChunk) << 32) | ov, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Sum of a uint32 over the wave (all lanes must be active).
MK_DEV uint32_t wave_sum_u32(uint32_t v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
MK_DEV uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t lo = __shfl_xor(uint32_t(v), off, 64);
    const uint32_t hi = __shfl_xor(uint32_t(v >> 32), off, 64);
    v += (uint64_t(hi) << 32) | lo;
  }
  return v;
}
MK_DEV double wave_sum_f64(double v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
MK_DEV float wave_sum_f32(float v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace mk

// ---------------------------------------------------------------- host
namespace mk {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
// Settle a graph's deferred mutation (graph.hip): wait for it and return its
// outcome once; MK_OK when none is queued.
int graph_settle(mk_graph *g);
// The error of the graph's last mutation batch when it failed and (src, dst,
// n) names that batch (same device pointers and size), else MK_OK.  An
// algorithm call given a failed batch returns this instead of running on
// edges the graph never applied (every handle of the graph, not only the
// one whose call settled the batch).
int graph_failed_batch(const mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n);
// Streams of the library's handles, per device (traverse.hip).  A frontier
// loop's grid barrier needs every block resident; it is launched as an
// ordinary kernel only while a single stream of the device is registered
// (no other handle's kernels can hold compute units meanwhile), and
// cooperatively otherwise.
void loop_stream_add(int device, hipStream_t s, int delta);
// Update-aware iteration (updated.hip): allocate a graph's new-key masks,
// release them, and mark the keys an insert batch applied (statuses in
// g->status) as new, on stream s (no-op unless the graph tracks updates).
int track_enable(mk_graph *g);
void track_release(mk_graph *g);
void mark_new(mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n, hipStream_t s);
int device_cus(int device);
int64_t tune_get(const char *name, int64_t dflt);
// Start keeping the in-edge mirror's flags of vertices with in-edges (its
// vdeg): set from the mirror's slabs now, then by every insert (queued on
// the graph's stream).
int mirror_flags_enable(mk_graph *rev);
// Captured update steps (mk_step): a mutation queued without host waits,
// and the outcome of its replay from the four counter words it published.
int graph_queue_mutation(mk_graph *g, int op, const uint32_t *src, const uint32_t *dst,
                         const uint32_t *w, uint32_t n, uint32_t *host_dev, uint32_t *epoch);
int graph_replay_outcome(mk_graph *g, int op, const uint32_t *src, const uint32_t *dst,
                         uint32_t n, const uint32_t *words, uint32_t epoch);
// Grid-stride kernels: the number of waves to launch on `device`.
inline int grid_waves(int device) { return device_cus(device) * int(tune_get("waves_per_cu", 16)); }

#define MK_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return ::mk::hip_fail(e_, #call);               \
  } while (0)

// A grow-only device buffer reused across calls (no allocation per call).
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Pinned host scratch for small device->host reads.
struct HostBuf {
  void *ptr = nullptr;
  void *dev = nullptr;  // the same memory as kernels address it
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace mk

struct mk_pool {
  int device = 0;
  mk::PoolView view{};
  uint64_t total_slabs = 0;
  int graphs = 0;
  // per pool slab, the mask of its words holding keys inserted since the
  // owning graph's last UpdateSlabPointers (allocated with the first graph
  // created with MK_GRAPH_TRACK_UPDATES; updated.hip)
  uint64_t *slab_new = nullptr;
};

struct mk_graph {
  mk_pool *pool = nullptr;
  int device = 0;
  uint32_t n = 0;
  int weighted = 0;
  double load_factor = 0.7;
  uint32_t total_buckets = 0;
  uint32_t *heads = nullptr;
  uint32_t *bucket_off = nullptr;
  uint32_t *bucket_owner = nullptr;
  uint32_t *vdeg = nullptr;
  // the in-edge mirror's vdeg holds "ever given an in-edge" flags, kept up
  // to date by its inserts, once a bottom-up BFS handle asked for them
  // (mirror_flags_enable); until then the mirror keeps no per-vertex array
  bool in_flags = false;
  // captured update steps (mk_step) on this graph: at most one, since its
  // replays use the graph's mutation epoch base (counters[1])
  uint32_t steps = 0;
  uint32_t *counters = nullptr;  // [0] range-check epoch [2..3] pool failures (u64)
  uint32_t mut_epoch = 0;        // epoch of the last mutation call
  hipStream_t own_stream = nullptr;  // mutations given no stream run here
  // deferred mode (mk_graph_set_deferred): a mutation call returns once
  // queued; `pend` describes it until the next call settles it
  bool deferred = false;
  // recorded after a queued mutation given another stream than own_stream
  // (BFS/SSSP calls run on own_stream and wait for it there)
  hipEvent_t mut_ev = nullptr;
  // recorded on own_stream when a mutation is given another stream, which
  // then waits for it: work queued on own_stream (deferred BFS/SSSP calls
  // included) comes first
  hipEvent_t trav_ev = nullptr;
  struct Pending {
    bool active;
    int op;  // 1 insert, 2 delete
    const uint32_t *src, *dst;
    uint32_t n;
    hipStream_t s;
    uint32_t seq;  // written after its counters in host_scratch (word 4)
  } pend = {false, 0, nullptr, nullptr, 0, nullptr, 0};
  std::vector<uint32_t> bucket_off_host;
  mk::DevBuf status;             // uint8 per edge of the last batch
  uint32_t status_n = 0;
  // identity of the last mutation batch, so an algorithm called with the
  // same batch can use its per-edge status (1 = insert, 2 = delete)
  int last_op = 0;
  const void *last_src = nullptr;
  const void *last_dst = nullptr;
  const void *last_w = nullptr;
  // the last batch, when it failed (graph_failed_batch): its identity and
  // error, kept until the next mutation
  struct Failed {
    int code;
    const void *src, *dst;
    uint32_t n;
    std::string msg;
  } failed = {0, nullptr, nullptr, 0, std::string()};
  mk::HostBuf host_scratch;
  int users = 0;                 // algorithm handles attached
  // update-aware iteration (MK_GRAPH_TRACK_UPDATES; updated.hip): per
  // bucket, the new-key mask of its head slab and an "inserted into" flag
  bool track = false;
  uint64_t *head_new = nullptr;
  uint8_t *bucket_upd = nullptr;
  mk_graph *rev = nullptr;       // in-edge mirror (MK_GRAPH_REVERSE), or null
  mk_graph *mirror_of = nullptr; // set on a mirror: the graph that owns it
  mk::GraphView view() const {
    mk::GraphView v;
    v.heads = heads;
    v.pool = pool->view.slabs;
    v.bucket_off = bucket_off;
    v.bucket_owner = bucket_owner;
    v.vdeg = vdeg;
    v.n = n;
    v.total_buckets = total_buckets;
    v.weighted = uint32_t(weighted);
    return v;
  }
};

// Launch-shape helpers.
namespace mk {
constexpr int kBlock = 256;  // 4 waves
inline unsigned blocks_for_threads(uint64_t threads) {
  uint64_t b = (threads + kBlock - 1) / kBlock;
  if (b == 0) b = 1;
  return unsigned(b);
}
inline unsigned blocks_for_waves(uint64_t waves) { return blocks_for_threads(waves * 64); }
// Grid-stride kernels bound by the latency of independent random loads
// (more waves keep more in flight): `want` waves per CU, capped at what a CU
// holds of `kernel` in kBlock-thread blocks, so that no part of the grid
// waits for a second round.  Sets *waves to the grid's waves on `device`.
inline hipError_t resident_waves(int device, const void *kernel, int want, int *waves) {
  int per_cu = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0);
  if (e != hipSuccess) return e;
  const int fit = (per_cu > 0 ? per_cu : 1) * (kBlock / 64);
  *waves = device_cus(device) * (want < fit ? want : fit);
  return hipSuccess;
}

// Device-wide fill (used instead of hipMemsetAsync so callers may capture).
void fill_u32(uint32_t *p, uint32_t v, uint64_t count, hipStream_t s);
void fill_u64(uint64_t *p, uint64_t v, uint64_t count, hipStream_t s);
void fill_u8(uint8_t *p, uint8_t v, uint64_t count, hipStream_t s);
// Copy n device words into pinned host memory by a kernel (`dst_host_dev`:
// HostBuf::dev), then `seq` into word n, after the others are visible.  A
// hipMemcpyAsync to pinned memory runs as a copy kernel, and an event record
// between kernels, each measured to leave ~10 us of device idle; the host
// learns of the kernel's completion from the sequence word instead
// (host_wait_word).
void publish_words(const uint32_t *src, uint32_t *dst_host_dev, uint32_t n, uint32_t seq,
                   hipStream_t s);
// Wait until the pinned word `*word` equals `want` (written by a publish
// kernel queued on `s`): spin briefly, then synchronise the stream.
int host_wait_word(const volatile uint32_t *word, uint32_t want, hipStream_t s);
}  // namespace mk
