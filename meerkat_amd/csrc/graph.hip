// Slab pool, dynamic graph construction and batched insert / delete /
// search kernels, plus the graph inspection calls used by tests.
//
// Reference: include/Graph/SlabGraph.cuh (DynamicGraph ctor :382-466 and
// :633-717), include/Graph/SlabGraphImpl.cuh (:6-62), device/Insert-
// EdgeKernel.cuh, device/DeleteEdgeKernel.cuh, warp/*.cuh, and the absent
// SlabHash submodule whose contract is restated in slabhash.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <mutex>
#include <string>

#include "profile.hpp"
#include "slabhash.hpp"

namespace mk {

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return MK_ERR_HIP;
}

int device_cus(int device) {
  static int cached[64] = {0};
  if (device < 0 || device >= 64) return 256;
  if (cached[device] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        v <= 0)
      v = 256;
    cached[device] = v;
  }
  return cached[device];
}

// ---------------------------------------------------------------- tuning
// Launch-shape knobs read when a handle is created (see mk_tune_set).
static std::mutex g_tune_mu;
static std::map<std::string, int64_t> g_tune = {
    {"waves_per_cu", 16},   // resident waves per CU the grid-stride kernels ask for
    {"ms_unroll", 8},       // neighbour rows in flight per wave in multi-source SSSP
    {"ms_waves_per_cu", 24},  // waves per CU of the multi-source BFS / SSSP steps
    {"pr_waves_per_cu", 32},  // waves per CU of the PageRank iteration
    {"orphan_repair", 0},   // decremental with the in-edge mirror re-parents orphans first
    {"persistent_steps", 1},  // BFS/SSSP frontier steps in one launch with grid barriers
    {"tail_enter", 128},    // frontiers up to this size run in one block's LDS (0: never)
    {"coop_launch", 0},     // 1: frontier loops always launched cooperatively
    {"mut_waves_per_cu", 32},  // waves per CU the mutation kernels' batch is spread over
    {"barrier_timeout_ms", 5000},  // longest grid-barrier wait of a frontier loop
    // static BFS over a graph with the in-edge mirror: bottom-up steps once
    // the frontier holds more than 1/pull_enter of the buckets (0: never),
    // until it holds fewer than 1/pull_exit of the vertices
    {"pull_enter", 14},
    {"pull_exit", 24},
    // static SSSP: near-far with this bucket width (0: label-correcting
    // without deferral); read when an SSSP handle is created
    {"sssp_delta", 0},
    {"sssp_far_cap", 0},  // entries per near-far list (0: one per vertex)
};

int64_t tune_get(const char *name, int64_t dflt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(name);
  return it == g_tune.end() ? dflt : it->second;
}

int DevBuf::ensure(size_t need) {
  if (need <= bytes && ptr) return MK_OK;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  size_t want = std::max<size_t>(need, 256);
  MK_HIP(hipMalloc(&ptr, want));
  bytes = want;
  return MK_OK;
}
void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}
int HostBuf::ensure(size_t need) {
  if (need <= bytes && ptr) return MK_OK;
  if (ptr) (void)hipHostFree(ptr);
  ptr = dev = nullptr;
  bytes = 0;
  size_t want = std::max<size_t>(need, 256);
  MK_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
  std::memset(ptr, 0, want);
  MK_HIP(hipHostGetDevicePointer(&dev, ptr, 0));
  bytes = want;
  return MK_OK;
}
void HostBuf::release() {
  if (ptr) (void)hipHostFree(ptr);
  ptr = dev = nullptr;
  bytes = 0;
}

// ---------------------------------------------------------------- fills
template <typename T> __global__ void k_fill(T *p, T v, uint64_t count) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}
static unsigned fill_blocks(uint64_t count) {
  uint64_t b = (count + kBlock - 1) / kBlock;
  return unsigned(std::min<uint64_t>(std::max<uint64_t>(b, 1), 65536));
}
void fill_u32(uint32_t *p, uint32_t v, uint64_t count, hipStream_t s) {
  if (count) k_fill<uint32_t><<<fill_blocks(count), kBlock, 0, s>>>(p, v, count);
}
void fill_u64(uint64_t *p, uint64_t v, uint64_t count, hipStream_t s) {
  if (count) k_fill<uint64_t><<<fill_blocks(count), kBlock, 0, s>>>(p, v, count);
}
void fill_u8(uint8_t *p, uint8_t v, uint64_t count, hipStream_t s) {
  if (count) k_fill<uint8_t><<<fill_blocks(count), kBlock, 0, s>>>(p, v, count);
}

// Device words copied straight into pinned host memory (`dst`: its device
// alias) and made visible to the host (system-scope fence); then `seq` in
// dst[n], so a host that reads seq there also reads the words.
__global__ void k_publish_words(const uint32_t *src, uint32_t *dst, uint32_t n, uint32_t seq) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    dst[n] = seq;
    __threadfence_system();
  }
}
void publish_words(const uint32_t *src, uint32_t *dst_host_dev, uint32_t n, uint32_t seq,
                   hipStream_t s) {
  k_publish_words<<<1, 64, 0, s>>>(src, dst_host_dev, n, seq);
}
int host_wait_word(const volatile uint32_t *word, uint32_t want, hipStream_t s) {
  for (uint32_t i = 0; i < (1u << 22); ++i) {  // tens of ms
    if (*word == want) return MK_OK;
    __builtin_ia32_pause();
  }
  MK_HIP(hipStreamSynchronize(s));  // a longer call: block instead
  if (*word == want) return MK_OK;
  return fail(MK_ERR_HIP, "a counter readback did not arrive");
}

// ---------------------------------------------------------------- kernels
// owner[b] = the vertex whose bucket range holds b (upper_bound - 1).
__global__ void k_bucket_owner(const uint32_t *off, uint32_t n, uint32_t total, uint32_t *owner) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= total) return;
  uint32_t lo = 0, hi = n;  // find largest v with off[v] <= b
  while (hi - lo > 1) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  owner[b] = lo;
}

// Mutation counters are versioned by the call's epoch instead of being
// cleared before each call (no clearing launch): counters[0] holds the
// epoch of the last batch that failed the range check; counters[2..3] (one
// 64-bit word) hold (epoch << 32) | slab-pool failures of that epoch.
// counters[1] is added to every epoch a kernel is given: zero for calls
// made through the API, the replay's base while a captured update step
// (mk_step) runs, so each replay has epochs of its own.

// Range check of a batch: counters[0] = epoch when an endpoint is >= nv.
__global__ void k_validate(const uint32_t *src, const uint32_t *dst, uint32_t n, uint32_t nv,
                           uint32_t *counters, uint32_t epoch) {
  epoch += counters[1];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n) bad = src[i] >= nv || dst[i] >= nv;
  uint64_t m = ballot(bad);
  if (m && lane_id() == ffs64(m)) counters[0] = epoch;
}

// Edges of a batch with an endpoint >= nv, added to *count (error path).
__global__ void k_count_bad(const uint32_t *src, const uint32_t *dst, uint32_t n, uint32_t nv,
                            uint32_t *count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n) bad = src[i] >= nv || dst[i] >= nv;
  uint64_t m = ballot(bad);
  if (m && lane_id() == ffs64(m)) atomicAdd(count, uint32_t(__popcll(m)));
}

// Add `add` to the (epoch << 32) | count word at p, restarting the count
// when the word holds an older epoch.
MK_DEV void add_epoch_count(unsigned long long *p, uint32_t epoch, uint32_t add) {
  unsigned long long old = coherent_load64(reinterpret_cast<const uint64_t *>(p));
  for (;;) {
    const unsigned long long nv = uint32_t(old >> 32) == epoch
                                      ? old + add
                                      : ((unsigned long long)epoch << 32) | add;
    const unsigned long long seen = atomicCAS(p, old, nv);
    if (seen == old) return;
    old = seen;
  }
}

// One side of a batch mutation: the graph it applies to, that graph's
// per-edge status and counters, and whether it takes (dst, src) (the in-edge
// mirror) instead of (src, dst).  Kernels pick a side field by field
// (pick_side): selecting a reference to one of two by-value kernel arguments
// would copy both to private memory.
struct MutSide {
  GraphView G;
  uint8_t *status;
  uint32_t *counters;
  uint32_t swap;
  // G.vdeg upkeep: 1 = live out-degree (atomic add / subtract); 2 = the
  // in-edge mirror's "ever inserted into" flag (set to 1 on an applied
  // insert, never cleared: a superset of the vertices with in-edges, read by
  // bottom-up BFS steps to skip vertices without any; the mirror's live
  // in-degrees are counted on request)
  uint32_t degrees;
};

MK_DEV MutSide pick_side(bool second, const MutSide &A, const MutSide &B) {
  MutSide m;
  m.G.heads = second ? B.G.heads : A.G.heads;
  m.G.pool = second ? B.G.pool : A.G.pool;
  m.G.bucket_off = second ? B.G.bucket_off : A.G.bucket_off;
  m.G.bucket_owner = second ? B.G.bucket_owner : A.G.bucket_owner;
  m.G.vdeg = second ? B.G.vdeg : A.G.vdeg;
  m.G.n = second ? B.G.n : A.G.n;
  m.G.total_buckets = second ? B.G.total_buckets : A.G.total_buckets;
  m.G.weighted = second ? B.G.weighted : A.G.weighted;
  m.status = second ? B.status : A.status;
  m.counters = second ? B.counters : A.counters;
  m.swap = second ? B.swap : A.swap;
  m.degrees = second ? B.degrees : A.degrees;
  return m;
}

// Waves [0, waves_a) apply the batch to side A, the rest to side B, so the
// forward graph and its mirror are updated by one launch with their
// (latency-bound) chain walks overlapping.  The batch is cut into chunks of
// csz = ceil(epw / 8) consecutive edges (epw <= 64 edges per wave); group q
// of a side (wave * 8 + group) takes chunk q * stride mod nchunks (stride
// coprime with nchunks: every chunk once), group lane i holding the chunk's
// edge i, and applies its edges one after another, walking and claiming in
// its own slab chain (the reference runs one edge at a time per warp,
// include/Graph/warp/InsertEdge.cuh:66-103).  Consecutive edges, which in a
// batch sorted by source share buckets, so stay with one group, and the
// groups in flight together work far-apart chunks instead of racing for
// the same free slots.
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_insert(MutSide A, MutSide B, uint32_t waves_a, uint32_t waves_b, PoolView P,
             const uint32_t *src, const uint32_t *dst, const uint32_t *wt, uint32_t n,
             uint32_t epw, uint32_t stride, const uint32_t *bad, uint32_t epoch) {
  const uint32_t lane = lane_id();
  const uint32_t gl = group_lane();
  const uint32_t wave_all = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const bool second = wave_all >= waves_a;
  // waves past both sides (the last block's spare waves) have nothing to do
  if (second && wave_all - waves_a >= waves_b) return;
  const MutSide M = pick_side(second, A, B);
  const uint32_t wave = second ? wave_all - waves_a : wave_all;
  const uint32_t csz = (epw + kGroupsPerWave - 1) / kGroupsPerWave;
  const uint64_t nchunks = (uint64_t(n) + csz - 1) / csz;
  const uint64_t q = uint64_t(wave) * kGroupsPerWave;
  if (q >= nchunks) return;
  uint64_t base = 0;
  uint32_t cnt = 0;  // this group's edges
  if (q + (lane >> 3) < nchunks) {
    base = ((q + (lane >> 3)) * stride % nchunks) * csz;
    cnt = uint32_t(min<uint64_t>(csz, n - base));
  }
  epoch += bad[1];  // the replay base of a captured step (0 otherwise)
  if (*bad == epoch) {  // the batch failed the range check: leave the graph as it is
    if (gl < cnt) M.status[base + gl] = MK_EDGE_UNCHANGED;
    return;
  }
  uint32_t s = 0, d = 0, w = 0;
  bool act = false;
  uint8_t st = MK_EDGE_UNCHANGED;
  if (gl < cnt) {
    s = src[base + gl];
    d = dst[base + gl];
    if (M.swap) {
      const uint32_t t = s;
      s = d;
      d = t;
    }
    w = (MAP && wt) ? wt[base + gl] : 0u;
    act = s != d;
    if (!act) st = MK_EDGE_SELF_LOOP;
  }
  const GraphView &G = M.G;
  // every lane's bucket at once, off the groups' sequential chains
  const uint32_t bk = act ? bucket_of(G, s, d) : 0u;
  uint32_t pending = group_bits(ballot(act));
  uint32_t pool_fail = 0;
  const uint32_t seed_base = (second ? 0x9E3779B9u : 0u) + wave * 64u + group_first_lane();
  // head slabs with plain loads: an older view of a slab is safe (a claim
  // made from it fails, and the walk is redone from agent-scope loads)
  uint32_t cur[kWordsPerLane];
  if (pending) load_slab_plain(G.heads + size_t(group_read(bk, ffs32(pending))) * kSlabWords, cur);
  while (pending) {
    const uint32_t j = ffs32(pending);
    pending &= pending - 1;
    // the group's next edge's head slab is in flight while this edge's
    // claim is
    uint32_t nxt[kWordsPerLane] = {};
    if (pending) load_slab_plain(G.heads + size_t(group_read(bk, ffs32(pending))) * kSlabWords, nxt);
    const uint32_t sj = group_read(s, j), dj = group_read(d, j), wj = group_read(w, j);
    const int r = slab_insert_group<MAP>(G, P, group_read(bk, j), dj, wj, seed_base + j, cur);
#pragma unroll
    for (uint32_t k = 0; k < kWordsPerLane; ++k) cur[k] = nxt[k];
    if (gl == j) {
      if (r == 1) st = MK_EDGE_APPLIED;
      else if (r < 0) st = MK_EDGE_POOL_FULL;  // rolled back by the host
    }
    if (r == 1) {
      if (gl == 0 && M.degrees == 1) atomicAdd(G.vdeg + sj, 1u);
      else if (gl == 0 && M.degrees == 2) G.vdeg[sj] = 1u;
    } else if (r < 0) {
      ++pool_fail;
    }
  }
  if (gl < cnt) M.status[base + gl] = st;
  // pool failures (rare): one update per group that had any
  if (gl == 0 && pool_fail)
    add_epoch_count(reinterpret_cast<unsigned long long *>(M.counters + 2), epoch, pool_fail);
}

// Delete the batch edges whose status equals `only_status` (all edges when
// only_status == 0xFF); records MK_EDGE_APPLIED / MK_EDGE_UNCHANGED.  Sides,
// chunks and groups as in k_insert (in a batch sorted by source, groups
// in flight together would otherwise swap words of the same slabs).
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_delete(MutSide A, MutSide B, uint32_t waves_a, uint32_t waves_b, const uint32_t *src,
             const uint32_t *dst, uint32_t n, uint32_t epw, uint32_t stride,
             uint32_t only_status, const uint32_t *bad, uint32_t epoch) {
  const uint32_t lane = lane_id();
  const uint32_t gl = group_lane();
  const uint32_t wave_all = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const bool second = wave_all >= waves_a;
  // waves past both sides (the last block's spare waves) have nothing to do
  if (second && wave_all - waves_a >= waves_b) return;
  const MutSide M = pick_side(second, A, B);
  const uint32_t wave = second ? wave_all - waves_a : wave_all;
  const uint32_t csz = (epw + kGroupsPerWave - 1) / kGroupsPerWave;
  const uint64_t nchunks = (uint64_t(n) + csz - 1) / csz;
  const uint64_t q = uint64_t(wave) * kGroupsPerWave;
  if (q >= nchunks) return;
  uint64_t base = 0;
  uint32_t cnt = 0;  // this group's edges
  if (q + (lane >> 3) < nchunks) {
    base = ((q + (lane >> 3)) * stride % nchunks) * csz;
    cnt = uint32_t(min<uint64_t>(csz, n - base));
  }
  epoch += bad[1];  // the replay base of a captured step (0 otherwise)
  if (*bad == epoch) {  // the batch failed the range check: leave the graph as it is
    if (gl < cnt && only_status == 0xFFu) M.status[base + gl] = MK_EDGE_UNCHANGED;
    return;
  }
  uint32_t s = 0, d = 0;
  bool act = false;
  uint8_t st = MK_EDGE_UNCHANGED;
  if (gl < cnt) {
    s = src[base + gl];
    d = dst[base + gl];
    if (M.swap) {
      const uint32_t t = s;
      s = d;
      d = t;
    }
    if (only_status != 0xFFu) {
      act = M.status[base + gl] == only_status;
    } else {
      act = s != d;
      if (!act) st = MK_EDGE_SELF_LOOP;
    }
  }
  const GraphView &G = M.G;
  const uint32_t bk = act ? bucket_of(G, s, d) : 0u;
  uint32_t pending = group_bits(ballot(act));
  uint32_t cur[kWordsPerLane];
  if (pending) load_slab_plain(G.heads + size_t(group_read(bk, ffs32(pending))) * kSlabWords, cur);
  while (pending) {
    const uint32_t j = ffs32(pending);
    pending &= pending - 1;
    // the next edge's head slab is in flight while this edge's swap is
    uint32_t nxt[kWordsPerLane] = {};
    if (pending) load_slab_plain(G.heads + size_t(group_read(bk, ffs32(pending))) * kSlabWords, nxt);
    const uint32_t sj = group_read(s, j), dj = group_read(d, j);
    const int r = slab_delete_group<MAP>(G, group_read(bk, j), dj, cur);
#pragma unroll
    for (uint32_t k = 0; k < kWordsPerLane; ++k) cur[k] = nxt[k];
    if (gl == j && r == 1) st = MK_EDGE_APPLIED;
    if (r == 1 && gl == 0 && M.degrees == 1) atomicSub(G.vdeg + sj, 1u);
  }
  if (only_status != 0xFFu) {
    if (gl < cnt && act) M.status[base + gl] = MK_EDGE_UNCHANGED;
  } else if (gl < cnt) {
    M.status[base + gl] = st;
  }
}

// Weight update of a map graph (and of its mirror's entries): every batch
// edge present gets the batch's weight (MK_EDGE_APPLIED); an absent edge is
// not inserted (MK_EDGE_UNCHANGED); self-loops are skipped, as the
// reference's UpdateEdgesKernel does (device/UpdateEdgeKernel.cuh:28).
// Sides as in k_insert; each wave takes `epw` consecutive edges, lane i
// holding edge i, and its eight groups apply their own lanes' edges one
// after another.
__global__ void __launch_bounds__(kBlock)
    k_update_weights(MutSide A, MutSide B, uint32_t waves_a, uint32_t waves_b,
                     const uint32_t *src, const uint32_t *dst, const uint32_t *wt, uint32_t n,
                     uint32_t epw, const uint32_t *bad, uint32_t epoch) {
  const uint32_t lane = lane_id();
  const uint32_t gl = group_lane();
  const uint32_t wave_all = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const bool second = wave_all >= waves_a;
  if (second && wave_all - waves_a >= waves_b) return;
  const MutSide M = pick_side(second, A, B);
  const uint32_t wave = second ? wave_all - waves_a : wave_all;
  const uint64_t base = uint64_t(wave) * epw;
  if (base >= n) return;
  const uint32_t cnt = uint32_t(min<uint64_t>(epw, n - base));
  epoch += bad[1];  // the replay base of a captured step (0 otherwise)
  if (*bad == epoch) {  // the batch failed the range check: leave the graph as it is
    if (lane < cnt) M.status[base + lane] = MK_EDGE_UNCHANGED;
    return;
  }
  uint32_t s = 0, d = 0, w = 0;
  bool act = false;
  uint8_t st = MK_EDGE_UNCHANGED;
  if (lane < cnt) {
    s = src[base + lane];
    d = dst[base + lane];
    if (M.swap) {
      const uint32_t t = s;
      s = d;
      d = t;
    }
    w = wt[base + lane];
    act = s != d;
    if (!act) st = MK_EDGE_SELF_LOOP;
  }
  const GraphView &G = M.G;
  const uint32_t bk = act ? bucket_of(G, s, d) : 0u;
  uint32_t pending = group_bits(ballot(act));
  while (pending) {
    const uint32_t j = ffs32(pending);
    pending &= pending - 1;
    const int r = slab_update_group(G, group_read(bk, j), group_read(d, j), group_read(w, j));
    if (gl == j && r == 1) st = MK_EDGE_APPLIED;
  }
  if (lane < cnt) M.status[base + lane] = st;
}

template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_search(GraphView G, const uint32_t *src, const uint32_t *dst, uint32_t n, uint32_t epw,
             uint8_t *found, uint32_t *wout) {
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t base = uint64_t(wave) * epw;
  if (base >= n) return;
  const uint32_t cnt = uint32_t(min<uint64_t>(epw, n - base));
  uint32_t s = 0, d = 0;
  bool act = false;
  if (lane < cnt) {
    s = src[base + lane];
    d = dst[base + lane];
    act = s < G.n && d < G.n;
  }
  uint8_t f = 0;
  uint32_t wv = 0;
  const uint32_t gl = group_lane();
  const uint32_t bk = act ? bucket_of(G, s, d) : 0u;
  uint32_t pending = group_bits(ballot(act));
  while (pending) {
    const uint32_t j = ffs32(pending);
    pending &= pending - 1;
    uint32_t wj = 0;
    const int r = slab_search_group<MAP>(G, group_read(bk, j), group_read(d, j), &wj);
    if (gl == j) {
      f = uint8_t(r);
      wv = wj;
    }
  }
  if (lane < cnt) {
    found[base + lane] = f;
    if (wout) wout[base + lane] = f ? wv : kEmpty;
  }
}

// Per-bucket walk, one 8-lane group per bucket (grid-stride, eight chains in
// flight per wave): live keys and tombstones added to the owner's counters;
// the overflow slab count and the longest chain gathered per wave and added
// with one atomic each (an atomic per bucket on one counter serialises: 0.3 s
// over RMAT-24's 21M buckets).
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_count_slots(GraphView G, uint32_t *live, uint32_t *tomb, unsigned long long *overflow,
                  uint32_t *max_chain) {
  constexpr uint32_t kGroups = 64 / kGroupLanes;
  const uint32_t gl = group_lane();
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t W = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  uint64_t ovf = 0;
  uint32_t mx = 0;
  for (uint64_t b0 = gw * kGroups; b0 < G.total_buckets; b0 += W * kGroups) {
    const uint64_t b = b0 + lane_id() / kGroupLanes;
    if (b >= G.total_buckets) continue;  // the whole group
    const uint32_t *slab = G.heads + b * kSlabWords;
    uint32_t nl = 0, nt = 0, chain = 1;
    for (;;) {
      uint32_t v[kWordsPerLane];
      load_words(slab + kWordsPerLane * gl, v);
#pragma unroll
      for (uint32_t k = 0; k < kWordsPerLane; ++k)
        if (key_word<MAP>(kWordsPerLane * gl + k)) {
          nl += v[k] < kTomb ? 1u : 0u;
          nt += v[k] == kTomb ? 1u : 0u;
        }
      const uint32_t nx = group_read(v[kWordsPerLane - 1], kGroupLanes - 1);
      if (nx == kEmpty) break;
      slab = slab_ptr(G.pool, nx);
      ++chain;
    }
#pragma unroll
    for (uint32_t o = kGroupLanes / 2; o >= 1; o >>= 1) {  // sums over the group
      nl += uint32_t(__shfl_xor(int(nl), int(o), 64));
      nt += uint32_t(__shfl_xor(int(nt), int(o), 64));
    }
    if (gl == 0) {
      const uint32_t u = G.bucket_owner[b];
      if (live && nl) atomicAdd(live + u, nl);
      if (tomb && nt) atomicAdd(tomb + u, nt);
      ovf += chain - 1;
      mx = max(mx, chain);
    }
  }
  ovf = wave_sum_u64(ovf);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = max(mx, uint32_t(__shfl_xor(int(mx), o, 64)));
  if (lane_id() == 0) {
    if (overflow && ovf) atomicAdd(overflow, (unsigned long long)ovf);
    if (max_chain && mx) atomicMax(max_chain, mx);
  }
}

// flag[owner] = 1 for every bucket holding a live key (one wave per bucket).
template <bool MAP>
__global__ void __launch_bounds__(kBlock) k_live_flags(GraphView G, uint32_t *flag) {
  const uint32_t lane = lane_id();
  const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= G.total_buckets) return;
  const uint32_t *slab = G.heads + size_t(b) * kSlabWords;
  for (;;) {
    const uint32_t v = slab[lane];
    if (ballot(v < kTomb) & key_mask<MAP>()) {
      if (lane == 0) flag[G.bucket_owner[b]] = 1u;
      return;
    }
    const uint32_t nx = read_lane(v, kNextLane);
    if (nx == kEmpty) return;
    slab = slab_ptr(G.pool, nx);
  }
}

int mirror_flags_enable(mk_graph *rev) {
  if (rev->in_flags) return MK_OK;
  MK_HIP(hipSetDevice(rev->device));
  mk_graph *g = rev->mirror_of;
  hipStream_t s = g ? g->own_stream : nullptr;
  fill_u32(rev->vdeg, 0u, rev->n, s);
  auto k = rev->weighted ? k_live_flags<true> : k_live_flags<false>;
  k<<<blocks_for_waves(rev->total_buckets), kBlock, 0, s>>>(rev->view(), rev->vdeg);
  MK_HIP(hipGetLastError());
  MK_HIP(hipStreamSynchronize(s));
  // every later insert keeps the flags
  rev->in_flags = true;
  return MK_OK;
}

// Write each bucket's live keys (and values) at cursor[owner]++.
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_export(GraphView G, unsigned long long *cursor, uint32_t *col, uint32_t *wcol) {
  const uint32_t lane = lane_id();
  const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= G.total_buckets) return;
  const uint64_t km = key_mask<MAP>();
  const uint32_t u = G.bucket_owner[b];
  const uint32_t *slab = G.heads + size_t(b) * kSlabWords;
  for (;;) {
    const uint32_t v = slab[lane];
    const uint32_t val = MAP ? __shfl_down(v, 1, 64) : 0u;
    const bool live = ((km >> lane) & 1ull) && v < kTomb;
    const uint64_t m = ballot(live);
    if (m) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(cursor + u, (unsigned long long)__popcll(m));
      base = __shfl(base, 0, 64);
      if (live) {
        const uint64_t pos = base + lanes_below(m);
        col[pos] = v;
        if (MAP && wcol) wcol[pos] = val;
      }
    }
    const uint32_t nx = read_lane(v, kNextLane);
    if (nx == kEmpty) break;
    slab = slab_ptr(G.pool, nx);
  }
}

// Return every overflow slab of every chain to the pool.
__global__ void __launch_bounds__(kBlock) k_release(GraphView G, PoolView P) {
  const uint32_t lane = lane_id();
  const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= G.total_buckets) return;
  uint32_t nx = G.heads[size_t(b) * kSlabWords + kNextLane];
  while (nx != kEmpty) {
    const uint32_t next = slab_ptr(G.pool, nx)[kNextLane];
    if (lane == 0) pool_free_lane(P, nx);
    nx = next;
  }
}

// A multiplier coprime with nchunks near 0.618 nchunks: q -> q * stride mod
// nchunks visits every chunk once and sends neighbouring q far apart.
static uint32_t chunk_stride(uint64_t nchunks) {
  if (nchunks <= 2) return 1;
  uint64_t p = uint64_t(double(nchunks) * 0.6180339887) | 1u;
  while (std::gcd(p, nchunks) != 1) p += 2;
  return uint32_t(p % nchunks);
}

static uint32_t edges_per_wave(uint32_t n, int device) {
  // enough waves to fill every CU `mut_waves_per_cu` deep (default 32; eight
  // slab operations in flight per wave), at most 64 edges (one per lane)
  const uint64_t target_waves = uint64_t(device_cus(device)) * uint64_t(tune_get("mut_waves_per_cu", 32));
  uint64_t epw = (n + target_waves - 1) / target_waves;
  return uint32_t(std::min<uint64_t>(std::max<uint64_t>(epw, 1), 64));
}

}  // namespace mk

using namespace mk;

// ================================================================ C ABI
extern "C" {

const char *mk_last_error(void) { return g_err.c_str(); }
int mk_version(void) { return 100; }

int mk_tune_set(const char *name, int64_t value) {
  if (!name) return fail(MK_ERR_BAD_ARG, "mk_tune_set: NULL name");
  const std::string k(name);
  if (k == "waves_per_cu" && (value < 1 || value > 64))
    return fail(MK_ERR_BAD_ARG, "waves_per_cu must be in [1, 64]");
  if ((k == "ms_waves_per_cu" || k == "pr_waves_per_cu") && (value < 1 || value > 64))
    return fail(MK_ERR_BAD_ARG, k + " must be in [1, 64]");
  if (k == "ms_unroll" && value != 4 && value != 8)
    return fail(MK_ERR_BAD_ARG, "ms_unroll must be 4 or 8");
  if ((k == "orphan_repair" || k == "persistent_steps" || k == "coop_launch") && value != 0 &&
      value != 1)
    return fail(MK_ERR_BAD_ARG, k + " must be 0 or 1");
  if (k == "barrier_timeout_ms" && (value < 0 || value > 600000))
    return fail(MK_ERR_BAD_ARG, "barrier_timeout_ms must be in [0, 600000]");
  if (k == "mut_waves_per_cu" && (value < 1 || value > 256))
    return fail(MK_ERR_BAD_ARG, "mut_waves_per_cu must be in [1, 256]");
  if (k == "tail_enter" && (value < 0 || value > 1024))
    return fail(MK_ERR_BAD_ARG, "tail_enter must be in [0, 1024]");
  if ((k == "pull_enter" || k == "pull_exit") && (value < 0 || value > 4096))
    return fail(MK_ERR_BAD_ARG, k + " must be in [0, 4096]");
  if (k == "pull_exit" && value == 0) return fail(MK_ERR_BAD_ARG, "pull_exit must be >= 1");
  if (k == "sssp_far_cap" && (value < 0 || value > 0xFFFFFFF0ll))
    return fail(MK_ERR_BAD_ARG, "sssp_far_cap must be in [0, 4294967280]");
  if (k == "sssp_delta" && (value < 0 || value > (1 << 24)))
    return fail(MK_ERR_BAD_ARG, "sssp_delta must be in [0, 16777216]");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(k);
  if (it == g_tune.end()) return fail(MK_ERR_BAD_ARG, "mk_tune_set: unknown knob " + k);
  it->second = value;
  return MK_OK;
}

int mk_tune_get(const char *name, int64_t *value) {
  if (!name || !value) return fail(MK_ERR_BAD_ARG, "mk_tune_get: NULL argument");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(name);
  if (it == g_tune.end()) return fail(MK_ERR_BAD_ARG, std::string("unknown knob ") + name);
  *value = it->second;
  return MK_OK;
}

int mk_device_info(int device, int *compute_units, int *wave_size) {
  int cu = 0, ws = 0;
  MK_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device));
  MK_HIP(hipDeviceGetAttribute(&ws, hipDeviceAttributeWarpSize, device));
  if (compute_units) *compute_units = cu;
  if (wave_size) *wave_size = ws;
  return MK_OK;
}

int mk_pool_create(uint64_t bytes, int device, mk_pool **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "mk_pool_create: out is NULL");
  *out = nullptr;
  if (bytes == 0) bytes = 1ull << 30;
  uint64_t slabs = bytes / (kSlabWords * 4);
  slabs = (slabs / kBlockSlabs) * kBlockSlabs;
  if (slabs == 0) slabs = kBlockSlabs;
  if (slabs > 0xFFFF0000ull) return fail(MK_ERR_BAD_ARG, "mk_pool_create: pool above 2^32 slabs");
  int wave = 0;
  MK_HIP(hipSetDevice(device));
  MK_HIP(hipDeviceGetAttribute(&wave, hipDeviceAttributeWarpSize, device));
  if (wave != 64) return fail(MK_ERR_HIP, "mk_pool_create: device wavefront is not 64 lanes");
  mk_pool *p = new mk_pool();
  p->device = device;
  p->total_slabs = slabs;
  p->view.nblocks = uint32_t(slabs / kBlockSlabs);
  p->view.total_slabs = uint32_t(slabs);
  hipError_t e = hipMalloc(&p->view.slabs, slabs * kSlabWords * 4);
  if (e == hipSuccess) e = hipMalloc(&p->view.bitmap, slabs / 8);
  if (e == hipSuccess) e = hipMalloc(&p->view.counters, 64);
  if (e != hipSuccess) {
    (void)hipFree(p->view.slabs);
    (void)hipFree(p->view.bitmap);
    delete p;
    return hip_fail(e, "mk_pool_create: hipMalloc");
  }
  fill_u32(p->view.bitmap, 0u, slabs / 32, nullptr);
  fill_u32(p->view.counters, 0u, 16, nullptr);
  MK_HIP(hipDeviceSynchronize());
  *out = p;
  return MK_OK;
}

int mk_pool_destroy(mk_pool *p) {
  if (!p) return fail(MK_ERR_BAD_ARG, "mk_pool_destroy: NULL pool");
  if (p->graphs) return fail(MK_ERR_STATE, "mk_pool_destroy: graphs still use this pool");
  (void)hipSetDevice(p->device);
  (void)hipFree(p->view.slabs);
  (void)hipFree(p->view.bitmap);
  (void)hipFree(p->view.counters);
  (void)hipFree(p->slab_new);
  delete p;
  return MK_OK;
}

int mk_pool_stats(mk_pool *p, uint64_t *used, uint64_t *total) {
  if (!p) return fail(MK_ERR_BAD_ARG, "mk_pool_stats: NULL pool");
  uint32_t c[2] = {0, 0};
  MK_HIP(hipSetDevice(p->device));
  MK_HIP(hipMemcpy(c, p->view.counters, sizeof(c), hipMemcpyDeviceToHost));
  if (used) *used = c[0];
  if (total) *total = p->total_slabs;
  return MK_OK;
}

static int graph_alloc(mk_pool *pool, uint32_t n, int weighted, double load_factor,
                       const uint32_t *hints, mk_graph **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "graph create: out is NULL");
  *out = nullptr;
  if (!pool) return fail(MK_ERR_BAD_ARG, "graph create: NULL pool");
  if (n == 0 || n >= kTomb) return fail(MK_ERR_BAD_ARG, "graph create: n out of range");
  if (!(load_factor > 0.0) || load_factor > 64.0)
    return fail(MK_ERR_BAD_ARG, "graph create: load_factor must be in (0, 64]");
  MK_HIP(hipSetDevice(pool->device));
  const double per = (weighted ? kMapKeysPerSlab : kSetKeysPerSlab) * load_factor;
  std::vector<uint32_t> off(size_t(n) + 1);
  uint64_t total = 0;
  for (uint32_t v = 0; v < n; ++v) {
    off[v] = uint32_t(total);
    uint64_t nb = hints ? uint64_t(std::ceil(double(hints[v]) / per)) : 1;
    total += nb ? nb : 1;
    if (total >= 0xFFFFFFF0ull) return fail(MK_ERR_BAD_ARG, "graph create: >2^32 buckets");
  }
  off[n] = uint32_t(total);
  mk_graph *g = new mk_graph();
  g->pool = pool;
  g->device = pool->device;
  g->n = n;
  g->weighted = weighted ? 1 : 0;
  g->load_factor = load_factor;
  g->total_buckets = uint32_t(total);
  hipError_t e = hipMalloc(&g->heads, total * kSlabWords * 4);
  if (e == hipSuccess) e = hipMalloc(&g->bucket_off, (size_t(n) + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&g->bucket_owner, total * 4);
  if (e == hipSuccess) e = hipMalloc(&g->vdeg, size_t(n) * 4);
  if (e == hipSuccess) e = hipMalloc(&g->counters, 64);
  if (e == hipSuccess)
    e = hipMemcpy(g->bucket_off, off.data(), (size_t(n) + 1) * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(g->heads);
    (void)hipFree(g->bucket_off);
    (void)hipFree(g->bucket_owner);
    (void)hipFree(g->vdeg);
    (void)hipFree(g->counters);
    delete g;
    return hip_fail(e, "graph create: allocation");
  }
  fill_u32(g->heads, kEmpty, total * kSlabWords, nullptr);
  fill_u32(g->vdeg, 0u, n, nullptr);
  fill_u32(g->counters, 0u, 16, nullptr);
  k_bucket_owner<<<blocks_for_threads(total), kBlock>>>(g->bucket_off, n, uint32_t(total),
                                                       g->bucket_owner);
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error(std::string("graph create: init: ") + hipGetErrorString(e));
    return MK_ERR_HIP;
  }
  // mutations the caller gives no stream run on this one (default flags, so
  // it still orders after the caller's NULL-stream work): launching on the
  // NULL stream itself makes the runtime order it against every stream
  e = hipStreamCreate(&g->own_stream);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->mut_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->trav_ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipFree(g->heads);
    (void)hipFree(g->bucket_off);
    (void)hipFree(g->bucket_owner);
    (void)hipFree(g->vdeg);
    (void)hipFree(g->counters);
    delete g;
    return hip_fail(e, "graph create: stream");
  }
  g->bucket_off_host.swap(off);
  pool->graphs++;
  *out = g;
  return MK_OK;
}


int mk_graph_create(mk_pool *pool, uint32_t n, int weighted, double load_factor,
                    const uint32_t *hints, mk_graph **out) {
  return graph_alloc(pool, n, weighted ? 1 : 0, load_factor, hints, out);
}

int mk_graph_create_ex(mk_pool *pool, uint32_t n, uint32_t flags, double load_factor,
                       const uint32_t *out_hints, const uint32_t *in_hints, mk_graph **out) {
  if (flags & ~uint32_t(MK_GRAPH_WEIGHTED | MK_GRAPH_REVERSE | MK_GRAPH_TRACK_UPDATES))
    return fail(MK_ERR_BAD_ARG, "mk_graph_create_ex: unknown flag bits");
  const int weighted = (flags & MK_GRAPH_WEIGHTED) ? 1 : 0;
  if (int r = graph_alloc(pool, n, weighted, load_factor, out_hints, out)) return r;
  if (flags & MK_GRAPH_TRACK_UPDATES) {
    if (int r = track_enable(*out)) {
      track_release(*out);
      (void)mk_graph_destroy(*out);
      *out = nullptr;
      return r;
    }
  }
  if (flags & MK_GRAPH_REVERSE) {
    mk_graph *rev = nullptr;
    int r = graph_alloc(pool, n, weighted, load_factor, in_hints ? in_hints : out_hints, &rev);
    if (r) {
      mk_graph_destroy(*out);
      *out = nullptr;
      return r;
    }
    rev->mirror_of = *out;
    (*out)->rev = rev;
  }
  return MK_OK;
}

int mk_graph_reverse(mk_graph *g, mk_graph **rev) {
  if (!g || !rev) return fail(MK_ERR_BAD_ARG, "mk_graph_reverse: NULL argument");
  *rev = g->rev;
  if (!g->rev) return fail(MK_ERR_STATE, "mk_graph_reverse: graph has no in-edge mirror");
  return MK_OK;
}

int mk_graph_destroy(mk_graph *g) {
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_graph_destroy: NULL graph");
  if (g->users) return fail(MK_ERR_STATE, "mk_graph_destroy: algorithm handles still attached");
  if (g->mirror_of) return fail(MK_ERR_STATE, "mk_graph_destroy: an in-edge mirror is owned by its graph");
  (void)graph_settle(g);
  (void)hipSetDevice(g->device);
  if (g->rev) {
    g->rev->mirror_of = nullptr;
    mk_graph_destroy(g->rev);
    g->rev = nullptr;
  }
  k_release<<<blocks_for_waves(g->total_buckets), kBlock>>>(g->view(), g->pool->view);
  (void)hipDeviceSynchronize();
  (void)hipFree(g->heads);
  (void)hipFree(g->bucket_off);
  (void)hipFree(g->bucket_owner);
  (void)hipFree(g->vdeg);
  (void)hipFree(g->counters);
  track_release(g);
  if (g->own_stream) (void)hipStreamDestroy(g->own_stream);
  if (g->mut_ev) (void)hipEventDestroy(g->mut_ev);
  if (g->trav_ev) (void)hipEventDestroy(g->trav_ev);
  g->status.release();
  g->host_scratch.release();
  g->pool->graphs--;
  delete g;
  return MK_OK;
}

// Shared front half of insert / delete: argument checks, range validation.
static int batch_prologue(mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n,
                          hipStream_t s, const char *who) {
  if (!g) return fail(MK_ERR_BAD_ARG, std::string(who) + ": NULL graph");
  if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, std::string(who) + ": NULL batch");
  MK_HIP(hipSetDevice(g->device));
  if (int r = g->status.ensure(std::max<size_t>(n, 1))) return r;
  if (int r = g->host_scratch.ensure(64)) return r;
  g->status_n = n;
  g->last_op = 0;
  g->last_src = g->last_dst = g->last_w = nullptr;
  g->failed.code = 0;
  if (n == 0) return MK_OK;
  if (++g->mut_epoch == 0) {  // epochs wrapped: clear the versioned counters once
    fill_u32(g->counters, 0u, 4, s);
    g->mut_epoch = 1;
  }
  // the range check's verdict (counters[0] == epoch) is read by the mutation
  // kernels, which leave the graph unchanged on a failure; the host reads
  // it with the mutation's results (range_error)
  k_validate<<<blocks_for_threads(n), kBlock, 0, s>>>(src, dst, n, g->n, g->counters,
                                                      g->mut_epoch);
  return MK_OK;
}

// After a mutation's counters were copied to host_scratch and the stream
// synchronised: MK_ERR_RANGE when the batch failed the range check (the
// offending edges are counted for the message).
static int range_error(mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n,
                       hipStream_t s, const char *who) {
  if (g->host_scratch.as<uint32_t>()[0] != g->mut_epoch) return MK_OK;
  // counters[4] counts the offending edges (counters[1] is the epoch base)
  fill_u32(g->counters + 4, 0u, 1, s);
  k_count_bad<<<blocks_for_threads(n), kBlock, 0, s>>>(src, dst, n, g->n, g->counters + 4);
  uint32_t bad = 0;
  MK_HIP(hipMemcpyAsync(&bad, g->counters + 4, 4, hipMemcpyDeviceToHost, s));
  MK_HIP(hipStreamSynchronize(s));
  return fail(MK_ERR_RANGE, std::string(who) + ": " + std::to_string(bad) +
                                " edges have a vertex id >= n; graph unchanged");
}

// Slab-pool failures of this call's mutation kernels (host_scratch read).
static uint32_t pool_failures(mk_graph *g) {
  const uint64_t w = g->host_scratch.as<uint64_t>()[1];
  return uint32_t(w >> 32) == g->mut_epoch ? uint32_t(w) : 0u;
}

// The mutation side of graph g (its own view and status; the counters of
// the graph the batch was given to, `counters`).  The mirror keeps no
// live degree array (mk_graph_degrees counts its live keys instead), only,
// once bottom-up BFS asked for it (mirror_flags_enable), the flag of
// vertices that were ever given an in-edge: a store per applied edge that
// costs the insert ~4% while nothing reads it.
static MutSide side_of(mk_graph *g, bool swap, uint32_t *counters) {
  return MutSide{g->view(), g->status.as<uint8_t>(), counters, swap ? 1u : 0u,
                 g->mirror_of ? (g->in_flags ? 2u : 0u) : 1u};
}

// Launch the insert kernel over g and, when it has one, its mirror (one
// launch, half the waves each).
static void launch_insert(mk_graph *g, const uint32_t *src, const uint32_t *dst,
                          const uint32_t *w, uint32_t n, hipStream_t s, const char *name,
                          const uint32_t *bad) {
  const uint32_t epw = edges_per_wave(n, g->device);
  const uint32_t csz = (epw + kGroupsPerWave - 1) / kGroupsPerWave;
  const uint64_t nchunks = (uint64_t(n) + csz - 1) / csz;
  const uint32_t stride = chunk_stride(nchunks);
  const uint64_t waves = (nchunks + kGroupsPerWave - 1) / kGroupsPerWave;
  const MutSide a = side_of(g, false, g->counters),
                b = g->rev ? side_of(g->rev, true, g->counters) : a;
  const uint64_t total = g->rev ? 2 * waves : waves;
  // per edge: batch entry, one slab probe, key (+value) write, degree atomic,
  // status byte; the mirror side the same without the degree atomic
  const double per = g->weighted ? 12 + 256 + 8 + 8 + 1 : 8 + 256 + 4 + 8 + 1;
  const double bytes = double(n) * (g->rev ? 2 * per - 8 : per);
  prof_launch(name, bytes, g->weighted ? k_insert<true> : k_insert<false>,
              dim3(blocks_for_waves(total)), dim3(kBlock), 0u, s, a, b, uint32_t(waves),
              uint32_t(total - waves), g->pool->view, src, dst, g->weighted ? w : nullptr, n, epw,
              stride, bad, g->mut_epoch);
}

// Launch the delete kernel over g (and its mirror); only_status != 0xFF
// restricts it to edges whose status equals it (roll-back of an insert).
static void launch_delete(mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n,
                          hipStream_t s, uint32_t only_status, const char *name,
                          const uint32_t *bad) {
  const uint32_t epw = edges_per_wave(n, g->device);
  const uint32_t csz = (epw + kGroupsPerWave - 1) / kGroupsPerWave;
  const uint64_t nchunks = (uint64_t(n) + csz - 1) / csz;
  const uint32_t stride = chunk_stride(nchunks);
  const uint64_t waves = (nchunks + kGroupsPerWave - 1) / kGroupsPerWave;
  const MutSide a = side_of(g, false, g->counters),
                b = g->rev ? side_of(g->rev, true, g->counters) : a;
  const uint64_t total = g->rev ? 2 * waves : waves;
  // per edge: batch entry, one slab probe, tombstone write, degree atomic,
  // status byte; the mirror side the same without the degree atomic
  prof_launch(name, double(n) * (g->rev ? 2 * 277 - 8 : 277),
              g->weighted ? k_delete<true> : k_delete<false>, dim3(blocks_for_waves(total)),
              dim3(kBlock), 0u, s, a, b, uint32_t(waves), uint32_t(total - waves), src, dst, n,
              epw, stride, only_status, bad, g->mut_epoch);
}

// Launch the weight-update kernel over g (and its mirror).
static void launch_update_weights(mk_graph *g, const uint32_t *src, const uint32_t *dst,
                                  const uint32_t *w, uint32_t n, hipStream_t s,
                                  const uint32_t *bad) {
  const uint32_t epw = edges_per_wave(n, g->device);
  const uint64_t waves = (uint64_t(n) + epw - 1) / epw;
  const MutSide a = side_of(g, false, g->counters),
                b = g->rev ? side_of(g->rev, true, g->counters) : a;
  const uint64_t total = g->rev ? 2 * waves : waves;
  // per edge: batch entry (src, dst, weight), one slab probe, value write,
  // status byte; the mirror side the same
  prof_launch("k_update_weights", double(n) * (g->rev ? 2 : 1) * (12 + 256 + 4 + 1),
              k_update_weights, dim3(blocks_for_waves(total)), dim3(kBlock), 0u, s, a, b,
              uint32_t(waves), uint32_t(total - waves), src, dst, w, n, epw, bad, g->mut_epoch);
}

// Status buffer of the in-edge mirror for a batch of n edges.
static int mirror_prologue(mk_graph *g, uint32_t n) {
  if (!g->rev) return MK_OK;
  if (int r = g->rev->status.ensure(std::max<size_t>(n, 1))) return r;
  g->rev->status_n = n;
  return MK_OK;
}

// The outcome of a mutation whose counters are in host_scratch (its stream
// done): MK_ERR_RANGE (the kernels changed nothing), or MK_ERR_POOL_FULL
// after rolling the insert back on stream s, or MK_OK.
static int mutation_outcome(mk_graph *g, int op, const uint32_t *src, const uint32_t *dst,
                            uint32_t n, hipStream_t s) {
  const char *who = op == 1 ? "mk_graph_insert"
                   : op == 2 ? "mk_graph_delete"
                             : "mk_graph_update_weights";
  if (int r = range_error(g, src, dst, n, s, who)) return r;
  const uint32_t pool_fail = op == 1 ? pool_failures(g) : 0u;  // both sides count here
  if (pool_fail) {
    // roll back: delete what this batch inserted, restoring the edge set
    // (the range check passed, so the kernels run)
    launch_delete(g, src, dst, n, s, MK_EDGE_APPLIED, "k_insert_rollback", g->counters);
    MK_HIP(hipStreamSynchronize(s));
    return fail(MK_ERR_POOL_FULL, std::string(who) + ": slab pool exhausted (" +
                                      std::to_string(pool_fail) +
                                      " edges); batch rolled back, edge set unchanged");
  }
  return MK_OK;
}

}  // extern "C"

namespace mk {
// A failed batch: its statuses no longer describe the graph, and later
// algorithm calls given the same batch report its error (graph_failed_batch).
static void record_failure(mk_graph *g, int code, const uint32_t *src, const uint32_t *dst,
                           uint32_t n) {
  g->last_op = 0;
  g->failed.code = code;
  g->failed.src = src;
  g->failed.dst = dst;
  g->failed.n = n;
  g->failed.msg = g_err;
}

// Settle the graph's deferred mutation, if one is queued: wait for it and
// report its outcome once (restoring the graph on a pool failure).
int graph_settle(mk_graph *g) {
  if (g && g->mirror_of) g = g->mirror_of;  // a mirror changes with its graph
  if (!g || !g->pend.active) return MK_OK;
  g->pend.active = false;
  if (int r = host_wait_word(g->host_scratch.as<uint32_t>() + 4, g->pend.seq, g->pend.s))
    return r;
  const int r = mutation_outcome(g, g->pend.op, g->pend.src, g->pend.dst, g->pend.n, g->pend.s);
  if (r) record_failure(g, r, g->pend.src, g->pend.dst, g->pend.n);
  return r;
}

int graph_failed_batch(const mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n) {
  if (g && g->mirror_of) g = g->mirror_of;
  if (!g || !g->failed.code || n == 0) return MK_OK;
  if (g->failed.src != src || g->failed.dst != dst || g->failed.n != n) return MK_OK;
  return fail(g->failed.code, "this batch failed in its mutation call (" + g->failed.msg + ")");
}
}  // namespace mk

extern "C" {

// Shared back half of insert / delete: the outcome now, or (deferred mode)
// queued for the next call that settles the graph.
static int mutation_finish(mk_graph *g, int op, const uint32_t *src, const uint32_t *dst,
                           const uint32_t *w, uint32_t n, hipStream_t s) {
  MK_HIP(hipGetLastError());
  publish_words(g->counters, static_cast<uint32_t *>(g->host_scratch.dev), 4, g->mut_epoch, s);
  g->last_op = op;
  g->last_src = src;
  g->last_dst = dst;
  g->last_w = op == 1 || op == 3 ? w : nullptr;
  if (g->deferred) {
    // BFS/SSSP calls run on own_stream: only a mutation on another stream
    // needs an event for them to wait on
    if (s != g->own_stream) MK_HIP(hipEventRecord(g->mut_ev, s));
    g->pend = {true, op, src, dst, n, s, g->mut_epoch};
    return MK_OK;
  }
  MK_HIP(hipStreamSynchronize(s));
  const int r = mutation_outcome(g, op, src, dst, n, s);
  if (r) record_failure(g, r, src, dst, n);
  return r;
}

}  // extern "C"

namespace mk {
// A batch mutation queued for a captured update step (mk_step): the range
// check, the kernels (forward and mirror) and, with `host_dev` (pinned) set,
// the copy of the four counter words there, on the graph's own stream, with
// no host wait and nothing left for a later call to settle.  op 1 insert,
// 2 delete.
// *epoch is the epoch the kernels were given (a replay adds its base).
int graph_queue_mutation(mk_graph *g, int op, const uint32_t *src, const uint32_t *dst,
                         const uint32_t *w, uint32_t n, uint32_t *host_dev, uint32_t *epoch) {
  hipStream_t s = g->own_stream;
  const char *who = op == 1 ? "captured insert" : "captured delete";
  if (int r = batch_prologue(g, src, dst, n, s, who)) return r;
  if (int r = mirror_prologue(g, n)) return r;
  if (op == 1) {
    launch_insert(g, src, dst, w, n, s, "k_insert", g->counters);
    mark_new(g, src, dst, n, s);
  } else {
    launch_delete(g, src, dst, n, s, 0xFFu, "k_delete", g->counters);
  }
  MK_HIP(hipGetLastError());
  if (host_dev) publish_words(g->counters, host_dev, 4, g->mut_epoch, s);
  g->last_op = op;
  g->last_src = src;
  g->last_dst = dst;
  g->last_w = op == 1 ? w : nullptr;
  *epoch = g->mut_epoch;
  return MK_OK;
}

// The outcome of a replayed mutation from its four counter words (`words`,
// host) and the epoch its kernels ran with: MK_ERR_RANGE (nothing changed),
// MK_ERR_POOL_FULL after rolling an insert back, or MK_OK.
int graph_replay_outcome(mk_graph *g, int op, const uint32_t *src, const uint32_t *dst,
                         uint32_t n, const uint32_t *words, uint32_t epoch) {
  const char *who = op == 1 ? "captured insert" : "captured delete";
  if (words[0] == epoch) {
    const int r = fail(MK_ERR_RANGE, std::string(who) +
                                         ": the batch has a vertex id >= n; graph unchanged");
    record_failure(g, r, src, dst, n);
    return r;
  }
  const uint64_t pw = (uint64_t(words[3]) << 32) | words[2];
  const uint32_t pool_fail = op == 1 && uint32_t(pw >> 32) == epoch ? uint32_t(pw) : 0u;
  if (pool_fail) {
    // the insert's statuses are still the graph's last batch's
    launch_delete(g, src, dst, n, g->own_stream, MK_EDGE_APPLIED, "k_insert_rollback",
                  g->counters);
    MK_HIP(hipStreamSynchronize(g->own_stream));
    const int r = fail(MK_ERR_POOL_FULL, std::string(who) + ": slab pool exhausted (" +
                                             std::to_string(pool_fail) +
                                             " edges); batch rolled back, edge set unchanged");
    record_failure(g, r, src, dst, n);
    return r;
  }
  return MK_OK;
}
}  // namespace mk

extern "C" {

int mk_graph_insert(mk_graph *g, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                    uint32_t n, void *stream) {
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : (g ? g->own_stream : nullptr);
  if (g && g->mirror_of) return fail(MK_ERR_STATE, "mk_graph_insert: mirror is read-only");
  if (int r = graph_settle(g)) return r;
  if (int r = batch_prologue(g, src, dst, n, s, "mk_graph_insert")) return r;
  if (n && s != g->own_stream) {  // after the work queued on own_stream (deferred calls)
    MK_HIP(hipEventRecord(g->trav_ev, g->own_stream));
    MK_HIP(hipStreamWaitEvent(s, g->trav_ev, 0));
  }
  if (n == 0) return MK_OK;
  if (int r = mirror_prologue(g, n)) return r;
  launch_insert(g, src, dst, w, n, s, "k_insert", g->counters);
  mark_new(g, src, dst, n, s);  // graphs tracking updates: the applied keys are new
  return mutation_finish(g, 1, src, dst, w, n, s);
}

int mk_graph_delete(mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n,
                    void *stream) {
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : (g ? g->own_stream : nullptr);
  if (g && g->mirror_of) return fail(MK_ERR_STATE, "mk_graph_delete: mirror is read-only");
  if (int r = graph_settle(g)) return r;
  if (int r = batch_prologue(g, src, dst, n, s, "mk_graph_delete")) return r;
  if (n && s != g->own_stream) {  // after the work queued on own_stream (deferred calls)
    MK_HIP(hipEventRecord(g->trav_ev, g->own_stream));
    MK_HIP(hipStreamWaitEvent(s, g->trav_ev, 0));
  }
  if (n == 0) return MK_OK;
  if (int r = mirror_prologue(g, n)) return r;
  launch_delete(g, src, dst, n, s, 0xFFu, "k_delete", g->counters);
  return mutation_finish(g, 2, src, dst, nullptr, n, s);
}

int mk_graph_update_weights(mk_graph *g, const uint32_t *src, const uint32_t *dst,
                            const uint32_t *w, uint32_t n, void *stream) {
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : (g ? g->own_stream : nullptr);
  if (g && g->mirror_of)
    return fail(MK_ERR_STATE, "mk_graph_update_weights: mirror is read-only");
  if (g && !g->weighted)
    return fail(MK_ERR_BAD_ARG, "mk_graph_update_weights: the graph is unweighted (a set)");
  if (n && !w) return fail(MK_ERR_BAD_ARG, "mk_graph_update_weights: NULL weights");
  if (int r = graph_settle(g)) return r;
  if (int r = batch_prologue(g, src, dst, n, s, "mk_graph_update_weights")) return r;
  if (n && s != g->own_stream) {  // after the work queued on own_stream (deferred calls)
    MK_HIP(hipEventRecord(g->trav_ev, g->own_stream));
    MK_HIP(hipStreamWaitEvent(s, g->trav_ev, 0));
  }
  if (n == 0) return MK_OK;
  if (int r = mirror_prologue(g, n)) return r;
  launch_update_weights(g, src, dst, w, n, s, g->counters);
  return mutation_finish(g, 3, src, dst, w, n, s);
}

int mk_graph_set_deferred(mk_graph *g, int on) {
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_graph_set_deferred: NULL graph");
  if (g->mirror_of) return fail(MK_ERR_STATE, "mk_graph_set_deferred: mirror follows its graph");
  if (int r = graph_settle(g)) return r;
  g->deferred = on != 0;
  return MK_OK;
}

int mk_graph_sync(mk_graph *g) {
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_graph_sync: NULL graph");
  return graph_settle(g);
}

int mk_graph_search(mk_graph *g, const uint32_t *src, const uint32_t *dst, uint32_t n,
                    uint8_t *found, uint32_t *wout, void *stream) {
  if (int r = graph_settle(g)) return r;
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_graph_search: NULL graph");
  if (n && (!src || !dst || !found)) return fail(MK_ERR_BAD_ARG, "mk_graph_search: NULL buffer");
  if (n == 0) return MK_OK;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : (g ? g->own_stream : nullptr);
  MK_HIP(hipSetDevice(g->device));
  const uint32_t epw = edges_per_wave(n, g->device);
  const uint64_t waves = (uint64_t(n) + epw - 1) / epw;
  if (g->weighted)
    k_search<true><<<blocks_for_waves(waves), kBlock, 0, s>>>(g->view(), src, dst, n, epw, found,
                                                              wout);
  else
    k_search<false><<<blocks_for_waves(waves), kBlock, 0, s>>>(g->view(), src, dst, n, epw, found,
                                                               wout);
  MK_HIP(hipGetLastError());
  MK_HIP(hipStreamSynchronize(s));
  return MK_OK;
}

int mk_graph_last_status(mk_graph *g, const uint8_t **status, uint32_t *n) {
  if (int r = graph_settle(g)) return r;
  if (!g || !status || !n) return fail(MK_ERR_BAD_ARG, "mk_graph_last_status: NULL argument");
  *status = g->status.as<uint8_t>();
  *n = g->status_n;
  return MK_OK;
}

int mk_graph_num_vertices(mk_graph *g, uint32_t *n) {
  if (!g || !n) return fail(MK_ERR_BAD_ARG, "mk_graph_num_vertices: NULL argument");
  *n = g->n;
  return MK_OK;
}

int mk_graph_is_weighted(mk_graph *g, int *w) {
  if (!g || !w) return fail(MK_ERR_BAD_ARG, "mk_graph_is_weighted: NULL argument");
  *w = g->weighted;
  return MK_OK;
}

int mk_graph_degrees(mk_graph *g, uint32_t *deg) {
  if (int r = graph_settle(g)) return r;
  if (!g || !deg) return fail(MK_ERR_BAD_ARG, "mk_graph_degrees: NULL argument");
  if (g->mirror_of) return mk_graph_count_slots(g, deg, nullptr);  // counted from the slabs
  MK_HIP(hipSetDevice(g->device));
  MK_HIP(hipMemcpy(deg, g->vdeg, size_t(g->n) * 4, hipMemcpyDeviceToHost));
  return MK_OK;
}

int mk_graph_num_edges(mk_graph *g, uint64_t *m) {
  if (int r = graph_settle(g)) return r;
  if (!g || !m) return fail(MK_ERR_BAD_ARG, "mk_graph_num_edges: NULL argument");
  std::vector<uint32_t> deg(g->n);
  if (int r = mk_graph_degrees(g, deg.data())) return r;
  uint64_t t = 0;
  for (uint32_t d : deg) t += d;
  *m = t;
  return MK_OK;
}

int mk_graph_bucket_offsets(mk_graph *g, uint32_t *off) {
  if (int r = graph_settle(g)) return r;
  if (!g || !off) return fail(MK_ERR_BAD_ARG, "mk_graph_bucket_offsets: NULL argument");
  std::memcpy(off, g->bucket_off_host.data(), (size_t(g->n) + 1) * 4);
  return MK_OK;
}

static int count_slots(mk_graph *g, uint32_t *live_dev, uint32_t *tomb_dev,
                       unsigned long long *overflow_dev, uint32_t *max_dev) {
  // 32 waves per CU of eight chain walks each
  const unsigned grid = blocks_for_waves(std::max<uint64_t>(
      1, std::min<uint64_t>(uint64_t(device_cus(g->device)) * 32,
                            (uint64_t(g->total_buckets) + 7) / 8)));
  if (g->weighted)
    k_count_slots<true><<<grid, kBlock>>>(g->view(), live_dev, tomb_dev, overflow_dev, max_dev);
  else
    k_count_slots<false><<<grid, kBlock>>>(g->view(), live_dev, tomb_dev, overflow_dev, max_dev);
  MK_HIP(hipGetLastError());
  return MK_OK;
}

int mk_graph_count_slots(mk_graph *g, uint32_t *live, uint32_t *tomb) {
  if (int r = graph_settle(g)) return r;
  if (!g || (!live && !tomb)) return fail(MK_ERR_BAD_ARG, "mk_graph_count_slots: NULL argument");
  MK_HIP(hipSetDevice(g->device));
  DevBuf b;
  if (int r = b.ensure(size_t(g->n) * 8)) return r;
  uint32_t *l = b.as<uint32_t>(), *t = l + g->n;
  fill_u32(l, 0u, uint64_t(g->n) * 2, nullptr);
  int r = count_slots(g, l, t, nullptr, nullptr);
  if (r == MK_OK) {
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && live) e = hipMemcpy(live, l, size_t(g->n) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && tomb) e = hipMemcpy(tomb, t, size_t(g->n) * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = hip_fail(e, "mk_graph_count_slots");
  }
  b.release();
  return r;
}

}  // extern "C"

namespace mk {
// Scratch of mk_graph_count_slots_noalloc: the live and tombstone counts of
// up to kNoAllocMaxVertices vertices (2 x 16 MiB), one copy per device,
// held until the process ends.  Host code never passes `scratch` itself to
// a kernel or a copy (its host-side name is not a device address): it uses
// the address hipGetSymbolAddress returned for the device, kept below.
constexpr uint32_t kNoAllocMaxVertices = 1u << 22;
static __device__ __align__(256) uint32_t scratch[2 * size_t(kNoAllocMaxVertices)];

struct NoAllocScratch {
  std::mutex mu;         // one call at a time per device (they share the array)
  uint32_t *ptr = nullptr;  // device address of `scratch` on this device
};
static std::mutex noalloc_table_mu;
static std::map<int, NoAllocScratch> noalloc_table;

// The device's scratch, looked up on the first call on that device (with the
// device current); nullptr and an error set on failure.
static NoAllocScratch *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(noalloc_table_mu);
  NoAllocScratch &e = noalloc_table[device];  // node-based map: the address stays valid
  if (!e.ptr) {
    void *p = nullptr;
    const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p) {
      hip_fail(r == hipSuccess ? hipErrorInvalidSymbol : r,
               "mk_graph_count_slots_noalloc: hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    e.ptr = static_cast<uint32_t *>(p);
  }
  return &e;
}
}  // namespace mk

extern "C" {

int mk_graph_count_slots_noalloc(mk_graph *g, uint32_t *live, uint32_t *tomb) {
  if (int r = graph_settle(g)) return r;
  if (!g || (!live && !tomb))
    return fail(MK_ERR_BAD_ARG, "mk_graph_count_slots_noalloc: NULL argument");
  if (g->n > kNoAllocMaxVertices)
    return fail(MK_ERR_BAD_ARG, "mk_graph_count_slots_noalloc: the graph has " +
                                    std::to_string(g->n) + " vertices, more than the " +
                                    std::to_string(kNoAllocMaxVertices) +
                                    " its fixed scratch holds (use mk_graph_count_slots)");
  MK_HIP(hipSetDevice(g->device));
  NoAllocScratch *sc = noalloc_scratch(g->device);
  if (!sc) return MK_ERR_HIP;
  std::lock_guard<std::mutex> lk(sc->mu);
  // the previous call's counts are still there: clear this call's part
  uint32_t *l = sc->ptr, *t = sc->ptr + g->n;
  fill_u32(l, 0u, uint64_t(g->n) * 2, nullptr);
  int r = count_slots(g, l, t, nullptr, nullptr);
  if (r == MK_OK) {
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && live) e = hipMemcpy(live, l, size_t(g->n) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && tomb) e = hipMemcpy(tomb, t, size_t(g->n) * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = hip_fail(e, "mk_graph_count_slots_noalloc");
  }
  return r;
}

int mk_graph_slab_stats(mk_graph *g, uint64_t *heads, uint64_t *overflow, uint32_t *max_chain) {
  if (int r = graph_settle(g)) return r;
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_graph_slab_stats: NULL graph");
  MK_HIP(hipSetDevice(g->device));
  DevBuf b;
  if (int r = b.ensure(16)) return r;
  unsigned long long *ov = b.as<unsigned long long>();
  uint32_t *mx = reinterpret_cast<uint32_t *>(ov + 1);
  MK_HIP(hipMemset(b.ptr, 0, 16));
  int r = count_slots(g, nullptr, nullptr, ov, mx);
  unsigned long long host[2] = {0, 0};
  if (r == MK_OK) {
    hipError_t e = hipMemcpy(host, b.ptr, 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = hip_fail(e, "mk_graph_slab_stats");
  }
  b.release();
  if (heads) *heads = g->total_buckets;
  if (overflow) *overflow = host[0];
  if (max_chain) *max_chain = uint32_t(host[1]);
  return r;
}

int mk_graph_export_csr(mk_graph *g, uint64_t *row_off, uint32_t *col, uint32_t *wcol) {
  if (int r = graph_settle(g)) return r;
  if (!g || !row_off) return fail(MK_ERR_BAD_ARG, "mk_graph_export_csr: NULL argument");
  MK_HIP(hipSetDevice(g->device));
  std::vector<uint32_t> live(g->n);
  if (int r = mk_graph_count_slots(g, live.data(), nullptr)) return r;
  row_off[0] = 0;
  for (uint32_t v = 0; v < g->n; ++v) row_off[v + 1] = row_off[v] + live[v];
  const uint64_t m = row_off[g->n];
  if (!col) return MK_OK;
  DevBuf cur, c, w;
  int r = cur.ensure(size_t(g->n) * 8);
  if (!r) r = c.ensure(std::max<uint64_t>(m, 1) * 4);
  if (!r && g->weighted && wcol) r = w.ensure(std::max<uint64_t>(m, 1) * 4);
  if (!r) {
    hipError_t e =
        hipMemcpy(cur.ptr, row_off, size_t(g->n) * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      if (g->weighted)
        k_export<true><<<blocks_for_waves(g->total_buckets), kBlock>>>(
            g->view(), cur.as<unsigned long long>(), c.as<uint32_t>(),
            wcol ? w.as<uint32_t>() : nullptr);
      else
        k_export<false><<<blocks_for_waves(g->total_buckets), kBlock>>>(
            g->view(), cur.as<unsigned long long>(), c.as<uint32_t>(), nullptr);
      e = hipDeviceSynchronize();
    }
    if (e == hipSuccess && m) e = hipMemcpy(col, c.ptr, m * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && m && g->weighted && wcol)
      e = hipMemcpy(wcol, w.ptr, m * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = hip_fail(e, "mk_graph_export_csr");
  }
  cur.release();
  c.release();
  w.release();
  return r;
}

}  // extern "C"
