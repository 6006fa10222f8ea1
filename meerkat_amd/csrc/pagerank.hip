// PageRank (FP32 ranks, FP64 reductions) over a slab-hash in-adjacency.
//
// Reference: include/Algorithms/PageRank.cuh:9-68 and impl/PageRankImpl.cuh:
//   contrib[v] = outdeg[v] ? PR[v] / outdeg[v] : 0           (:133-144)
//   acc        = sum_{outdeg(x) = 0} PR[x] / N                 (:89-107)
//   PR'[u]     = (1 - d) / N + d * sum_{v in in(u)} contrib[v] + d * acc
//                                                    (:3-62, :109-115, :193)
//   delta      = sum |PR' - PR|; loop while delta > margin && it < max_iter
//                                                    (:117-131, :173)
// Ranks start at 1/N (PageRank.cuh:42-43); Static, Incremental and
// Decremental all continue from the current ranks (warm start), after the
// out-degree fix-up for the batch (CheckPresent, :64-87).
//
// Method: one fused kernel per iteration computes PR', the next
// contributions, the next dangling mass and the L1 delta (FP64 partial sums
// spread over kSlots addresses).  Each iteration kernel first evaluates the
// stopping rule from the previous kernel's delta, so the host launches
// iterations in groups without waiting and the iteration count is exactly
// the reference's.  Vertices whose in-adjacency spans more than kHubBuckets
// buckets have their buckets spread over all waves (partial sums added to a
// per-hub FP64 accumulator) and are finished by a second, small kernel.
//
// In-adjacency: either a graph holding reversed edges, as the reference's
// RevUnweightedEdgePolicyT graph (PageRank.cuh:11; the driver inserts (dst,
// src), src/PageRank.cu:137), whose out-degrees are host-provided and fixed
// up per batch; or a graph created with MK_GRAPH_REVERSE, whose mirror is
// the in-adjacency and whose live out-degree array is exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "profile.hpp"
#include "slabhash.hpp"

namespace mk {

constexpr uint32_t kSlots = 64;       // FP64 partial-sum addresses
constexpr uint32_t kHubBuckets = 4;   // more in-buckets than this: hub path
constexpr uint32_t kPrBatch = 4;      // slab loads in flight per wave

struct PrState {
  double delta[3][kSlots];   // L1 delta of iteration i in [i % 3]
  double dang[3][kSlots];    // dangling mass of the ranks iteration i reads
  uint32_t iters;            // iterations completed in this call
  uint32_t stopped;          // stopping rule met
  double last_delta;
};

struct PrView {
  float *pr[2];
  float *contrib[2];
  const uint32_t *outdeg;
  PrState *st;
  const uint32_t *hub_vertex;  // hub index -> vertex
  const uint32_t *hub_start;   // hub index -> first work item (sorted)
  double *hub_acc;             // per hub in-sum of the running iteration
  uint32_t n_hubs, hub_items;
  uint32_t n;
  float damping;
  double margin;
  uint32_t max_iter;
};

MK_DEV double wave_sum_slots(const double *slots) {
  double x = slots[lane_id()];
  return wave_sum_f64(x);
}

MK_DEV void add_slot(double *slots, double v) {
  if (v != 0.0) atomicAdd(&slots[blockIdx.x % kSlots], v);
}

// Stopping rule at the start of iteration i; every thread evaluates it.
MK_DEV bool pr_should_stop(const PrView &P, uint32_t i) {
  if (P.st->stopped) return true;
  if (i >= P.max_iter) return true;
  if (i == 0) return false;  // the reference's loop starts with delta = 1
  const double d = wave_sum_slots(P.st->delta[(i - 1) % 3]);
  return !(d > P.margin);
}

// Sum of contrib over one in-chain (first slab word x loaded).
template <bool MAP>
MK_DEV float chain_sum(const GraphView &R, const float *contrib, uint32_t x, uint32_t &slabs) {
  const uint32_t lane = lane_id();
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  float acc = 0.f;
  for (;;) {
    ++slabs;
    if (key_lane && x < kTomb) acc += contrib[x];
    const uint32_t nx = read_lane(x, kNextLane);
    if (nx == kEmpty) break;
    x = slab_ptr(R.pool, nx)[lane];
  }
  return wave_sum_f32(acc);
}

// Clears every reduction slot and the iteration counters of a compute call
// (one block of kSlots threads, launched before k_pr_prepare adds into
// dang[0]).
__global__ void k_pr_reset(PrState *st) {
  const uint32_t t = threadIdx.x;
  for (uint32_t k = 0; k < 3; ++k) {
    st->delta[k][t] = 0.0;
    st->dang[k][t] = 0.0;
  }
  if (t == 0) {
    st->iters = 0;
    st->stopped = 0;
  }
}

// Initial contributions and dangling mass of the current ranks.
__global__ void __launch_bounds__(kBlock) k_pr_prepare(PrView P, uint32_t cur) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double dang = 0.0;
  if (i < P.n) {
    const float r = P.pr[cur][i];
    const uint32_t od = P.outdeg[i];
    P.contrib[cur][i] = od ? r / float(od) : 0.f;
    if (od == 0) dang = double(r);
  }
  dang = wave_sum_f64(dang);
  if (lane_id() == 0) add_slot(P.st->dang[0], dang);
}

// One PageRank iteration i over non-hub vertices, plus the hubs' partial
// in-sums.  Reads pr/contrib[(cur + i) & 1], writes the other buffers.
template <bool MAP>
__global__ void __launch_bounds__(kBlock)
    k_pr_iterate(GraphView R, PrView P, uint32_t cur, uint32_t i) {
  if (pr_should_stop(P, i)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) P.st->stopped = 1;
    return;
  }
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const uint32_t a = (cur + i) & 1u, b = a ^ 1u;
  if (blockIdx.x == 0 && threadIdx.x < kSlots) {
    P.st->delta[(i + 1) % 3][threadIdx.x] = 0.0;  // written by iteration i + 1
    P.st->dang[(i + 2) % 3][threadIdx.x] = 0.0;   // written by iteration i + 1
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) P.st->iters = i + 1;
  const float *contrib = P.contrib[a];
  const float base = (1.0f - P.damping) / float(P.n);
  const float dang = float(wave_sum_slots(P.st->dang[i % 3]) / double(P.n));
  uint32_t slabs = 0;
  // hub buckets: work items spread over all waves, partial sums per hub
  for (uint32_t t0 = gw * 64; t0 < P.hub_items; t0 += W * 64) {
    const uint32_t t = t0 + lane;
    bool act = t < P.hub_items;
    uint32_t h = 0, bk = 0;
    if (act) {
      uint32_t lo = 0, hi = P.n_hubs;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (P.hub_start[mid] <= t) lo = mid;
        else hi = mid;
      }
      h = lo;
      bk = R.bucket_off[P.hub_vertex[h]] + (t - P.hub_start[h]);
    }
    uint64_t m = ballot(act);
    while (m) {
      uint32_t js[kPrBatch];
      uint32_t c = 0;
#pragma unroll
      for (uint32_t k = 0; k < kPrBatch; ++k) {
        js[k] = m ? ffs64(m) : 0u;
        if (m) {
          m &= m - 1;
          c = k + 1;
        }
      }
      uint32_t x[kPrBatch];
#pragma unroll
      for (uint32_t k = 0; k < kPrBatch; ++k)
        x[k] = k < c ? R.heads[size_t(read_lane(bk, js[k])) * kSlabWords + lane] : kEmpty;
#pragma unroll
      for (uint32_t k = 0; k < kPrBatch; ++k) {
        if (k >= c) continue;
        const float s = chain_sum<MAP>(R, contrib, x[k], slabs);
        if (lane == 0) atomicAdd(&P.hub_acc[read_lane(h, js[k])], double(s));
      }
    }
  }
  // ordinary vertices: 64 per wave step, one lane per vertex
  double delta = 0.0, dnext = 0.0;
  for (uint32_t u0 = gw * 64; u0 < P.n; u0 += W * 64) {
    const uint32_t u = u0 + lane;
    bool act = u < P.n;
    uint32_t b0 = 0, nb = 0;
    if (act) {
      b0 = R.bucket_off[u];
      nb = R.bucket_off[u + 1] - b0;
      if (nb > kHubBuckets) act = false;  // finished by k_pr_hubs
    }
    float acc = 0.f;
    for (uint32_t k = 0; ballot(act && nb > k); ++k) {
      uint64_t m = ballot(act && nb > k);
      while (m) {
        uint32_t js[kPrBatch];
        uint32_t c = 0;
#pragma unroll
        for (uint32_t q = 0; q < kPrBatch; ++q) {
          js[q] = m ? ffs64(m) : 0u;
          if (m) {
            m &= m - 1;
            c = q + 1;
          }
        }
        uint32_t x[kPrBatch];
#pragma unroll
        for (uint32_t q = 0; q < kPrBatch; ++q)
          x[q] = q < c ? R.heads[size_t(read_lane(b0, js[q]) + k) * kSlabWords + lane] : kEmpty;
#pragma unroll
        for (uint32_t q = 0; q < kPrBatch; ++q) {
          if (q >= c) continue;
          const float s = chain_sum<MAP>(R, contrib, x[q], slabs);
          if (lane == js[q]) acc += s;
        }
      }
    }
    if (act) {
      const float x = base + P.damping * acc + P.damping * dang;
      const float old = P.pr[a][u];
      delta += double(fabsf(x - old));
      P.pr[b][u] = x;
      const uint32_t od = P.outdeg[u];
      P.contrib[b][u] = od ? x / float(od) : 0.f;
      if (od == 0) dnext += double(x);
    }
  }
  delta = wave_sum_f64(delta);
  dnext = wave_sum_f64(dnext);
  if (lane == 0) {
    add_slot(P.st->delta[i % 3], delta);
    add_slot(P.st->dang[(i + 1) % 3], dnext);
  }
}

// Finish the hub vertices of iteration i (after k_pr_iterate).
__global__ void __launch_bounds__(kBlock) k_pr_hubs(PrView P, uint32_t cur, uint32_t i) {
  // k_pr_iterate of this iteration already evaluated the rule and set
  // `stopped`; iters tells whether iteration i ran
  if (P.st->iters != i + 1 || P.st->stopped) return;
  const uint32_t a = (cur + i) & 1u, b = a ^ 1u;
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  const float base = (1.0f - P.damping) / float(P.n);
  const float dang = float(wave_sum_slots(P.st->dang[i % 3]) / double(P.n));
  double delta = 0.0, dnext = 0.0;
  if (h < P.n_hubs) {
    const uint32_t u = P.hub_vertex[h];
    const float x = base + P.damping * float(P.hub_acc[h]) + P.damping * dang;
    P.hub_acc[h] = 0.0;
    delta = double(fabsf(x - P.pr[a][u]));
    P.pr[b][u] = x;
    const uint32_t od = P.outdeg[u];
    P.contrib[b][u] = od ? x / float(od) : 0.f;
    if (od == 0) dnext = double(x);
  }
  delta = wave_sum_f64(delta);
  dnext = wave_sum_f64(dnext);
  if (lane_id() == 0) {
    add_slot(P.st->delta[i % 3], delta);
    add_slot(P.st->dang[(i + 1) % 3], dnext);
  }
}

// Out-degree fix-up for a batch: +-1 for each edge the last mutation applied.
__global__ void k_pr_fix_outdeg(uint32_t *outdeg, const uint32_t *src, const uint8_t *status,
                                uint32_t n, int delta) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && status[i] == MK_EDGE_APPLIED) atomicAdd(outdeg + src[i], uint32_t(delta));
}

// Out-degree recount from the in-adjacency: every live key v of an
// in-bucket is one out-edge of v.
template <bool MAP>
__global__ void __launch_bounds__(kBlock) k_pr_recount(GraphView R, uint32_t *outdeg) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  for (uint32_t bk = gw; bk < R.total_buckets; bk += W) {
    uint32_t x = R.heads[size_t(bk) * kSlabWords + lane];
    for (;;) {
      if (key_lane && x < kTomb) atomicAdd(outdeg + x, 1u);
      const uint32_t nx = read_lane(x, kNextLane);
      if (nx == kEmpty) break;
      x = slab_ptr(R.pool, nx)[lane];
    }
  }
}

__global__ void k_fill_f32(float *p, float v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

struct PageRank {
  mk_graph *g = nullptr;   // graph passed at create
  mk_graph *in = nullptr;  // in-adjacency (g, or g's mirror)
  bool exact_outdeg = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float *pr[2] = {nullptr, nullptr};
  float *contrib[2] = {nullptr, nullptr};
  uint32_t *outdeg = nullptr;
  PrState *st = nullptr;
  uint32_t *hub_vertex = nullptr, *hub_start = nullptr;
  double *hub_acc = nullptr;
  uint32_t n_hubs = 0, hub_items = 0;
  uint32_t cur = 0;
  float damping = 0.85f;
  float margin = 1e-7f;
  uint32_t max_iter = 100;
  HostBuf host;
  uint32_t last_iters = 0;
  double last_delta = 0.0;
  int max_waves = 0;

  PrView view() const {
    PrView v;
    v.pr[0] = pr[0];
    v.pr[1] = pr[1];
    v.contrib[0] = contrib[0];
    v.contrib[1] = contrib[1];
    v.outdeg = outdeg;
    v.st = st;
    v.hub_vertex = hub_vertex;
    v.hub_start = hub_start;
    v.hub_acc = hub_acc;
    v.n_hubs = n_hubs;
    v.hub_items = hub_items;
    v.n = g->n;
    v.damping = damping;
    v.margin = double(margin);
    v.max_iter = max_iter;
    return v;
  }

  int init(mk_graph *graph, float d, float eps, uint32_t iters, const uint32_t *host_outdeg) {
    g = graph;
    in = g->rev ? g->rev : g;
    exact_outdeg = g->rev != nullptr;
    damping = d;
    margin = eps;
    max_iter = iters;
    const size_t n = g->n;
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipStreamCreate(&stream));
    loop_stream_add(g->device, stream, 1);
    MK_HIP(hipEventCreate(&ev0));
    MK_HIP(hipEventCreate(&ev1));
    for (int k = 0; k < 2; ++k) {
      MK_HIP(hipMalloc(&pr[k], n * 4));
      MK_HIP(hipMalloc(&contrib[k], n * 4));
    }
    MK_HIP(hipMalloc(&outdeg, n * 4));
    MK_HIP(hipMalloc(&st, sizeof(PrState)));
    if (int r = host.ensure(sizeof(PrState))) return r;
    // hubs: vertices with more than kHubBuckets in-buckets (bucket counts are
    // fixed at graph creation)
    std::vector<uint32_t> hv, hs;
    uint64_t items = 0;
    const std::vector<uint32_t> &off = in->bucket_off_host;
    for (uint32_t v = 0; v < g->n; ++v) {
      const uint32_t nb = off[v + 1] - off[v];
      if (nb > kHubBuckets) {
        hv.push_back(v);
        hs.push_back(uint32_t(items));
        items += nb;
      }
    }
    n_hubs = uint32_t(hv.size());
    hub_items = uint32_t(items);
    MK_HIP(hipMalloc(&hub_vertex, std::max<size_t>(hv.size(), 1) * 4));
    MK_HIP(hipMalloc(&hub_start, std::max<size_t>(hs.size(), 1) * 4));
    MK_HIP(hipMalloc(&hub_acc, std::max<size_t>(hv.size(), 1) * 8));
    if (n_hubs) {
      MK_HIP(hipMemcpy(hub_vertex, hv.data(), hv.size() * 4, hipMemcpyHostToDevice));
      MK_HIP(hipMemcpy(hub_start, hs.data(), hs.size() * 4, hipMemcpyHostToDevice));
    }
    fill_u64(reinterpret_cast<uint64_t *>(hub_acc), 0ull, std::max<uint32_t>(n_hubs, 1), stream);
    k_fill_f32<<<blocks_for_threads(n), kBlock, 0, stream>>>(pr[0], 1.0f / float(n), g->n);
    if (exact_outdeg) {
      MK_HIP(hipMemcpyAsync(outdeg, g->vdeg, n * 4, hipMemcpyDeviceToDevice, stream));
    } else if (host_outdeg) {
      MK_HIP(hipMemcpyAsync(outdeg, host_outdeg, n * 4, hipMemcpyHostToDevice, stream));
    } else {
      if (int r = recount()) return r;
    }
    MK_HIP(hipStreamSynchronize(stream));
    // the iteration gathers 4-byte contributions at random: as many waves
    // as a CU holds (`pr_waves_per_cu`, 32) keep the most loads in flight
    const void *kern = in->weighted ? reinterpret_cast<const void *>(&k_pr_iterate<true>)
                                    : reinterpret_cast<const void *>(&k_pr_iterate<false>);
    MK_HIP(resident_waves(g->device, kern, int(tune_get("pr_waves_per_cu", 32)), &max_waves));
    cur = 0;
    return MK_OK;
  }

  int recount() {
    fill_u32(outdeg, 0u, g->n, stream);
    const unsigned grid = blocks_for_waves(std::min<uint64_t>(in->total_buckets,
                                                              uint64_t(device_cus(g->device)) * 16));
    if (in->weighted)
      k_pr_recount<true><<<grid, kBlock, 0, stream>>>(in->view(), outdeg);
    else
      k_pr_recount<false><<<grid, kBlock, 0, stream>>>(in->view(), outdeg);
    MK_HIP(hipGetLastError());
    return MK_OK;
  }

  void destroy() {
    (void)hipSetDevice(g->device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (int k = 0; k < 2; ++k) {
      (void)hipFree(pr[k]);
      (void)hipFree(contrib[k]);
    }
    (void)hipFree(outdeg);
    (void)hipFree(st);
    (void)hipFree(hub_vertex);
    (void)hipFree(hub_start);
    (void)hipFree(hub_acc);
    host.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) {
      loop_stream_add(g->device, stream, -1);
      (void)hipStreamDestroy(stream);
    }
  }

  int compute(float *ms) {
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipEventRecord(ev0, stream));
    k_pr_reset<<<1, kSlots, 0, stream>>>(st);
    const PrView P = view();
    const GraphView R = in->view();
    k_pr_prepare<<<blocks_for_threads(g->n), kBlock, 0, stream>>>(P, cur);
    const unsigned grid = blocks_for_waves(
        std::min<uint64_t>((uint64_t(g->n) + 63) / 64 + hub_items / 64 + 1, uint64_t(max_waves)));
    const unsigned hgrid = blocks_for_threads(std::max<uint32_t>(n_hubs, 1));
    uint32_t i = 0, group = 8;
    for (;;) {
      for (uint32_t k = 0; k < group; ++k, ++i) {
        // bytes: 256 per in-slab + 4 per key (contribution gather) and 20 per
        // vertex (bucket range, old rank, new rank, outdeg, contribution),
        // accounted with the in-adjacency size below
        ProfScope prof(stream, "k_pr_iterate", iter_bytes());
        if (in->weighted)
          k_pr_iterate<true><<<grid, kBlock, 0, stream>>>(R, P, cur, i);
        else
          k_pr_iterate<false><<<grid, kBlock, 0, stream>>>(R, P, cur, i);
        if (n_hubs) k_pr_hubs<<<hgrid, kBlock, 0, stream>>>(P, cur, i);
      }
      MK_HIP(hipGetLastError());
      MK_HIP(hipMemcpyAsync(host.ptr, st, sizeof(PrState), hipMemcpyDeviceToHost, stream));
      MK_HIP(hipStreamSynchronize(stream));
      const PrState &h = *host.as<PrState>();
      if (h.stopped || h.iters < i || i >= max_iter) {
        last_iters = h.iters;
        double d = 0.0;
        if (h.iters > 0)
          for (uint32_t k = 0; k < kSlots; ++k) d += h.delta[(h.iters - 1) % 3][k];
        last_delta = d;
        break;
      }
      group = 8;
    }
    cur = (cur + last_iters) & 1u;
    MK_HIP(hipEventRecord(ev1, stream));
    MK_HIP(hipEventSynchronize(ev1));
    float t = 0.f;
    MK_HIP(hipEventElapsedTime(&t, ev0, ev1));
    if (ms) *ms = t;
    return MK_OK;
  }

  double iter_bytes() const {
    // in-slabs (heads + overflow are read once), one 4 B gather per edge and
    // 20 B per vertex; the edge count is the live out-degree total
    return 256.0 * double(in->total_buckets) + 20.0 * double(g->n) + 4.0 * edge_estimate;
  }
  double edge_estimate = 0.0;

  int batch(const uint32_t *src, const uint32_t *dst, uint32_t n, int sign, float *ms) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "pagerank: NULL batch");
    MK_HIP(hipSetDevice(g->device));
    if (n) {
      if (exact_outdeg) {
        MK_HIP(hipMemcpyAsync(outdeg, g->vdeg, size_t(g->n) * 4, hipMemcpyDeviceToDevice,
                              stream));
      } else {
        // the reverse graph saw (dst, src); use its per-edge status when this
        // is the batch it last applied, else recount from the in-adjacency
        const int op = sign > 0 ? 1 : 2;
        const bool same = g->last_op == op && g->status_n == n &&
                          ((g->last_src == dst && g->last_dst == src) ||
                           (g->last_src == src && g->last_dst == dst));
        if (same) {
          k_pr_fix_outdeg<<<blocks_for_threads(n), kBlock, 0, stream>>>(
              outdeg, src, g->status.as<uint8_t>(), n, sign);
          MK_HIP(hipGetLastError());
        } else if (int r = recount()) {
          return r;
        }
      }
    }
    return compute(ms);
  }

  int read(float *out) {
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipStreamSynchronize(stream));
    MK_HIP(hipMemcpy(out, pr[cur], size_t(g->n) * 4, hipMemcpyDeviceToHost));
    return MK_OK;
  }
};

}  // namespace mk

struct mk_pagerank {
  mk::PageRank p;
};

using namespace mk;

extern "C" {

int mk_pagerank_create(mk_graph *g, float damping, float error_margin, uint32_t max_iter,
                       const uint32_t *host_outdeg, mk_pagerank **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "mk_pagerank_create: out is NULL");
  *out = nullptr;
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_pagerank_create: NULL graph");
  if (!(damping >= 0.f && damping < 1.f)) return fail(MK_ERR_BAD_ARG, "damping must be in [0, 1)");
  if (!(error_margin >= 0.f)) return fail(MK_ERR_BAD_ARG, "error margin must be >= 0");
  mk_pagerank *h = new mk_pagerank();
  int r = h->p.init(g, damping, error_margin, max_iter, host_outdeg);
  if (r) {
    h->p.destroy();
    delete h;
    return r;
  }
  uint64_t m = 0;
  if (mk_graph_num_edges(g, &m) == MK_OK) h->p.edge_estimate = double(m);
  g->users++;
  *out = h;
  return MK_OK;
}

int mk_pagerank_destroy(mk_pagerank *p) {
  if (!p) return fail(MK_ERR_BAD_ARG, "mk_pagerank_destroy: NULL handle");
  p->p.g->users--;
  p->p.destroy();
  delete p;
  return MK_OK;
}

int mk_pagerank_static(mk_pagerank *p, float *ms) {
  if (!p) return fail(MK_ERR_BAD_ARG, "mk_pagerank_static: NULL handle");
  if (int r = graph_settle(p->p.g)) return r;
  return p->p.compute(ms);
}

int mk_pagerank_incremental(mk_pagerank *p, const uint32_t *src, const uint32_t *dst, uint32_t n,
                            float *ms) {
  if (!p) return fail(MK_ERR_BAD_ARG, "mk_pagerank_incremental: NULL handle");
  if (int r = graph_settle(p->p.g)) return r;
  // a graph of reversed edges saw the batch as (dst, src)
  if (int r = graph_failed_batch(p->p.g, src, dst, n)) return r;
  if (int r = graph_failed_batch(p->p.g, dst, src, n)) return r;
  return p->p.batch(src, dst, n, +1, ms);
}

int mk_pagerank_decremental(mk_pagerank *p, const uint32_t *src, const uint32_t *dst, uint32_t n,
                            float *ms) {
  if (!p) return fail(MK_ERR_BAD_ARG, "mk_pagerank_decremental: NULL handle");
  if (int r = graph_settle(p->p.g)) return r;
  if (int r = graph_failed_batch(p->p.g, src, dst, n)) return r;
  if (int r = graph_failed_batch(p->p.g, dst, src, n)) return r;
  return p->p.batch(src, dst, n, -1, ms);
}

int mk_pagerank_read(mk_pagerank *p, float *values) {
  if (!p || !values) return fail(MK_ERR_BAD_ARG, "mk_pagerank_read: NULL argument");
  return p->p.read(values);
}

int mk_pagerank_last_stats(mk_pagerank *p, uint32_t *iterations, double *last_delta) {
  if (!p) return fail(MK_ERR_BAD_ARG, "mk_pagerank_last_stats: NULL handle");
  if (iterations) *iterations = p->p.last_iters;
  if (last_delta) *last_delta = p->p.last_delta;
  return MK_OK;
}

}  // extern "C"
