// Connected components by concurrent union-find over the slab-hash graph.
//
// Reference: include/Util/UnionFind.cuh (primitives only; no caller and no
// CC driver in the reference).  UnionAsync (:122-149) finds both roots and
// links the larger root under the smaller with atomicCAS; this keeps every
// parent pointer <= its vertex, so a component's root is its smallest vertex
// and the labels are canonical (min vertex id) with no relabelling pass.
// Find uses path halving with plain stores (find_op::FindCompress / halving
// variants, :24-86): a halving store only replaces a parent by one of its
// ancestors, which keeps parent[x] <= x, so concurrent halving is safe.
//
// Edges are taken as undirected (weak components).  Static = union over
// every edge in the slabs, then a compress pass (label = root).  Incremental
// = union over the inserted batch, then compress.  Decremental recomputes
// (union-find cannot split components).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "profile.hpp"
#include "slabhash.hpp"

namespace mk {

MK_DEV uint32_t uf_find(uint32_t *parent, uint32_t x) {
  uint32_t p = parent[x];
  while (p != x) {
    const uint32_t gp = parent[p];
    if (gp != p) parent[x] = gp;  // halving: gp is an ancestor, gp <= p
    x = p;
    p = parent[x];
  }
  return x;
}

// Root of x reading parents with agent-coherent loads (no halving).
MK_DEV uint32_t uf_find_coherent(uint32_t *parent, uint32_t x) {
  for (;;) {
    const uint32_t p = coherent_load(parent + x);
    if (p == x) return x;
    x = p;
  }
}

// Link the roots of a and b, the larger under the smaller.  A plain load may
// return an older parent; older parents are still ancestors (pointers only
// move up), so a stale walk ends at a vertex on the true path and a lost CAS
// is the only consequence.  After a lost CAS the walk resumes from the
// parent the CAS returned, with coherent loads, so every retry climbs
// strictly and the loop ends.
MK_DEV void uf_union(uint32_t *parent, uint32_t a, uint32_t b) {
  bool fresh = false;
  for (;;) {
    uint32_t ra = fresh ? uf_find_coherent(parent, a) : uf_find(parent, a);
    uint32_t rb = fresh ? uf_find_coherent(parent, b) : uf_find(parent, b);
    if (ra == rb) return;
    if (ra < rb) {
      const uint32_t t = ra;
      ra = rb;
      rb = t;
    }
    // ra > rb: hang ra under rb if ra is still a root
    const uint32_t old = atomicCAS(parent + ra, ra, rb);
    if (old == ra) return;
    a = old;  // ra's current parent (smaller than ra)
    b = rb;
    fresh = true;
  }
}

__global__ void k_cc_init(uint32_t *parent, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = i;
}

// Union over every edge: waves stream over buckets 64 at a time; each slab
// is loaded once and its keys united with the bucket's owner lane-parallel.
template <bool MAP>
__global__ void __launch_bounds__(kBlock) k_cc_edges(GraphView G, uint32_t *parent) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t W = (gridDim.x * blockDim.x) >> 6;
  const bool key_lane = (key_mask<MAP>() >> lane) & 1ull;
  for (uint64_t b0 = uint64_t(gw) * 64; b0 < G.total_buckets; b0 += uint64_t(W) * 64) {
    const uint64_t b = b0 + lane;
    bool act = b < G.total_buckets;
    uint32_t u = 0;
    if (act) {
      u = G.bucket_owner[b];
      act = G.vdeg[u] != 0;
    }
    uint64_t m = ballot(act);
    while (m) {
      uint32_t js[4];
      uint32_t c = 0;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        js[k] = m ? ffs64(m) : 0u;
        if (m) {
          m &= m - 1;
          c = k + 1;
        }
      }
      uint32_t x[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        x[k] = k < c ? G.heads[(b0 + js[k]) * kSlabWords + lane] : kEmpty;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        if (k >= c) continue;
        const uint32_t uj = read_lane(u, js[k]);
        uint32_t v = x[k];
        for (;;) {
          if (key_lane && v < kTomb) uf_union(parent, uj, v);
          const uint32_t nx = read_lane(v, kNextLane);
          if (nx == kEmpty) break;
          v = slab_ptr(G.pool, nx)[lane];
        }
      }
    }
  }
}

// Union over a batch of edges (thread per edge).
__global__ void k_cc_batch(uint32_t *parent, const uint32_t *src, const uint32_t *dst,
                           uint32_t n, uint32_t nv) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t u = src[i], v = dst[i];
  if (u != v && u < nv && v < nv) uf_union(parent, u, v);
}

// label = root: every parent pointer jumps to its root.
__global__ void k_cc_compress(uint32_t *parent, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t r = parent[i];
  while (parent[r] != r) r = parent[r];
  parent[i] = r;
}

struct CC {
  mk_graph *g = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t *parent = nullptr;

  int init(mk_graph *graph) {
    g = graph;
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipStreamCreate(&stream));
    loop_stream_add(g->device, stream, 1);
    MK_HIP(hipEventCreate(&ev0));
    MK_HIP(hipEventCreate(&ev1));
    MK_HIP(hipMalloc(&parent, size_t(g->n) * 4));
    k_cc_init<<<blocks_for_threads(g->n), kBlock, 0, stream>>>(parent, g->n);
    MK_HIP(hipStreamSynchronize(stream));
    return MK_OK;
  }

  void destroy() {
    (void)hipSetDevice(g->device);
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(parent);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) {
      loop_stream_add(g->device, stream, -1);
      (void)hipStreamDestroy(stream);
    }
  }

  int finish(float *ms) {
    {
      // per vertex: parent read + root read + write
      ProfScope prof(stream, "k_cc_compress", 12.0 * g->n);
      k_cc_compress<<<blocks_for_threads(g->n), kBlock, 0, stream>>>(parent, g->n);
    }
    MK_HIP(hipGetLastError());
    MK_HIP(hipEventRecord(ev1, stream));
    MK_HIP(hipEventSynchronize(ev1));
    float t = 0.f;
    MK_HIP(hipEventElapsedTime(&t, ev0, ev1));
    if (ms) *ms = t;
    return MK_OK;
  }

  int run_static(float *ms) {
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipEventRecord(ev0, stream));
    k_cc_init<<<blocks_for_threads(g->n), kBlock, 0, stream>>>(parent, g->n);
    const unsigned grid = blocks_for_waves(std::min<uint64_t>(
        (uint64_t(g->total_buckets) + 63) / 64, uint64_t(device_cus(g->device)) * 16));
    {
      // per bucket: owner + degree (8 B) and its slabs; per edge two root
      // reads (the 20 B/edge union estimate)
      uint64_t m = 0;
      if (profiling_on()) (void)mk_graph_num_edges(g, &m);
      ProfScope prof(stream, "k_cc_edges", 264.0 * g->total_buckets + 20.0 * double(m));
      if (g->weighted)
        k_cc_edges<true><<<grid, kBlock, 0, stream>>>(g->view(), parent);
      else
        k_cc_edges<false><<<grid, kBlock, 0, stream>>>(g->view(), parent);
    }
    MK_HIP(hipGetLastError());
    return finish(ms);
  }

  int run_incremental(const uint32_t *src, const uint32_t *dst, uint32_t n, float *ms) {
    if (n && (!src || !dst)) return fail(MK_ERR_BAD_ARG, "cc incremental: NULL batch");
    MK_HIP(hipSetDevice(g->device));
    MK_HIP(hipEventRecord(ev0, stream));
    if (n) {
      ProfScope prof(stream, "k_cc_batch", 20.0 * n);
      k_cc_batch<<<blocks_for_threads(n), kBlock, 0, stream>>>(parent, src, dst, n, g->n);
    }
    MK_HIP(hipGetLastError());
    return finish(ms);
  }
};

}  // namespace mk

struct mk_cc {
  mk::CC c;
};

using namespace mk;

extern "C" {

int mk_cc_create(mk_graph *g, mk_cc **out) {
  if (!out) return fail(MK_ERR_BAD_ARG, "mk_cc_create: out is NULL");
  *out = nullptr;
  if (!g) return fail(MK_ERR_BAD_ARG, "mk_cc_create: NULL graph");
  mk_cc *h = new mk_cc();
  int r = h->c.init(g);
  if (r) {
    h->c.destroy();
    delete h;
    return r;
  }
  g->users++;
  *out = h;
  return MK_OK;
}

int mk_cc_destroy(mk_cc *c) {
  if (!c) return fail(MK_ERR_BAD_ARG, "mk_cc_destroy: NULL handle");
  c->c.g->users--;
  c->c.destroy();
  delete c;
  return MK_OK;
}

int mk_cc_static(mk_cc *c, float *ms) {
  if (!c) return fail(MK_ERR_BAD_ARG, "mk_cc_static: NULL handle");
  if (int r = graph_settle(c->c.g)) return r;
  return c->c.run_static(ms);
}

int mk_cc_incremental(mk_cc *c, const uint32_t *src, const uint32_t *dst, uint32_t n,
                      float *ms) {
  if (!c) return fail(MK_ERR_BAD_ARG, "mk_cc_incremental: NULL handle");
  if (int r = graph_settle(c->c.g)) return r;
  if (int r = graph_failed_batch(c->c.g, src, dst, n)) return r;
  return c->c.run_incremental(src, dst, n, ms);
}

int mk_cc_decremental(mk_cc *c, const uint32_t *src, const uint32_t *dst, uint32_t n,
                      float *ms) {
  (void)src;
  (void)dst;
  (void)n;
  if (!c) return fail(MK_ERR_BAD_ARG, "mk_cc_decremental: NULL handle");
  if (int r = graph_settle(c->c.g)) return r;
  return c->c.run_static(ms);
}

int mk_cc_labels(mk_cc *c, uint32_t *labels) {
  if (!c || !labels) return fail(MK_ERR_BAD_ARG, "mk_cc_labels: NULL argument");
  MK_HIP(hipSetDevice(c->c.g->device));
  MK_HIP(hipStreamSynchronize(c->c.stream));
  MK_HIP(hipMemcpy(labels, c->c.parent, size_t(c->c.g->n) * 4, hipMemcpyDeviceToHost));
  return MK_OK;
}

}  // extern "C"
