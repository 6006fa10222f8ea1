// SlabGraph.hpp - C++ classes with the reference's names over the C ABI of
// libmeerkat_hip.so: the slab allocator, the edge policies and
// DynamicGraph (include/Graph/SlabGraph.cuh:382-484, 633-743;
// include/Graph/SlabGraphImpl.cuh:6-62 in the reference tree).
//
// A program written against the reference's Graph API compiles against
// this header with its includes changed; the differences are:
//  - a non-zero status throws meerkat::Error (a std::runtime_error), where
//    the reference prints and exits (CHECK_ERROR);
//  - InsertEdges / DeleteEdges complete before they return (the reference
//    launches them asynchronously and the caller synchronises; a
//    synchronisation after the call stays correct);
//  - UpdateSlabPointers / UpdatedEdges track inserted keys by a per-slab
//    mask instead of a position marker (a tombstoned slot is reused here, so
//    slabs are not append-only); a graph tracks them when created with
//    track_updates;
//  - as in the reference (include/Graph/warp/InsertEdge.cuh:125-165), a
//    graph with a Rev* policy stores each edge the caller inserts or deletes
//    reversed: (src, dst) goes to dst's adjacency as key src.  PageRank reads
//    such a graph as its in-adjacency.
#ifndef MEERKAT_SLABGRAPH_HPP_
#define MEERKAT_SLABGRAPH_HPP_

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../meerkat.h"

namespace meerkat {

// A non-zero status of the C ABI; code() is the MK_ERR_* value.
class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char *call) {
  if (rc != MK_OK) throw Error(rc, std::string(call) + ": " + mk_last_error());
}

// The slab pool: 256-byte slabs handed out from an HBM-resident bitmap.
// Replaces the SlabHash dynamic allocator the reference constructs with
// `new Allocator` (src/BFS.cu:53-55) and passes by reference to every graph.
class SlabPool {
 public:
  explicit SlabPool(uint64_t bytes = 0, int device = 0) {
    check(mk_pool_create(bytes, device, &pool_), "mk_pool_create");
  }
  ~SlabPool() {
    if (pool_) (void)mk_pool_destroy(pool_);
  }
  SlabPool(const SlabPool &) = delete;
  SlabPool &operator=(const SlabPool &) = delete;

  mk_pool *handle() const { return pool_; }
  // (slabs in use, slabs in the pool)
  std::pair<uint64_t, uint64_t> Stats() const {
    uint64_t used = 0, total = 0;
    check(mk_pool_stats(pool_, &used, &total), "mk_pool_stats");
    return {used, total};
  }

 private:
  mk_pool *pool_ = nullptr;
};

// FullAllocatorPolicy<LogMemBlocks, SuperBlocks, Replicas> as named in the
// reference drivers (src/BFS.cu:53).  The pool holds as many slabs as that
// allocator's configuration: SuperBlocks * 2^LogMemBlocks memory blocks of
// 1024 slabs (<9, 48, 1>: 25,165,824 slabs, 6 GiB at 256 B per slab).
template <uint32_t LogMemBlocks = 9, uint32_t SuperBlocks = 48, uint32_t Replicas = 1>
struct FullAllocatorPolicy {
  static constexpr uint64_t kSlabs = (uint64_t(SuperBlocks) << LogMemBlocks) * 1024u;
  struct DynamicAllocatorT : SlabPool {
    explicit DynamicAllocatorT(int device = 0) : SlabPool(kSlabs * 256u, device) {}
  };
};
// LightAllocatorPolicy<8, 32, 1> (src/PageRank.cu:31, src/TC.cu:29): the same
// slab count rule (<8, 32, 1>: 8,388,608 slabs, 2 GiB).
template <uint32_t LogMemBlocks = 8, uint32_t SuperBlocks = 32, uint32_t Replicas = 1>
struct LightAllocatorPolicy : FullAllocatorPolicy<LogMemBlocks, SuperBlocks, Replicas> {};

// Edge policies (SlabGraph.cuh:92-240): adjacency kind only.
template <typename AllocPolicy, typename VertexTy = uint32_t>
struct UnweightedEdgePolicyT {
  using AllocPolicyT = AllocPolicy;
  using VertexT = VertexTy;
  using EdgeValueT = uint32_t;
  using CountT = uint32_t;
  static constexpr bool IsWeighted = false;
  static constexpr bool IsReversed = false;
};
template <typename AllocPolicy, typename VertexTy = uint32_t>
struct RevUnweightedEdgePolicyT : UnweightedEdgePolicyT<AllocPolicy, VertexTy> {
  static constexpr bool IsReversed = true;
};
template <typename AllocPolicy, typename VertexTy = uint32_t, typename ValueTy = uint32_t>
struct WeightedEdgePolicyMapT {
  using AllocPolicyT = AllocPolicy;
  using VertexT = VertexTy;
  using EdgeValueT = ValueTy;
  using CountT = uint32_t;
  static constexpr bool IsWeighted = true;
  static constexpr bool IsReversed = false;
};
template <typename AllocPolicy, typename VertexTy = uint32_t, typename ValueTy = uint32_t>
struct RevWeightedEdgePolicyMapT : WeightedEdgePolicyMapT<AllocPolicy, VertexTy, ValueTy> {
  static constexpr bool IsReversed = true;
};

// Per-vertex slab-hash adjacency (ConcurrentMap dst -> weight when
// weighted, ConcurrentSet otherwise).  Common part of both specialisations.
class GraphBase {
 public:
  GraphBase(uint32_t N, SlabPool &pool, double LoadFactor, const uint32_t *VertexDegreeHints,
            uint32_t DeviceIdx, bool weighted, bool mirror, bool track_updates = false)
      : n_(N) {
    (void)DeviceIdx;  // the pool's device is the graph's device
    uint32_t flags = (weighted ? MK_GRAPH_WEIGHTED : 0u) | (mirror ? MK_GRAPH_REVERSE : 0u) |
                     (track_updates ? MK_GRAPH_TRACK_UPDATES : 0u);
    check(mk_graph_create_ex(pool.handle(), N, flags, LoadFactor, VertexDegreeHints, nullptr, &g_),
          "mk_graph_create_ex");
  }
  ~GraphBase() {
    if (g_) (void)mk_graph_destroy(g_);
  }
  GraphBase(const GraphBase &) = delete;
  GraphBase &operator=(const GraphBase &) = delete;

  mk_graph *handle() const { return g_; }
  uint32_t NumVertices() const { return n_; }
  uint64_t NumEdges() const {
    uint64_t m = 0;
    check(mk_graph_num_edges(g_, &m), "mk_graph_num_edges");
    return m;
  }
  // GetVertexBucketsOffsets (SlabGraph.cuh:483): n + 1 bucket offsets.
  uint32_t *GetVertexBucketsOffsets() {
    if (offsets_.empty()) {
      offsets_.resize(size_t(n_) + 1);
      check(mk_graph_bucket_offsets(g_, offsets_.data()), "mk_graph_bucket_offsets");
    }
    return offsets_.data();
  }
  // GetBucketsPerVertex (SlabGraph.cuh:484).
  uint32_t *GetBucketsPerVertex() {
    if (per_vertex_.empty()) {
      const uint32_t *off = GetVertexBucketsOffsets();
      per_vertex_.resize(n_);
      for (uint32_t v = 0; v < n_; ++v) per_vertex_[v] = off[v + 1] - off[v];
    }
    return per_vertex_.data();
  }
  // Live out-degrees (the reference's VertexDegrees array, SlabGraph.cuh:392).
  std::vector<uint32_t> VertexDegrees() const {
    std::vector<uint32_t> d(n_);
    check(mk_graph_degrees(g_, d.data()), "mk_graph_degrees");
    return d;
  }
  // Batched SearchEdge (include/Graph/warp/SearchEdge.cuh:5-63): found_dev[i]
  // = 1 when (src[i], dst[i]) is present.
  void SearchEdges(const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                   uint8_t *found_dev, uint32_t *w_out_dev = nullptr) {
    check(mk_graph_search(g_, src_dev, dst_dev, n, found_dev, w_out_dev, nullptr),
          "mk_graph_search");
  }
  // UpdateSlabPointers (SlabGraphImpl.cuh:64-82): edges inserted so far stop
  // being reported by UpdatedEdges (a graph created with track_updates).
  void UpdateSlabPointers() {
    check(mk_graph_update_slab_pointers(g_), "mk_graph_update_slab_pointers");
  }
  // The update-aware iteration: the edges inserted since the last
  // UpdateSlabPointers and still present, into caller-owned device arrays of
  // `capacity` entries; returns the total (a capacity of 0 only counts).
  uint64_t UpdatedEdges(uint32_t *src_dev, uint32_t *dst_dev, uint32_t *w_dev, uint64_t capacity) {
    uint64_t count = 0;
    check(mk_graph_updated_edges(g_, src_dev, dst_dev, w_dev, capacity, &count),
          "mk_graph_updated_edges");
    return count;
  }

  // Deferred mutation checks (mk_graph_set_deferred): InsertEdges /
  // DeleteEdges return once queued; a failed batch throws from the graph's
  // next call, or from the BFS/SSSP call that follows it.  Sync() settles.
  void SetDeferred(bool on) { check(mk_graph_set_deferred(g_, on ? 1 : 0), "mk_graph_set_deferred"); }
  void Sync() { check(mk_graph_sync(g_), "mk_graph_sync"); }

 protected:
  mk_graph *g_ = nullptr;
  uint32_t n_;
  std::vector<uint32_t> offsets_, per_vertex_;
};

template <typename EdgePolicy, bool IsWeighted>
class DynamicGraph;

// Weighted: ConcurrentMap adjacency (SlabGraph.cuh:382-466).
template <typename EdgePolicy>
class DynamicGraph<EdgePolicy, true> : public GraphBase {
 public:
  using VertexT = typename EdgePolicy::VertexT;
  using EdgeValueT = typename EdgePolicy::EdgeValueT;
  using CountT = typename EdgePolicy::CountT;
  using EdgeDynAllocator = typename EdgePolicy::AllocPolicyT::DynamicAllocatorT;

  // `mirror` (no reference counterpart) also keeps an in-edge mirror, which
  // decremental BFS/SSSP use to re-seed and PageRank as its in-adjacency.
  DynamicGraph(uint32_t N, SlabPool &TheEdgeAllocator, double LoadFactor,
               uint32_t *VertexDegreeHints, uint32_t DeviceIdx = 0, bool mirror = false,
               bool track_updates = false)
      : GraphBase(N, TheEdgeAllocator, LoadFactor, VertexDegreeHints, DeviceIdx, true, mirror,
                  track_updates) {}

  // SlabGraphImpl.cuh:6-20.  EdgeValues may be null (weight 0).
  void InsertEdges(VertexT *SourceVertices, VertexT *DestinationVertices, CountT NumberOfEdges,
                   EdgeValueT *EdgeValues) {
    if (EdgePolicy::IsReversed) std::swap(SourceVertices, DestinationVertices);
    check(mk_graph_insert(g_, SourceVertices, DestinationVertices, EdgeValues, NumberOfEdges,
                          nullptr),
          "mk_graph_insert");
  }
  // SlabGraphImpl.cuh:36-48.
  void DeleteEdges(VertexT *SourceVertices, VertexT *DestinationVertices, CountT NumberOfEdges) {
    if (EdgePolicy::IsReversed) std::swap(SourceVertices, DestinationVertices);
    check(mk_graph_delete(g_, SourceVertices, DestinationVertices, NumberOfEdges, nullptr),
          "mk_graph_delete");
  }
  // UpdateEdges (SlabGraphImpl.cuh:84-103, compiled out in the reference):
  // present edges take NewEdgeValues[i]; absent ones are not inserted.  The
  // reference's FilterMaps argument (a PhaseConcurrentMap detail of the
  // absent SlabHash) is accepted only as null.
  template <typename FilterMapTy = void>
  void UpdateEdges(VertexT *SourceVertices, VertexT *DestinationVertices, CountT NumberOfEdges,
                   EdgeValueT *NewEdgeValues, FilterMapTy *EdgeFilterMaps = nullptr) {
    if (EdgeFilterMaps != nullptr)
      throw Error(MK_ERR_BAD_ARG, "UpdateEdges: FilterMaps are not supported");
    if (EdgePolicy::IsReversed) std::swap(SourceVertices, DestinationVertices);
    check(mk_graph_update_weights(g_, SourceVertices, DestinationVertices, NewEdgeValues,
                                  NumberOfEdges, nullptr),
          "mk_graph_update_weights");
  }
};

// Unweighted: ConcurrentSet adjacency (SlabGraph.cuh:633-717).
template <typename EdgePolicy>
class DynamicGraph<EdgePolicy, false> : public GraphBase {
 public:
  using VertexT = typename EdgePolicy::VertexT;
  using CountT = typename EdgePolicy::CountT;
  using EdgeDynAllocator = typename EdgePolicy::AllocPolicyT::DynamicAllocatorT;

  DynamicGraph(uint32_t N, SlabPool &TheEdgeAllocator, double LoadFactor,
               uint32_t *VertexDegreeHints, uint32_t DeviceIdx = 0, bool mirror = false,
               bool track_updates = false)
      : GraphBase(N, TheEdgeAllocator, LoadFactor, VertexDegreeHints, DeviceIdx, false, mirror,
                  track_updates) {}

  // SlabGraphImpl.cuh:22-34.
  void InsertEdges(VertexT *SourceVertices, VertexT *DestinationVertices, CountT NumberOfEdges) {
    if (EdgePolicy::IsReversed) std::swap(SourceVertices, DestinationVertices);
    check(mk_graph_insert(g_, SourceVertices, DestinationVertices, nullptr, NumberOfEdges,
                          nullptr),
          "mk_graph_insert");
  }
  // SlabGraphImpl.cuh:50-62.
  void DeleteEdges(VertexT *SourceVertices, VertexT *DestinationVertices, CountT NumberOfEdges) {
    if (EdgePolicy::IsReversed) std::swap(SourceVertices, DestinationVertices);
    check(mk_graph_delete(g_, SourceVertices, DestinationVertices, NumberOfEdges, nullptr),
          "mk_graph_delete");
  }
};

// DynGraph<Policy> (SlabGraph.cuh:768-769).
template <typename EdgePolicy>
using DynGraph = DynamicGraph<EdgePolicy, EdgePolicy::IsWeighted>;

}  // namespace meerkat

#endif  // MEERKAT_SLABGRAPH_HPP_
