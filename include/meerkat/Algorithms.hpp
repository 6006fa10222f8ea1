// Algorithms.hpp - the reference's algorithm classes (include/Algorithms/
// BFS.cuh, SSSP.cuh, PageRank.cuh, TC.cuh in the reference tree) over the C
// ABI of libmeerkat_hip.so, plus CC and MultiSource, which have no class in
// the reference.
//
// Same construction, protocol and returns as the reference: mutate the
// graph first, then call Incremental / Decremental with the same device
// batch; every compute call blocks and returns elapsed milliseconds.
// Differences: Distances() returns the real uint32 distances (the
// reference copies sizeof(float) * N bytes out of a uint64 array,
// BFS.cuh:55-60) and Parents() is added; errors throw meerkat::Error.
#ifndef MEERKAT_ALGORITHMS_HPP_
#define MEERKAT_ALGORITHMS_HPP_

#include <cstdint>
#include <memory>
#include <vector>

#include "SlabGraph.hpp"

namespace meerkat {

// BFS<AllocPolicy> (BFS.cuh:14-68).
template <typename AllocPolicy>
struct BFS {
  using EdgePolicy = UnweightedEdgePolicyT<AllocPolicy>;
  using GraphT = DynGraph<EdgePolicy>;

  BFS(GraphT &G, uint32_t SourceVertex, uint32_t VertexN) : n_(VertexN) {
    check(mk_bfs_create(G.handle(), SourceVertex, &h_), "mk_bfs_create");
  }
  ~BFS() {
    if (h_) (void)mk_bfs_destroy(h_);
  }
  BFS(const BFS &) = delete;
  BFS &operator=(const BFS &) = delete;

  float Static(uint32_t EdgesN = 0) {
    (void)EdgesN;  // frontier buffers are sized from the graph
    float ms = 0.f;
    check(mk_bfs_static(h_, &ms), "mk_bfs_static");
    return ms;
  }
  float Incremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_bfs_incremental(h_, EdgesSrcDev, EdgesDstDev, EdgesN, &ms), "mk_bfs_incremental");
    return ms;
  }
  float Decremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_bfs_decremental(h_, EdgesSrcDev, EdgesDstDev, EdgesN, &ms), "mk_bfs_decremental");
    return ms;
  }
  std::vector<uint32_t> Distances() { return read(true); }
  std::vector<uint32_t> Parents() { return read(false); }
  void SetSource(uint32_t source) { check(mk_bfs_set_source(h_, source), "mk_bfs_set_source"); }
  // Deferred calls (mk_bfs_set_deferred): Incremental / Decremental return 0
  // once queued; the next call on the handle, or Sync(), settles them.
  void SetDeferred(bool on) { check(mk_bfs_set_deferred(h_, on ? 1 : 0), "mk_bfs_set_deferred"); }
  void Sync() { check(mk_bfs_sync(h_), "mk_bfs_sync"); }

 private:
  std::vector<uint32_t> read(bool dist) {
    std::vector<uint32_t> d(n_), p(n_);
    check(mk_bfs_read(h_, d.data(), p.data()), "mk_bfs_read");
    return dist ? d : p;
  }
  mk_bfs *h_ = nullptr;
  uint32_t n_;
};

// SSSP<AllocPolicy> (SSSP.cuh:10-69).  Weights are read from the graph; the
// weight pointers of Incremental / Decremental may be null.
template <typename AllocPolicy>
struct SSSP {
  using EdgePolicy = WeightedEdgePolicyMapT<AllocPolicy>;
  using GraphT = DynGraph<EdgePolicy>;

  SSSP(GraphT &G, uint32_t SourceVertex, uint32_t VertexN) : n_(VertexN) {
    check(mk_sssp_create(G.handle(), SourceVertex, &h_), "mk_sssp_create");
  }
  ~SSSP() {
    if (h_) (void)mk_sssp_destroy(h_);
  }
  SSSP(const SSSP &) = delete;
  SSSP &operator=(const SSSP &) = delete;

  float Static(uint32_t EdgesN = 0) {
    (void)EdgesN;
    float ms = 0.f;
    check(mk_sssp_static(h_, &ms), "mk_sssp_static");
    return ms;
  }
  float Incremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t *EdgesWgtDev,
                    uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_sssp_incremental(h_, EdgesSrcDev, EdgesDstDev, EdgesWgtDev, EdgesN, &ms),
          "mk_sssp_incremental");
    return ms;
  }
  float Decremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t *EdgesWgtDev,
                    uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_sssp_decremental(h_, EdgesSrcDev, EdgesDstDev, EdgesWgtDev, EdgesN, &ms),
          "mk_sssp_decremental");
    return ms;
  }
  // After DynamicGraph::UpdateEdges with the same batch (no reference
  // counterpart): heavier tree edges are handled as deletions, lighter edges
  // as insertions (mk_sssp_reweight).
  float Reweight(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t *NewWgtDev,
                 uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_sssp_reweight(h_, EdgesSrcDev, EdgesDstDev, NewWgtDev, EdgesN, &ms),
          "mk_sssp_reweight");
    return ms;
  }
  std::vector<uint32_t> Distances() { return read(true); }
  std::vector<uint32_t> Parents() { return read(false); }
  void SetSource(uint32_t source) { check(mk_sssp_set_source(h_, source), "mk_sssp_set_source"); }
  // As BFS::SetDeferred / BFS::Sync.
  void SetDeferred(bool on) {
    check(mk_sssp_set_deferred(h_, on ? 1 : 0), "mk_sssp_set_deferred");
  }
  void Sync() { check(mk_sssp_sync(h_), "mk_sssp_sync"); }

 private:
  std::vector<uint32_t> read(bool dist) {
    std::vector<uint32_t> d(n_), p(n_);
    check(mk_sssp_read(h_, d.data(), p.data()), "mk_sssp_read");
    return dist ? d : p;
  }
  mk_sssp *h_ = nullptr;
  uint32_t n_;
};

// PageRank<AllocPolicy> (PageRank.cuh:9-68).  G is a Rev-policy graph (its
// InsertEdges stores each edge reversed, as the reference's GraphT) with
// VertexOutDegree the out-degrees; or G was built with an in-edge mirror
// (then the out-degrees are exact and VertexOutDegree may be null).
// Incremental / Decremental take the batch in its original direction.
template <typename AllocPolicy>
struct PageRank {
  using EdgePolicy = RevUnweightedEdgePolicyT<AllocPolicy>;
  using GraphT = DynGraph<EdgePolicy>;

  PageRank(GraphBase &G, float DampingFactor, float ErrorMargin, uint32_t VertexN,
           uint32_t *VertexOutDegree, uint32_t MaxIter = 100)
      : n_(VertexN) {
    check(mk_pagerank_create(G.handle(), DampingFactor, ErrorMargin, MaxIter, VertexOutDegree,
                             &h_),
          "mk_pagerank_create");
  }
  ~PageRank() {
    if (h_) (void)mk_pagerank_destroy(h_);
  }
  PageRank(const PageRank &) = delete;
  PageRank &operator=(const PageRank &) = delete;

  float Static() {
    float ms = 0.f;
    check(mk_pagerank_static(h_, &ms), "mk_pagerank_static");
    return ms;
  }
  float Incremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_pagerank_incremental(h_, EdgesSrcDev, EdgesDstDev, EdgesN, &ms),
          "mk_pagerank_incremental");
    return ms;
  }
  float Decremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_pagerank_decremental(h_, EdgesSrcDev, EdgesDstDev, EdgesN, &ms),
          "mk_pagerank_decremental");
    return ms;
  }
  std::vector<float> Values() {
    std::vector<float> v(n_);
    check(mk_pagerank_read(h_, v.data()), "mk_pagerank_read");
    return v;
  }

 private:
  mk_pagerank *h_ = nullptr;
  uint32_t n_;
};

// TC<AllocPolicy> (TC.cuh:9-40).  The Allocator argument of Incremental /
// Decremental is accepted for signature parity; no temporary graph is built.
template <typename AllocPolicy>
struct TC {
  using EdgePolicy = UnweightedEdgePolicyT<AllocPolicy>;
  using GraphT = DynGraph<EdgePolicy>;
  using Allocator = typename AllocPolicy::DynamicAllocatorT;

  TC(GraphT &G, uint32_t VertexN) {
    (void)VertexN;
    check(mk_tc_create(G.handle(), &h_), "mk_tc_create");
  }
  ~TC() {
    if (h_) (void)mk_tc_destroy(h_);
  }
  TC(const TC &) = delete;
  TC &operator=(const TC &) = delete;

  unsigned long long int GetCount() const {
    uint64_t c = 0;
    check(mk_tc_count(h_, &c), "mk_tc_count");
    return c;
  }
  float Static(uint32_t EdgesN = 0) {
    (void)EdgesN;
    float ms = 0.f;
    check(mk_tc_static(h_, &ms), "mk_tc_static");
    return ms;
  }
  float Incremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgeN,
                    std::unique_ptr<Allocator> &) {
    float ms = 0.f;
    check(mk_tc_incremental(h_, EdgesSrcDev, EdgesDstDev, EdgeN, &ms), "mk_tc_incremental");
    return ms;
  }
  float Decremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgeN,
                    std::unique_ptr<Allocator> &) {
    float ms = 0.f;
    check(mk_tc_decremental(h_, EdgesSrcDev, EdgesDstDev, EdgeN, &ms), "mk_tc_decremental");
    return ms;
  }

 private:
  mk_tc *h_ = nullptr;
};

// Connected components by union-find (the reference has only the
// primitives, include/Util/UnionFind.cuh:24-253).  Labels are the smallest
// vertex id of each component.
template <typename AllocPolicy>
struct CC {
  using EdgePolicy = UnweightedEdgePolicyT<AllocPolicy>;

  CC(GraphBase &G, uint32_t VertexN) : n_(VertexN) {
    check(mk_cc_create(G.handle(), &h_), "mk_cc_create");
  }
  ~CC() {
    if (h_) (void)mk_cc_destroy(h_);
  }
  CC(const CC &) = delete;
  CC &operator=(const CC &) = delete;

  float Static() {
    float ms = 0.f;
    check(mk_cc_static(h_, &ms), "mk_cc_static");
    return ms;
  }
  float Incremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_cc_incremental(h_, EdgesSrcDev, EdgesDstDev, EdgesN, &ms), "mk_cc_incremental");
    return ms;
  }
  float Decremental(uint32_t *EdgesSrcDev, uint32_t *EdgesDstDev, uint32_t EdgesN) {
    float ms = 0.f;
    check(mk_cc_decremental(h_, EdgesSrcDev, EdgesDstDev, EdgesN, &ms), "mk_cc_decremental");
    return ms;
  }
  std::vector<uint32_t> Labels() {
    std::vector<uint32_t> l(n_);
    check(mk_cc_labels(h_, l.data()), "mk_cc_labels");
    return l;
  }

 private:
  mk_cc *h_ = nullptr;
  uint32_t n_;
};

// Up to 64 BFS or SSSP queries in one traversal (no reference class).
class MultiSource {
 public:
  MultiSource(GraphBase &G, bool sssp) : n_(G.NumVertices()) {
    check(mk_msq_create(G.handle(), sssp ? 1 : 0, &h_), "mk_msq_create");
  }
  ~MultiSource() {
    if (h_) (void)mk_msq_destroy(h_);
  }
  MultiSource(const MultiSource &) = delete;
  MultiSource &operator=(const MultiSource &) = delete;

  float Run(const std::vector<uint32_t> &sources) {
    float ms = 0.f;
    check(mk_msq_run(h_, sources.data(), uint32_t(sources.size()), &ms), "mk_msq_run");
    return ms;
  }
  std::vector<uint32_t> Distances(uint32_t slot) {
    std::vector<uint32_t> d(n_);
    check(mk_msq_read(h_, slot, d.data()), "mk_msq_read");
    return d;
  }
  // Continuous batching: sources join a running traversal in free slots and
  // each slot finishes on its own step (mk_msq_open / admit / launch / poll).
  void Open() { check(mk_msq_open(h_), "mk_msq_open"); }
  void Admit(const std::vector<uint32_t> &sources, const std::vector<uint32_t> &slots) {
    if (sources.size() != slots.size()) throw std::runtime_error("Admit: size mismatch");
    check(mk_msq_admit(h_, sources.data(), slots.data(), uint32_t(sources.size())),
          "mk_msq_admit");
  }
  void Launch(uint32_t steps = 1) { check(mk_msq_launch(h_, steps), "mk_msq_launch"); }
  // Waits for the launched steps; the admitted slots with work left.
  uint64_t Poll() {
    uint64_t active = 0;
    check(mk_msq_poll(h_, &active, nullptr), "mk_msq_poll");
    return active;
  }

 private:
  mk_msq *h_ = nullptr;
  uint32_t n_;
};

}  // namespace meerkat

#endif  // MEERKAT_ALGORITHMS_HPP_
