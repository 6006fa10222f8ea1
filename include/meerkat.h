/*
 * meerkat.h - C ABI of libmeerkat_hip.so, the MI355X (gfx950) dynamic-graph
 * library.
 *
 * The reference (Meerkat) is a header-only CUDA C++ template library; its
 * "boundary" is the set of class methods its drivers and tests call.  Each
 * entry point below replaces one of those methods and cites it (paths are
 * relative to the reference tree).  The C++ classes in include/meerkat/ (SlabGraph.hpp, Algorithms.hpp)
 * restore the reference's class and method names on top of this ABI.
 *
 * Conventions
 *  - Every function returns an int status: MK_OK (0) or one of MK_ERR_*.
 *    mk_last_error() returns the calling thread's last error text.
 *  - Pointers named *_dev are device (HBM) pointers owned by the caller and
 *    only read during the call; *_host pointers are host memory.
 *  - Handles own their device state and are externally serialised (not
 *    thread-safe); distinct handles may be driven from distinct threads.
 *  - A pool must outlive every graph built on it, and a graph every
 *    algorithm handle built on it (checked: destroy returns MK_ERR_STATE).
 *  - Vertex ids are uint32 in [0, n); 0xFFFFFFFF and 0xFFFFFFFE are reserved
 *    (empty / tombstone slab words).
 *  - `stream` arguments are hipStream_t values passed as void*.  Graph
 *    mutations are ordered on that stream and complete before the call
 *    returns (unless deferred: mk_graph_set_deferred); NULL selects the
 *    graph's own stream, which (created with default flags) still orders
 *    after work queued on the null stream.  Algorithm compute calls block
 *    until done and report elapsed milliseconds (HIP events), like the
 *    reference's float returns (include/Algorithms/impl/BFSImpl.cuh:644-661),
 *    except deferred BFS/SSSP calls (mk_bfs_set_deferred).
 */
#ifndef MEERKAT_H_
#define MEERKAT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MK_OK 0
#define MK_ERR_BAD_ARG 1   /* null handle/pointer, bad size or parameter */
#define MK_ERR_POOL_FULL 2 /* slab pool exhausted; graph edge set unchanged */
#define MK_ERR_HIP 3       /* a HIP runtime call failed */
#define MK_ERR_STATE 4     /* call out of protocol order, or handle in use */
#define MK_ERR_RANGE 5     /* vertex id >= n in a batch; graph unchanged */

/* Per-edge outcome codes written by mk_graph_insert / mk_graph_delete. */
#define MK_EDGE_UNCHANGED 0 /* insert: already present; delete: absent   */
#define MK_EDGE_APPLIED 1   /* inserted as new / deleted                  */
#define MK_EDGE_SELF_LOOP 2 /* skipped (InsertEdgeKernel.cuh:25,54)       */
#define MK_EDGE_POOL_FULL 3 /* not inserted: pool exhausted (rolled back)  */

typedef struct mk_pool mk_pool;
typedef struct mk_graph mk_graph;
typedef struct mk_bfs mk_bfs;
typedef struct mk_sssp mk_sssp;
typedef struct mk_pagerank mk_pagerank;
typedef struct mk_tc mk_tc;
typedef struct mk_cc mk_cc;
typedef struct mk_msq mk_msq;

/* Last error text of the calling thread ("" when none).  Replaces the
 * reference's CHECK_ERROR print-and-exit (SlabHash macros used at e.g.
 * include/Graph/SlabGraph.cuh:392-440). */
const char *mk_last_error(void);
/* ABI version, major*10000 + minor*100 + patch. */
int mk_version(void);
/* Device facts used by launch sizing: compute units and wavefront width. */
int mk_device_info(int device, int *compute_units, int *wave_size);

/* ---------------------------------------------------------------- A1 pool
 * Slab pool: 256-byte slabs handed out from an HBM-resident bitmap.  Replaces
 * AllocPolicy::DynamicAllocatorT (default-constructed at src/BFS.cu:53-55,
 * held by reference in include/Graph/SlabGraph.cuh:369,620).  `bytes` = 0
 * picks 1 GiB.  Stats report slabs in use and capacity. */
int mk_pool_create(uint64_t bytes, int device, mk_pool **out);
int mk_pool_destroy(mk_pool *pool);
int mk_pool_stats(mk_pool *pool, uint64_t *used_slabs, uint64_t *total_slabs);

/* --------------------------------------------------------------- A2 graph
 * DynamicGraph(N, Allocator&, LoadFactor, VertexDegreeHints, DeviceIdx)
 * (include/Graph/SlabGraph.cuh:382-466 weighted, 633-717 unweighted).
 * weighted != 0 gives ConcurrentMap adjacency (dst -> uint32 weight), else
 * ConcurrentSet.  Buckets per vertex = max(1, ceil(hint / (keys_per_slab *
 * load_factor))) with 63 keys per set slab and 31 pairs per map slab.
 * host_degree_hints may be NULL (one bucket per vertex). */
int mk_graph_create(mk_pool *pool, uint32_t n, int weighted, double load_factor,
                    const uint32_t *host_degree_hints, mk_graph **out);
int mk_graph_destroy(mk_graph *g);
/* Flags of mk_graph_create_ex. */
#define MK_GRAPH_WEIGHTED 1u /* ConcurrentMap adjacency (dst -> weight)        */
#define MK_GRAPH_REVERSE 2u  /* also keep an in-edge mirror (dst -> src [, w]) */
#define MK_GRAPH_TRACK_UPDATES 4u /* mark inserted keys for mk_graph_updated_edges */
/* DynamicGraph constructor with options.  MK_GRAPH_REVERSE keeps a second
 * slab-hash graph of in-edges on the same pool, updated by every
 * mk_graph_insert / mk_graph_delete; decremental BFS/SSSP then re-seed from
 * in-neighbours instead of scanning every edge.  in_hints sizes the mirror
 * (NULL: out_hints).  No reference counterpart (the reference builds a
 * separate reverse DynamicGraph for PageRank, include/Algorithms/
 * PageRank.cuh:11, and updates it by hand, src/PageRank.cu:137). */
int mk_graph_create_ex(mk_pool *pool, uint32_t n, uint32_t flags, double load_factor,
                       const uint32_t *out_hints, const uint32_t *in_hints, mk_graph **out);
/* The in-edge mirror of g (read-only: inspection calls only; its degrees,
 * the in-degrees of g, are counted from its slabs on request). */
int mk_graph_reverse(mk_graph *g, mk_graph **rev);

/* DynamicGraph::InsertEdges(src, dst, n[, w])
 * (include/Graph/SlabGraphImpl.cuh:6-34 -> device/InsertEdgeKernel.cuh:4-63).
 * Self-loops are skipped.  An edge already present is left unchanged (its
 * weight too).  w_dev may be NULL (weight 0, InsertEdgeKernel.cuh:24).  On
 * MK_ERR_RANGE or MK_ERR_POOL_FULL the edge set is left as it was. */
int mk_graph_insert(mk_graph *g, const uint32_t *src_dev, const uint32_t *dst_dev,
                    const uint32_t *w_dev, uint32_t n, void *stream);
/* DynamicGraph::DeleteEdges(src, dst, n)
 * (include/Graph/SlabGraphImpl.cuh:36-62 -> device/DeleteEdgeKernel.cuh:4-26).
 * A deleted key becomes a tombstone; later inserts reuse tombstoned slots. */
int mk_graph_delete(mk_graph *g, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                    void *stream);
/* Deferred mutation checks (no reference counterpart; the reference's
 * mutations are asynchronous and unchecked).  With on != 0, mk_graph_insert
 * and mk_graph_delete return once the batch is queued on the device; the
 * batch's range check and slab-pool outcome are settled by the graph's next
 * call (any mk_graph_* call that reads or changes it, mk_graph_sync, or an
 * algorithm compute call on it), which waits for the batch and, if it
 * failed, restores the graph as the synchronous call would have and returns
 * that batch's error (MK_ERR_RANGE / MK_ERR_POOL_FULL) instead of doing its
 * own work.  A BFS/SSSP compute call queued after the batch orders its
 * kernels after it without a host wait and settles it at its end; if the
 * batch failed, the call returns its error and the handle's results stay
 * invalid until its next Static.  Any later algorithm call (on any handle
 * of the graph) given a batch that failed, synchronous or deferred, returns
 * that batch's error instead of running on edges the graph never applied,
 * until the graph's next mutation.  The batch buffers and the stream must
 * stay valid until the batch is settled. */
int mk_graph_set_deferred(mk_graph *g, int on);
int mk_graph_sync(mk_graph *g);

/* DynamicGraph::UpdateSlabPointers() (include/Graph/SlabGraphImpl.cuh:64-82
 * -> device/UpdateSlablistPointers.cuh:4-104): the edges inserted so far stop
 * being "updated"; mk_graph_updated_edges then reports only edges inserted
 * after this call.  A no-op on a graph created without
 * MK_GRAPH_TRACK_UPDATES (as the reference's with nothing inserted). */
int mk_graph_update_slab_pointers(mk_graph *g);
/* The update-aware iteration the reference's markers serve
 * (SlabGraph.cuh:377-379, 425-439: a bucket is iterated from its first
 * updated slab and lane): every edge inserted since the last
 * mk_graph_update_slab_pointers (since creation before the first) and still
 * present, written as (src_dev[i], dst_dev[i][, w_dev[i]]) for i <
 * min(*count, capacity) in no particular order; *count receives the total,
 * so a call with capacity 0 sizes the buffers.  An edge deleted and
 * inserted again counts as inserted; a weight update does not.  Needs a
 * graph created with MK_GRAPH_TRACK_UPDATES (MK_ERR_STATE otherwise), whose
 * inserts then also mark each applied key (one more chain walk per applied
 * edge).  w_dev may be NULL; it is ignored on an unweighted graph. */
int mk_graph_updated_edges(mk_graph *g, uint32_t *src_dev, uint32_t *dst_dev, uint32_t *w_dev,
                           uint64_t capacity, uint64_t *count);

/* DynamicGraph::UpdateEdges(src, dst, n, NewEdgeValues, FilterMaps)
 * (include/Graph/SlabGraphImpl.cuh:84-103, compiled out with #if 0 in the
 * reference; device/UpdateEdgeKernel.cuh:7-37 -> warp/UpdateEdge.cuh:7-33)
 * on a weighted graph: each batch edge present in the graph (and its
 * in-edge mirror entry) takes the weight w_dev[i]; an absent edge is not
 * inserted; self-loops are skipped.  Per-edge outcomes (mk_graph_last_status):
 * MK_EDGE_APPLIED (present, weight set), MK_EDGE_UNCHANGED (absent),
 * MK_EDGE_SELF_LOOP.  A batch naming an edge twice keeps one of its
 * weights.  The reference's per-edge FilterMaps belong to the absent
 * SlabHash's PhaseConcurrentMap and have no counterpart.  Errors, streams
 * and deferred checks as mk_graph_insert (MK_ERR_RANGE leaves the graph
 * unchanged); MK_ERR_BAD_ARG on an unweighted graph or NULL w_dev. */
int mk_graph_update_weights(mk_graph *g, const uint32_t *src_dev, const uint32_t *dst_dev,
                            const uint32_t *w_dev, uint32_t n, void *stream);

/* Batched DynamicGraphContext::SearchEdge (include/Graph/warp/SearchEdge.cuh:
 * 5-63, 118-142): found_dev[i] = 1 if (src[i], dst[i]) is present; for a
 * weighted graph w_out_dev[i] (nullable) receives its weight. */
int mk_graph_search(mk_graph *g, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                    uint8_t *found_dev, uint32_t *w_out_dev, void *stream);
/* Device array of MK_EDGE_* codes for the most recent insert/delete batch.
 * No reference counterpart: the reference's SlabHash insertKey returns this
 * per lane (include/Graph/warp/InsertEdge.cuh:86-98) and discards it. */
int mk_graph_last_status(mk_graph *g, const uint8_t **status_dev, uint32_t *n);

/* Queries used by tests and algorithms.  VertexDegrees is the live
 * out-degree array (SlabGraph.cuh:392-395). */
int mk_graph_num_vertices(mk_graph *g, uint32_t *n);
int mk_graph_num_edges(mk_graph *g, uint64_t *m);
int mk_graph_is_weighted(mk_graph *g, int *weighted);
int mk_graph_degrees(mk_graph *g, uint32_t *deg_host);
/* GetVertexBucketsOffsets (SlabGraph.cuh:483-484), with n+1 entries. */
int mk_graph_bucket_offsets(mk_graph *g, uint32_t *off_host);
/* Per-vertex count of live keys and of tombstones found by walking every
 * slab (the counting kernels of test/GraphTest.cu:58-200). */
int mk_graph_count_slots(mk_graph *g, uint32_t *live_host, uint32_t *tomb_host);
/* mk_graph_count_slots without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot call hipMalloc / hipFree
 * (code under stream capture, tight serving loops, harnesses that forbid
 * allocation).  The per-vertex counts are staged in a fixed 32 MiB device
 * array of the library (one per device, held until the process ends)
 * instead of a buffer allocated per call, so graphs of more than 4,194,304
 * vertices are refused with MK_ERR_BAD_ARG.  Every call on a device shares
 * that array: the library serialises its own calls with a per-device lock,
 * and a call clears the counts it uses before counting.  Results and
 * errors otherwise as mk_graph_count_slots. */
int mk_graph_count_slots_noalloc(mk_graph *g, uint32_t *live_host, uint32_t *tomb_host);
/* Adjacency as CSR on the host: row_off_host[n+1], col_host[m] (and
 * w_host[m] when weighted and non-NULL), neighbours in slab order
 * (the collection kernels of test/GraphTest.cu:202-483). */
int mk_graph_export_csr(mk_graph *g, uint64_t *row_off_host, uint32_t *col_host, uint32_t *w_host);
/* Head slabs, overflow slabs in use by this graph, and longest chain. */
int mk_graph_slab_stats(mk_graph *g, uint64_t *head_slabs, uint64_t *overflow_slabs,
                        uint32_t *max_chain);

/* ------------------------------------------------------ A8-A10 BFS
 * BFS(G, Src, N) (include/Algorithms/BFS.cuh:34-53).  Static
 * (impl/BFSImpl.cuh:590-662), Incremental after InsertEdges with the same
 * batch (:664-753), Decremental after DeleteEdges with the same batch
 * (:755-846).  Distances are levels, UINT32_MAX when unreachable.  read
 * returns real uint32 distances and parents (the reference's Distances()
 * copies sizeof(float)*N bytes of a uint64 array, BFS.cuh:55-60). */
int mk_bfs_create(mk_graph *g, uint32_t source, mk_bfs **out);
int mk_bfs_destroy(mk_bfs *b);
int mk_bfs_static(mk_bfs *b, float *ms);
int mk_bfs_incremental(mk_bfs *b, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                       float *ms);
int mk_bfs_decremental(mk_bfs *b, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                       float *ms);
int mk_bfs_read(mk_bfs *b, uint32_t *dist_host, uint32_t *parent_host);
/* Rebind the source vertex (no reference counterpart: the reference fixes
 * it at construction).  Distances are those of the old source until the
 * next mk_bfs_static; lets one handle answer many single-source queries
 * without reallocating its n-sized buffers. */
int mk_bfs_set_source(mk_bfs *b, uint32_t source);
/* Frontier iterations, adjacency slabs and live keys (edges) read by the
 * last compute call; edges/ms gives its traversed-edges rate. */
int mk_bfs_last_stats(mk_bfs *b, uint32_t *iterations, uint64_t *slabs_read,
                      uint64_t *edges_read);
/* Deferred calls (no reference counterpart; the reference's compute calls
 * end with a device synchronisation).  With on != 0, mk_bfs_incremental and
 * mk_bfs_decremental return once the call is queued on the graph's own
 * stream (*ms = 0: not measured); the host does not wait for it, so the
 * caller's next mutation and call are queued while it runs.  Its outcome is
 * settled by the handle's next call (compute, read, last_stats, set_source,
 * mk_bfs_sync, or set_deferred(0)), which waits for it and, if it failed,
 * returns its error (MK_ERR_STATE: a loop gave up; the results stay invalid
 * until Static) instead of doing its own work.  A call whose relax loop
 * needs more steps than one launch reserves is recomputed from scratch at
 * settling (exact for the graph as it is then).  The batch buffers must stay
 * valid until the call is settled; a mutation of the graph queued on another
 * stream waits for it on the device.  Static calls are never deferred. */
int mk_bfs_set_deferred(mk_bfs *b, int on);
int mk_bfs_sync(mk_bfs *b);
/* Deferred calls settled since the handle was created, and how many of them
 * were recomputed from scratch at settling (their relax loop needed more
 * steps than one launch reserves). */
int mk_bfs_deferred_stats(mk_bfs *b, uint64_t *settled, uint64_t *recomputed);

/* ------------------------------------------------------ A11 SSSP
 * SSSP(G, Src, N) (include/Algorithms/SSSP.cuh:34-69; impl/SSSPImpl.cuh
 * Static :607-682, Incremental :684-778, Decremental :780-876) on a weighted
 * graph.  Edge weights are read from the graph, which holds them; w_dev is
 * accepted for signature parity and may be NULL.  Weights must be >= 1 for
 * Decremental (a zero-weight cycle can carry a stale parent chain). */
int mk_sssp_create(mk_graph *g, uint32_t source, mk_sssp **out);
int mk_sssp_destroy(mk_sssp *s);
int mk_sssp_static(mk_sssp *s, float *ms);
int mk_sssp_incremental(mk_sssp *s, const uint32_t *src_dev, const uint32_t *dst_dev,
                        const uint32_t *w_dev, uint32_t n, float *ms);
int mk_sssp_decremental(mk_sssp *s, const uint32_t *src_dev, const uint32_t *dst_dev,
                        const uint32_t *w_dev, uint32_t n, float *ms);
/* Shortest paths kept exact across a weight update (no reference
 * counterpart: the reference's UpdateEdges is compiled out and no algorithm
 * follows it).  Call after mk_graph_update_weights with the same batch
 * (w_dev: its new weights).  A tree edge that became heavier invalidates
 * its subtree, which is re-seeded from valid in-neighbours (as
 * Decremental); every batch edge is then relaxed with its weight in the
 * graph (as Incremental), which covers the edges that became lighter.
 * Deferred like Incremental / Decremental. */
int mk_sssp_reweight(mk_sssp *s, const uint32_t *src_dev, const uint32_t *dst_dev,
                     const uint32_t *w_dev, uint32_t n, float *ms);
int mk_sssp_read(mk_sssp *s, uint32_t *dist_host, uint32_t *parent_host);
/* As mk_bfs_set_source. */
int mk_sssp_set_source(mk_sssp *s, uint32_t source);
int mk_sssp_last_stats(mk_sssp *s, uint32_t *iterations, uint64_t *slabs_read,
                       uint64_t *edges_read);
/* As mk_bfs_set_deferred / mk_bfs_sync / mk_bfs_deferred_stats. */
int mk_sssp_set_deferred(mk_sssp *s, int on);
int mk_sssp_sync(mk_sssp *s);
int mk_sssp_deferred_stats(mk_sssp *s, uint64_t *settled, uint64_t *recomputed);

/* Captured update steps (no reference counterpart: the reference launches
 * every kernel from the host; here the update -> incremental-recompute step
 * can be replayed as one hipGraph).  mk_*_capture_step records one step
 * of the handle on its graph, DeleteEdges(del) -> Decremental(del) ->
 * InsertEdges(ins) -> Incremental(ins), into a hipGraph, reading the batch
 * buffers given here (device memory the caller refills between replays; the
 * sizes are fixed).  Nothing runs at capture.  mk_step_launch queues one
 * replay on the graph's stream and returns; replays may be queued back to
 * back.  mk_step_sync waits for the queued replays and reports their
 * outcome: MK_ERR_RANGE if a batch named a vertex >= n (that batch changed
 * nothing), MK_ERR_POOL_FULL if an insert ran out of slabs (rolled back when
 * it was the last replay's, else left partly applied), MK_ERR_STATE if a
 * frontier loop gave up; after any of these the handle needs Static().  A
 * relax loop that ran out of its steps is recomputed from scratch by the
 * sync.  Calls through the API may follow a sync.  One step per graph; the
 * handle cannot be destroyed before its step.  Needs persistent_steps = 1
 * and profiling off at capture. */
typedef struct mk_step mk_step;
int mk_sssp_capture_step(mk_sssp *s, const uint32_t *del_src_dev, const uint32_t *del_dst_dev,
                         uint32_t n_del, const uint32_t *ins_src_dev, const uint32_t *ins_dst_dev,
                         const uint32_t *ins_w_dev, uint32_t n_ins, mk_step **out);
int mk_bfs_capture_step(mk_bfs *b, const uint32_t *del_src_dev, const uint32_t *del_dst_dev,
                        uint32_t n_del, const uint32_t *ins_src_dev, const uint32_t *ins_dst_dev,
                        uint32_t n_ins, mk_step **out);
int mk_step_launch(mk_step *st);
/* Queue copies of the next step's batches (device arrays of the captured
 * sizes; ins_w_dev NULL keeps the weights) into the step's buffers, on the
 * graph's stream: ordered after the replays queued so far and before the
 * next, so replays with different batches can be queued back to back. */
int mk_step_load(mk_step *st, const uint32_t *del_src_dev, const uint32_t *del_dst_dev,
                 const uint32_t *ins_src_dev, const uint32_t *ins_dst_dev,
                 const uint32_t *ins_w_dev);
int mk_step_sync(mk_step *st);
int mk_step_destroy(mk_step *st);

/* ------------------------------------------------------ A12 PageRank
 * PageRank(G, DampingFactor, ErrorMargin, N, VertexOutDegree, MaxIter)
 * (include/Algorithms/PageRank.cuh:32-67; impl/PageRankImpl.cuh:146-251).
 * g is either a graph of reversed edges, as the reference's RevUnweighted
 * policy graph (the caller inserts (dst, src)), with host_outdeg giving the
 * out-degrees (NULL: counted from g); or a graph created with
 * MK_GRAPH_REVERSE, whose in-edge mirror is used and whose live out-degrees
 * are exact (host_outdeg ignored).  Ranks are FP32, reductions FP64; every
 * compute call continues from the current ranks (initially 1/N) and stops
 * when the L1 delta <= error_margin or after max_iter iterations, as the
 * reference loop (:173).  Incremental / Decremental take the batch just
 * inserted / deleted (original direction) and fix the out-degrees first
 * (CheckPresent, :64-87). */
int mk_pagerank_create(mk_graph *g, float damping, float error_margin, uint32_t max_iter,
                       const uint32_t *host_outdeg, mk_pagerank **out);
int mk_pagerank_destroy(mk_pagerank *p);
int mk_pagerank_static(mk_pagerank *p, float *ms);
int mk_pagerank_incremental(mk_pagerank *p, const uint32_t *src_dev, const uint32_t *dst_dev,
                            uint32_t n, float *ms);
int mk_pagerank_decremental(mk_pagerank *p, const uint32_t *src_dev, const uint32_t *dst_dev,
                            uint32_t n, float *ms);
/* Values() (PageRank.cuh:50-55): n FP32 ranks. */
int mk_pagerank_read(mk_pagerank *p, float *values_host);
/* Iterations run and final L1 delta of the last compute call. */
int mk_pagerank_last_stats(mk_pagerank *p, uint32_t *iterations, double *last_delta);

/* ------------------------------------------------------ A13 TC
 * TC(G, N) (include/Algorithms/TC.cuh:28-40; impl/TCImpl.cuh) on an
 * undirected graph stored in both directions.  Static counts every triangle
 * once (the reference's /6, :368); Incremental / Decremental take the batch
 * just inserted / deleted, in both directions, and update the count exactly
 * (the reference combines three counts in closed form, :488-489, 556-557).
 * The batch's changed edges are those the graph's last mutation applied
 * when it is the same batch (same device pointers and size); otherwise
 * every non-loop entry is taken as changed, as the reference assumes.  The
 * reference's Allocator argument (used there for a temporary batch graph,
 * :457) has no counterpart: no temporary graph is built. */
int mk_tc_create(mk_graph *g, mk_tc **out);
int mk_tc_destroy(mk_tc *t);
int mk_tc_static(mk_tc *t, float *ms);
int mk_tc_incremental(mk_tc *t, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                      float *ms);
int mk_tc_decremental(mk_tc *t, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                      float *ms);
/* GetCount() (TC.cuh:30). */
int mk_tc_count(mk_tc *t, uint64_t *count);

/* ------------------------------------------------------ A14 CC
 * Connected components (edges taken as undirected) by concurrent
 * union-find.  The reference has only primitives (include/Util/
 * UnionFind.cuh:24-253; UnionAsync :122-149 links the larger root under the
 * smaller), no caller and no driver.  Labels are canonical: the smallest
 * vertex id of each component.  Incremental unions the inserted batch;
 * Decremental recomputes from the graph (union-find cannot split). */
int mk_cc_create(mk_graph *g, mk_cc **out);
int mk_cc_destroy(mk_cc *c);
int mk_cc_static(mk_cc *c, float *ms);
int mk_cc_incremental(mk_cc *c, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                      float *ms);
int mk_cc_decremental(mk_cc *c, const uint32_t *src_dev, const uint32_t *dst_dev, uint32_t n,
                      float *ms);
int mk_cc_labels(mk_cc *c, uint32_t *labels_host);

/* ------------------------------------------------------ A17 multi-source
 * Up to 64 BFS (sssp = 0) or SSSP (sssp != 0, weighted graph) queries
 * answered by one traversal, for the serving loop's batching.  No reference
 * counterpart (the reference answers one source per BFS/SSSP object,
 * include/Algorithms/BFS.cuh:34-53).  Source slot s is bit s of per-vertex
 * masks; sources may repeat.  run blocks until all k traversals are done;
 * read returns slot s's distances (UINT32_MAX = unreachable), identical to
 * a single-source static BFS / SSSP from that source. */
int mk_msq_create(mk_graph *g, int sssp, mk_msq **out);
int mk_msq_destroy(mk_msq *q);
int mk_msq_run(mk_msq *q, const uint32_t *sources_host, uint32_t k, float *ms);
int mk_msq_read(mk_msq *q, uint32_t slot, uint32_t *dist_host);
int mk_msq_last_stats(mk_msq *q, uint32_t *steps, uint64_t *edges_read);
/* Continuous batching (no reference counterpart): a session admits sources
 * into free slots of a traversal that is already running, and each slot
 * finishes on its own step.
 *  - mk_msq_open starts a session on the graph as it is now (its mutations
 *    must wait until the session's slots have finished: a slot's distances
 *    are those of the graph version it ran on);
 *  - mk_msq_admit puts sources_host[i] into slot slots_host[i] (< 64, each
 *    slot once per session), joining the frontier of the next step (BFS
 *    levels then count from that step);
 *  - mk_msq_launch queues `steps` more steps (no-ops once the frontier is
 *    empty) and returns; mk_msq_ready tells without waiting whether they
 *    have run; mk_msq_poll waits for them and sets *active to the
 *    admitted slots whose traversal still has work and *steps (nullable) to
 *    the steps queued so far: an admitted slot not in *active is final, and
 *    mk_msq_read(q, slot, ...) returns its distances.
 * mk_msq_run ends any session.  A BFS session whose level log fills up
 * (sources admitted at many different steps put a vertex in many levels)
 * returns MK_ERR_STATE from mk_msq_poll; its slots still active are then
 * incomplete and belong in a new session, which gets a log four times
 * larger, up to the 64 entries per vertex a session can need; the slots
 * reported finished stay readable until then. */
int mk_msq_open(mk_msq *q);
int mk_msq_admit(mk_msq *q, const uint32_t *sources_host, const uint32_t *slots_host, uint32_t k);
int mk_msq_launch(mk_msq *q, uint32_t steps);
int mk_msq_poll(mk_msq *q, uint64_t *active, uint32_t *steps);
/* *ready = 1 when the steps launched so far have run (mk_msq_poll will not
 * wait), else 0; never waits. */
int mk_msq_ready(mk_msq *q, int *ready);

/* ------------------------------------------------------ tuning knobs
 * Launch-shape knobs, read when a handle is created (no reference
 * counterpart; the reference hard-codes its shapes, e.g. BFSImpl.cuh:
 * 620-625).  "waves_per_cu" (1..64, default 16): waves the grid-stride
 * frontier, CC and triangle kernels launch per compute unit.  "ms_unroll"
 * (4 or 8, default 8): neighbour distance rows in flight per wave in
 * multi-source SSSP.  "ms_waves_per_cu" (1..64, default 24) and
 * "pr_waves_per_cu" (1..64, default 32): waves per compute unit of the
 * multi-source steps and of the PageRank iteration, each capped at what a
 * compute unit holds of the kernel.
 * "orphan_repair" (0 or 1, default 0): a decremental BFS/SSSP call on a
 * graph with the in-edge mirror first offers each vertex whose tree edge
 * was deleted another parent at the same distance, sparing its subtree
 * (pays on unweighted graphs with small batches; on RMAT-24 batches of
 * 100k+ edges the extra pass costs more than the shorter walk saves).
 * "persistent_steps" (0 or 1, default 1): BFS/SSSP frontier steps run in
 * one cooperative launch with a grid barrier between steps (0: one launch
 * per step, the host polling the frontier counters).  "tail_enter"
 * (0..1024, default 128): frontiers of at most this many vertices (and no
 * hub) are expanded by one block step after step in LDS (0: never).
 * "coop_launch" (0 or 1, default 0): 1 launches the frontier loops
 * cooperatively always; 0 does so only while BFS/SSSP handles exist on more
 * than one graph of a device (their loops could then overlap), and launches
 * them as ordinary kernels, with a grid sized to stay resident, otherwise.
 * "mut_waves_per_cu" (1..256, default 32, read at each mutation call): the
 * waves per compute unit a mutation batch is spread over.
 * "barrier_timeout_ms" (0..600000, default 5000): the longest a frontier
 * loop's grid barrier waits before the call gives up with MK_ERR_STATE.
 * "pull_enter" / "pull_exit" (0..4096, defaults 14 / 24): static BFS over a
 * graph with the in-edge mirror runs a step bottom-up once the frontier
 * holds more than 1/pull_enter of the buckets (0: never) and until it holds
 * fewer than 1/pull_exit of the vertices.  "sssp_delta" (0..2^24, default
 * 0): static SSSP runs near-far with this bucket width (0: label-correcting
 * without deferral); "sssp_far_cap" (default 0: one entry per vertex)
 * bounds each far list.
 * The Python autotuner (meerkat_amd/autotune.py) sets them from
 * measurements. */
int mk_tune_set(const char *name, int64_t value);
int mk_tune_get(const char *name, int64_t *value);

/* ------------------------------------------------------ kernel profiling
 * No reference counterpart (the reference times whole calls with
 * cudaEvent, BFSImpl.cuh:644-661).  When enabled, every instrumented kernel
 * launch is bracketed by HIP events on its own stream; totals (launches,
 * milliseconds, algorithmic bytes) accumulate per kernel name until reset.
 * mk_profile_count snapshots the names; mk_profile_get reads entry i.
 * on: 0 off, 1 every instrumented kernel, 2 only the persistent BFS/SSSP
 * loops (timed by their own device clock readings, nearly free). */
int mk_profile_enable(int on);
int mk_profile_reset(void);
int mk_profile_count(int *count);
int mk_profile_get(int i, const char **name, uint64_t *launches, double *total_ms,
                   double *algo_bytes);
/* Frontier steps of the persistent BFS/SSSP loops while profiling (the
 * first 64 steps of each loop launch): kernel name, step index, frontier
 * size (vertices, hub bucket items) and the step's duration in
 * microseconds from the device wall clock, grid barrier included. */
int mk_profile_step_count(int *count);
int mk_profile_step_get(int i, const char **name, uint32_t *step, uint32_t *small,
                        uint32_t *hub_items, double *us);

/* ------------------------------------------------------ workload generators
 * Host-side, seeded, thread-count independent.  They replace the text edge
 * lists and ".batch.<i>" files the reference drivers read (src/BFS.cu:23-36,
 * 76-91).  mk_gen_rmat / mk_gen_er leave the sorted, de-duplicated,
 * self-loop-free edge list in a per-thread buffer and return its length;
 * mk_gen_fetch copies it out (weights 1 + draw mod 100, as src/SSSP.cu:52). */
int mk_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                uint64_t seed, int undirected, uint64_t *out_edges);
int mk_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint64_t *out_edges);
int mk_gen_fetch(uint32_t *src_host, uint32_t *dst_host, uint32_t *w_host, uint64_t wseed);
/* New edges absent from sorted_keys (src<<32|dst ascending) and distinct. */
int mk_gen_insert_batch(const uint64_t *sorted_keys, uint64_t nkeys, uint32_t n, uint32_t count,
                        uint64_t seed, uint32_t *src_host, uint32_t *dst_host, uint32_t *w_host,
                        uint64_t wseed);
/* Distinct edges sampled without replacement from sorted_keys. */
int mk_gen_delete_batch(const uint64_t *sorted_keys, uint64_t nkeys, uint32_t count, uint64_t seed,
                        uint32_t *src_host, uint32_t *dst_host);

#ifdef __cplusplus
}
#endif

#endif /* MEERKAT_H_ */
