/*
 * oracle.c - CPU restatement of the reference's graph algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * (tests/, __graft_entry__.smoke) and the timed CPU baseline (bench.py's
 * cpu_baseline leg).  Nothing in the product (meerkat_amd/, include/) links,
 * loads or calls it.
 *
 * Every function is a from-scratch static recompute on an edge list / CSR,
 * restating the semantics of the reference (paths relative to its tree):
 *   orc_csr_build   set semantics of the slab hash (duplicates stored once,
 *                   self-loops skipped: include/Graph/device/
 *                   InsertEdgeKernel.cuh:25,54; test/GraphTest.cu:506-507
 *                   std::map<uint32_t, std::set<uint32_t>> oracle); for
 *                   weighted duplicates the first occurrence is kept.
 *   orc_bfs         levels from the source, source = 0
 *                   (include/Algorithms/impl/BFSImpl.cuh:498-588); sequential
 *                   FIFO queue.  orc_bfs_par: level-synchronous, OpenMP.
 *   orc_dijkstra    distances with uint32 weights, INF = 0xFFFFFFFF
 *                   (impl/SSSPImpl.cuh:286,571; definitions.cuh:3-6);
 *                   binary heap, 64-bit accumulation.  orc_sssp_par:
 *                   frontier label-correcting relaxation (the reference's
 *                   SSSP method, SSSPImpl.cuh:490-605) on OpenMP threads, the
 *                   multithreaded CPU baseline; equal distances.
 *   orc_pagerank    PR'[u] = (1-d)/N + d*sum_{v in in(u)} PR[v]/outdeg[v]
 *                            + d*sum_{outdeg(x)=0} PR[x]/N,
 *                   delta = sum |PR' - PR|, stop when delta <= margin or
 *                   after max_iter (impl/PageRankImpl.cuh:3-62,89-144,
 *                   146-216), in FP64.
 *   orc_tc          triangles of an undirected (symmetric) graph, each
 *                   counted once (impl/TCImpl.cuh:149-214 counts every
 *                   directed wedge and divides by 6, :368).
 *   orc_cc          connected components with label = minimum vertex id,
 *                   union by linking the larger root under the smaller
 *                   (include/Util/UnionFind.cuh:122-149).
 *
 * Pinning (tests/test_oracle.py): every function is checked against
 * library routines (scipy.sparse.csgraph BFS/Dijkstra/connected_components,
 * numpy dense algebra for PageRank and trace(A^3)/6), closed forms and brute
 * force on tiny graphs, and the golden fixtures in tests/golden/.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF 0xFFFFFFFFu

/* ------------------------------------------------------------------ CSR */
typedef struct {
  uint32_t col;
  uint32_t w;
  uint64_t ord;
} orc_ent;

static int ent_cmp(const void *a, const void *b) {
  const orc_ent *x = (const orc_ent *)a, *y = (const orc_ent *)b;
  if (x->col != y->col) return x->col < y->col ? -1 : 1;
  if (x->ord != y->ord) return x->ord < y->ord ? -1 : 1;
  return 0;
}

/* Build a CSR with sorted, unique columns per row.  row_off has n+1 entries;
 * col/wcol need room for m entries (wcol may be NULL, w may be NULL).
 * Returns the number of stored edges, or -1 on a vertex id >= n. */
int64_t orc_csr_build(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                      const uint32_t *w, uint64_t *row_off, uint32_t *col, uint32_t *wcol) {
  uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  if (!cnt) return -1;
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] >= n || dst[i] >= n) {
      free(cnt);
      return -1;
    }
    if (src[i] != dst[i]) cnt[src[i] + 1]++;
  }
  for (uint32_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
  orc_ent *e = (orc_ent *)malloc(sizeof(orc_ent) * (cnt[n] ? cnt[n] : 1));
  uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  if (!e || !pos) {
    free(cnt);
    free(e);
    free(pos);
    return -1;
  }
  memcpy(pos, cnt, sizeof(uint64_t) * ((size_t)n + 1));
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] == dst[i]) continue;
    orc_ent *t = &e[pos[src[i]]++];
    t->col = dst[i];
    t->w = w ? w[i] : 0;
    t->ord = i;
  }
  /* rows are sorted (by column, then input order) and counted in parallel,
   * then written at their offsets: the same result as one sequential pass */
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; ++v) {
    const uint64_t lo = cnt[v], hi = cnt[v + 1];
    if (hi - lo > 1) qsort(e + lo, hi - lo, sizeof(orc_ent), ent_cmp);
    uint64_t u = 0;
    for (uint64_t j = lo; j < hi; ++j)
      if (j == lo || e[j].col != e[j - 1].col) ++u; /* keep first */
    pos[v] = u;
  }
  row_off[0] = 0;
  for (uint32_t v = 0; v < n; ++v) row_off[v + 1] = row_off[v] + pos[v];
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; ++v) {
    uint64_t out = row_off[v];
    for (uint64_t j = cnt[v]; j < cnt[v + 1]; ++j) {
      if (j > cnt[v] && e[j].col == e[j - 1].col) continue;
      col[out] = e[j].col;
      if (wcol) wcol[out] = e[j].w;
      ++out;
    }
  }
  const uint64_t out = row_off[n];
  free(cnt);
  free(e);
  free(pos);
  return (int64_t)out;
}

/* Reverse (transpose) of a CSR: in-adjacency.  rrow_off n+1, rcol m. */
void orc_csr_transpose(uint32_t n, const uint64_t *row_off, const uint32_t *col,
                       uint64_t *rrow_off, uint32_t *rcol) {
  memset(rrow_off, 0, sizeof(uint64_t) * ((size_t)n + 1));
  const uint64_t m = row_off[n];
  for (uint64_t j = 0; j < m; ++j) rrow_off[col[j] + 1]++;
  for (uint32_t v = 0; v < n; ++v) rrow_off[v + 1] += rrow_off[v];
  uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  memcpy(pos, rrow_off, sizeof(uint64_t) * ((size_t)n + 1));
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t j = row_off[u]; j < row_off[u + 1]; ++j) rcol[pos[col[j]]++] = u;
  free(pos);
}

/* ------------------------------------------------------------------ BFS */
void orc_bfs(uint32_t n, const uint64_t *row_off, const uint32_t *col, uint32_t src,
             uint32_t *dist) {
  for (uint32_t v = 0; v < n; ++v) dist[v] = ORC_INF;
  uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t head = 0, tail = 0;
  dist[src] = 0;
  q[tail++] = src;
  while (head < tail) {
    uint32_t u = q[head++];
    for (uint64_t j = row_off[u]; j < row_off[u + 1]; ++j) {
      uint32_t v = col[j];
      if (dist[v] == ORC_INF) {
        dist[v] = dist[u] + 1;
        q[tail++] = v;
      }
    }
  }
  free(q);
}

/* Level-synchronous parallel BFS (top-down, OpenMP).  Same result as
 * orc_bfs.  Returns the number of levels. */
uint32_t orc_bfs_par(uint32_t n, const uint64_t *row_off, const uint32_t *col, uint32_t src,
                     uint32_t *dist, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < (int64_t)n; ++v) dist[v] = ORC_INF;
  uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *nxt = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t ncur = 1, nnxt = 0;
  cur[0] = src;
  dist[src] = 0;
  uint32_t level = 0;
  while (ncur) {
    nnxt = 0;
#pragma omp parallel
    {
      uint32_t local[256];
      int nl = 0;
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < (int64_t)ncur; ++i) {
        uint32_t u = cur[i];
        for (uint64_t j = row_off[u]; j < row_off[u + 1]; ++j) {
          uint32_t v = col[j];
          if (__atomic_load_n(&dist[v], __ATOMIC_RELAXED) != ORC_INF) continue;
          uint32_t expect = ORC_INF;
          if (__atomic_compare_exchange_n(&dist[v], &expect, level + 1, 0, __ATOMIC_RELAXED,
                                          __ATOMIC_RELAXED)) {
            local[nl++] = v;
            if (nl == 256) {
              uint64_t at = __atomic_fetch_add(&nnxt, 256, __ATOMIC_RELAXED);
              memcpy(nxt + at, local, sizeof(local));
              nl = 0;
            }
          }
        }
      }
      if (nl) {
        uint64_t at = __atomic_fetch_add(&nnxt, (uint64_t)nl, __ATOMIC_RELAXED);
        memcpy(nxt + at, local, sizeof(uint32_t) * nl);
      }
    }
    uint32_t *t = cur;
    cur = nxt;
    nxt = t;
    ncur = nnxt;
    ++level;
  }
  free(cur);
  free(nxt);
  return level;
}

/* ------------------------------------------------------------- Dijkstra */
typedef struct {
  uint64_t d;
  uint32_t v;
} orc_hn;

static void heap_push(orc_hn *h, uint64_t *sz, orc_hn x) {
  uint64_t i = (*sz)++;
  while (i) {
    uint64_t p = (i - 1) / 2;
    if (h[p].d <= x.d) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = x;
}
static orc_hn heap_pop(orc_hn *h, uint64_t *sz) {
  orc_hn top = h[0];
  orc_hn x = h[--(*sz)];
  uint64_t i = 0, n = *sz;
  for (;;) {
    uint64_t c = 2 * i + 1;
    if (c >= n) break;
    if (c + 1 < n && h[c + 1].d < h[c].d) ++c;
    if (h[c].d >= x.d) break;
    h[i] = h[c];
    i = c;
  }
  if (n) h[i] = x;
  return top;
}

/* Returns 0, or -1 if some distance does not fit below 0xFFFFFFFF. */
int orc_dijkstra(uint32_t n, const uint64_t *row_off, const uint32_t *col, const uint32_t *w,
                 uint32_t src, uint32_t *dist) {
  uint64_t *d = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  for (uint32_t v = 0; v < n; ++v) d[v] = UINT64_MAX;
  orc_hn *h = (orc_hn *)malloc(sizeof(orc_hn) * (row_off[n] + 2));
  uint64_t sz = 0;
  d[src] = 0;
  orc_hn s0 = {0, src};
  heap_push(h, &sz, s0);
  while (sz) {
    orc_hn t = heap_pop(h, &sz);
    if (t.d != d[t.v]) continue;
    for (uint64_t j = row_off[t.v]; j < row_off[t.v + 1]; ++j) {
      uint64_t nd = t.d + w[j];
      if (nd < d[col[j]]) {
        d[col[j]] = nd;
        orc_hn x = {nd, col[j]};
        heap_push(h, &sz, x);
      }
    }
  }
  int rc = 0;
  for (uint32_t v = 0; v < n; ++v) {
    if (d[v] == UINT64_MAX) dist[v] = ORC_INF;
    else if (d[v] >= ORC_INF) {
      dist[v] = ORC_INF;
      rc = -1;
    } else dist[v] = (uint32_t)d[v];
  }
  free(d);
  free(h);
  return rc;
}

/* Parallel label-correcting SSSP: every step relaxes the out-edges of the
 * vertices whose distance dropped in the previous step, with an atomic
 * compare-and-swap minimum; a per-vertex stamp keeps each vertex once per
 * frontier.  Returns the number of steps, or -1 when a distance would
 * exceed uint32 (INF excluded). */
int64_t orc_sssp_par(uint32_t n, const uint64_t *row_off, const uint32_t *col, const uint32_t *w,
                     uint32_t src, uint32_t *dist, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  uint32_t *stamp = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *nxt = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!stamp || !cur || !nxt) {
    free(stamp);
    free(cur);
    free(nxt);
    return -1;
  }
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < (int64_t)n; ++v) dist[v] = ORC_INF;
  dist[src] = 0;
  cur[0] = src;
  uint64_t ncur = 1, nnxt = 0;
  int overflow = 0;
  int64_t steps = 0;
  while (ncur) {
    nnxt = 0;
    const uint32_t tag = (uint32_t)steps + 1u;
#pragma omp parallel
    {
      uint32_t local[256];
      int nl = 0;
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < (int64_t)ncur; ++i) {
        const uint32_t u = cur[i];
        const uint32_t du = __atomic_load_n(&dist[u], __ATOMIC_RELAXED);
        for (uint64_t j = row_off[u]; j < row_off[u + 1]; ++j) {
          const uint32_t v = col[j];
          const uint64_t nd = (uint64_t)du + w[j];
          if (nd >= ORC_INF) {
            __atomic_store_n(&overflow, 1, __ATOMIC_RELAXED);
            continue;
          }
          uint32_t old = __atomic_load_n(&dist[v], __ATOMIC_RELAXED);
          while ((uint32_t)nd < old) {
            if (__atomic_compare_exchange_n(&dist[v], &old, (uint32_t)nd, 1, __ATOMIC_RELAXED,
                                            __ATOMIC_RELAXED)) {
              if (__atomic_exchange_n(&stamp[v], tag, __ATOMIC_RELAXED) != tag) {
                local[nl++] = v;
                if (nl == 256) {
                  uint64_t at = __atomic_fetch_add(&nnxt, 256, __ATOMIC_RELAXED);
                  memcpy(nxt + at, local, sizeof(local));
                  nl = 0;
                }
              }
              break;
            }
          }
        }
      }
      if (nl) {
        uint64_t at = __atomic_fetch_add(&nnxt, (uint64_t)nl, __ATOMIC_RELAXED);
        memcpy(nxt + at, local, sizeof(uint32_t) * nl);
      }
    }
    uint32_t *t = cur;
    cur = nxt;
    nxt = t;
    ncur = nnxt;
    ++steps;
  }
  free(stamp);
  free(cur);
  free(nxt);
  return overflow ? -1 : steps;
}

/* ------------------------------------------------------------- PageRank */
/* in_off/in_col: in-adjacency (the reference runs on the reverse graph,
 * include/Algorithms/PageRank.cuh:11).  pr holds the start vector on entry
 * (the reference initialises 1/N, PageRank.cuh:42-43, and warm-starts
 * incremental runs) and the result on exit.  Returns iterations run. */
uint32_t orc_pagerank(uint32_t n, const uint64_t *in_off, const uint32_t *in_col,
                      const uint32_t *outdeg, double d, double margin, uint32_t max_iter,
                      double *pr, double *last_delta, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  double *contrib = (double *)malloc(sizeof(double) * (n ? n : 1));
  double *next = (double *)malloc(sizeof(double) * (n ? n : 1));
  double delta = 1.0 + margin;
  uint32_t it = 0;
  while (delta > margin && it < max_iter) {
    ++it;
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
      contrib[v] = outdeg[v] ? pr[v] / (double)outdeg[v] : 0.0;
      if (outdeg[v] == 0) acc += pr[v] / (double)n;
    }
    double dl = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : dl)
    for (int64_t u = 0; u < (int64_t)n; ++u) {
      double s = 0.0;
      for (uint64_t j = in_off[u]; j < in_off[u + 1]; ++j) s += contrib[in_col[j]];
      double x = (1.0 - d) / (double)n + d * s + d * acc;
      dl += x > pr[u] ? x - pr[u] : pr[u] - x;
      next[u] = x;
    }
    memcpy(pr, next, sizeof(double) * n);
    delta = dl;
  }
  if (last_delta) *last_delta = delta;
  free(contrib);
  free(next);
  return it;
}

/* ------------------------------------------------------------------- TC */
/* row_off/col: symmetric CSR with sorted unique rows.  Each triangle is
 * counted once: orient u -> v when (deg u, u) < (deg v, v), then intersect
 * the oriented out-lists. */
uint64_t orc_tc(uint32_t n, const uint64_t *row_off, const uint32_t *col, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  const uint64_t m = row_off[n];
  uint64_t *ooff = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  uint32_t *ocol = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
#define ORC_RANK_LT(a, b)                                                          \
  ((row_off[(a) + 1] - row_off[a]) < (row_off[(b) + 1] - row_off[b]) ||            \
   ((row_off[(a) + 1] - row_off[a]) == (row_off[(b) + 1] - row_off[b]) && (a) < (b)))
  ooff[0] = 0;
  for (uint32_t u = 0; u < n; ++u) {
    uint64_t k = ooff[u];
    for (uint64_t j = row_off[u]; j < row_off[u + 1]; ++j)
      if (ORC_RANK_LT(u, col[j])) ocol[k++] = col[j];
    ooff[u + 1] = k;
  }
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (int64_t u = 0; u < (int64_t)n; ++u) {
    for (uint64_t j = ooff[u]; j < ooff[u + 1]; ++j) {
      uint32_t v = ocol[j];
      uint64_t a = ooff[u], ae = ooff[u + 1], b = ooff[v], be = ooff[v + 1];
      while (a < ae && b < be) {
        if (ocol[a] < ocol[b]) ++a;
        else if (ocol[a] > ocol[b]) ++b;
        else {
          ++total;
          ++a;
          ++b;
        }
      }
    }
  }
#undef ORC_RANK_LT
  free(ooff);
  free(ocol);
  return total;
}

/* ------------------------------------------------------------------- CC */
static uint32_t uf_find(uint32_t *p, uint32_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

/* Components of the graph with edges taken as undirected.  labels[v] =
 * smallest vertex id in v's component.  Returns the component count. */
uint32_t orc_cc(uint32_t n, const uint64_t *row_off, const uint32_t *col, uint32_t *labels) {
  for (uint32_t v = 0; v < n; ++v) labels[v] = v;
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t j = row_off[u]; j < row_off[u + 1]; ++j) {
      uint32_t a = uf_find(labels, u), b = uf_find(labels, col[j]);
      if (a == b) continue;
      if (a < b) labels[b] = a; /* larger root under the smaller */
      else labels[a] = b;
    }
  uint32_t comps = 0;
  for (uint32_t v = 0; v < n; ++v) {
    labels[v] = uf_find(labels, v);
    if (labels[v] == v) ++comps;
  }
  return comps;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
