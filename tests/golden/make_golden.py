"""Write the golden fixtures in tests/golden/*.npz.

The reference ships no data (its GraphTest fixture `simple` is not in the
tree) and cannot be built here (CUDA-only, SlabHash submodule absent), so the
golden vectors come from independent library routines, not from this
project's code: graphs are drawn with numpy's PCG64 generator and the
expected answers are computed with scipy.sparse.csgraph (BFS levels,
Dijkstra, weak components) and dense numpy linear algebra (PageRank fixed
point, trace(A^3)/6 triangles).  Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0xFFFFFFFF


def dedup(n, src, dst, w=None):
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = None if w is None else w[keep]
    key = src.astype(np.int64) * n + dst
    _, first = np.unique(key, return_index=True)  # first occurrence per edge
    first.sort()
    return src[first], dst[first], (None if w is None else w[first])


def adjacency(n, src, dst, vals=None):
    v = np.ones(len(src)) if vals is None else vals.astype(np.float64)
    return sp.csr_matrix((v, (src, dst)), shape=(n, n))


def levels(n, src, dst, source):
    d = cg.shortest_path(adjacency(n, src, dst), unweighted=True, indices=source)
    return np.where(np.isinf(d), INF, d).astype(np.uint32)


def dijkstra(n, src, dst, w, source):
    d = cg.dijkstra(adjacency(n, src, dst, w), indices=source)
    return np.where(np.isinf(d), INF, d).astype(np.uint32)


def pagerank_fixed_point(n, src, dst, d=0.85):
    """Solve x = (1-d)/n + d*M x + d*(dangling . x)/n exactly."""
    outdeg = np.bincount(src, minlength=n).astype(np.float64)
    M = np.zeros((n, n))
    for u, v in zip(src, dst):
        M[v, u] += 1.0 / outdeg[u]
    dangling = (outdeg == 0).astype(np.float64)
    A = np.eye(n) - d * M - d * np.outer(np.ones(n), dangling) / n
    return np.linalg.solve(A, np.full(n, (1.0 - d) / n))


def triangles(n, src, dst):
    A = adjacency(n, src, dst).toarray()
    A = ((A + A.T) > 0).astype(np.float64)
    np.fill_diagonal(A, 0)
    return int(round(np.trace(A @ A @ A) / 6))


def components(n, src, dst):
    k, lab = cg.connected_components(adjacency(n, src, dst), directed=True, connection="weak")
    canon = np.full(k, INF, dtype=np.int64)
    for v in range(n):
        canon[lab[v]] = min(canon[lab[v]], v)
    return canon[lab].astype(np.uint32), int(k)


def make(name, n, m, seed, source=0, undirected=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    src = rng.integers(0, n, m, dtype=np.int64).astype(np.uint32)
    dst = rng.integers(0, n, m, dtype=np.int64).astype(np.uint32)
    w = rng.integers(1, 101, m, dtype=np.int64).astype(np.uint32)
    if undirected:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
        w = np.concatenate([w, w])
    s, d, ww = dedup(n, src, dst, w)
    out = dict(n=np.uint32(n), source=np.uint32(source), raw_src=src, raw_dst=dst, raw_w=w,
               src=s, dst=d, w=ww)
    out["bfs"] = levels(n, s, d, source)
    out["sssp"] = dijkstra(n, s, d, ww, source)
    out["pagerank"] = pagerank_fixed_point(n, s, d)
    out["outdeg"] = np.bincount(s, minlength=n).astype(np.uint32)
    lab, k = components(n, s, d)
    out["cc"], out["cc_count"] = lab, np.uint32(k)
    out["tc"] = np.uint64(triangles(n, s, d))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(name, n, len(s), "reached", int((out["bfs"] != INF).sum()), "tc", int(out["tc"]),
          "cc", k)


if __name__ == "__main__":
    make("sparse_directed", 300, 700, seed=11)
    make("dense_directed", 120, 2500, seed=12, source=5)
    make("undirected", 200, 900, seed=13, undirected=True)
    make("components", 400, 260, seed=14, source=3, undirected=True)
