"""Host copy of a graph's edge set under a sequence of update batches, for
checking GPU results at sizes where rebuilding the edge set with Python sets
would be too slow (RMAT-22/24).

The initial edges are sorted unique keys (src << 32 | dst) with weights.
Deletions mark initial keys removed or drop added ones; insertions follow
the slab hash's set semantics (include/Graph/device/InsertEdgeKernel.cuh:
22-26): self-loops skipped, a key already present keeps its weight, the
first occurrence of a key within a batch wins.  ``csr()`` hands the edge
set to the oracle."""
import numpy as np

import oracle
from meerkat_amd import gen


class HostEdges:
    def __init__(self, n, src, dst, w=None):
        self.n = n
        self.keys = gen.keys_of(src, dst)
        assert np.all(self.keys[1:] > self.keys[:-1]), "initial edges must be sorted and unique"
        self.w = None if w is None else np.array(w, np.uint32)  # a copy: weights change
        self.removed = np.zeros(len(self.keys), bool)
        self.added = {}  # key -> weight, edges not in the initial set

    def _lookup(self, k):
        pos = np.searchsorted(self.keys, k)
        pos = np.minimum(pos, len(self.keys) - 1)
        return pos, self.keys[pos] == k

    def delete(self, src, dst):
        k = gen.keys_of(np.asarray(src, np.uint32), np.asarray(dst, np.uint32))
        pos, hit = self._lookup(k)
        for key, p, h in zip(k.tolist(), pos.tolist(), hit.tolist()):
            if key in self.added:
                del self.added[key]
            elif h:
                self.removed[p] = True

    def insert(self, src, dst, w=None):
        src, dst = np.asarray(src, np.uint32), np.asarray(dst, np.uint32)
        k = gen.keys_of(src, dst)
        pos, hit = self._lookup(k)
        ww = np.zeros(len(k), np.uint32) if w is None else np.asarray(w, np.uint32)
        for key, p, h, wt, a, b in zip(k.tolist(), pos.tolist(), hit.tolist(), ww.tolist(),
                                       src.tolist(), dst.tolist()):
            if a == b or key in self.added or (h and not self.removed[p]):
                continue
            self.added[key] = wt

    def update_weights(self, src, dst, w):
        """Present edges take the new weight (a batch naming an edge twice
        must give it one weight); absent edges are not inserted."""
        src, dst = np.asarray(src, np.uint32), np.asarray(dst, np.uint32)
        k = gen.keys_of(src, dst)
        pos, hit = self._lookup(k)
        for key, p, h, wt in zip(k.tolist(), pos.tolist(), hit.tolist(),
                                 np.asarray(w, np.uint32).tolist()):
            if key in self.added:
                self.added[key] = wt
            elif h and not self.removed[p]:
                self.w[p] = wt

    def csr(self):
        keep = ~self.removed
        ak = np.fromiter(self.added.keys(), np.uint64, len(self.added))
        s, d = gen.split_keys(np.concatenate([self.keys[keep], ak]))
        w = None
        if self.w is not None:
            aw = np.fromiter(self.added.values(), np.uint32, len(self.added))
            w = np.concatenate([self.w[keep], aw])
        return oracle.csr(self.n, s, d, w)
