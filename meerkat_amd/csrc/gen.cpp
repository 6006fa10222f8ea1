// Seeded synthetic graph and update-batch generators (host C++, std::thread).
//
// These produce the workloads the reference drivers read from text files
// (src/BFS.cu:23-36 base graph, src/BFS.cu:76-91 batches): RMAT and
// Erdos-Renyi edge lists with duplicates and self-loops dropped, uniform
// integer weights in [1,100] as in src/SSSP.cu:52 (`1 + rand() % 100`), and
// update batches of new edges (rejected if present) or of deletions sampled
// from the current edge set.
//
// Every random draw comes from a counter-based stream (splitmix64 of
// seed/edge-index/draw-index), so the output is identical whatever the thread
// count, and the same seed gives the same graph on every machine.
#include "meerkat.h"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Draw `k` of the stream identified by (seed, id).
inline uint64_t draw(uint64_t seed, uint64_t id, uint64_t k) {
  return mix64(mix64(seed * 0xD1B54A32D192ED03ull + id) + k * 0x9E3779B97F4A7C15ull);
}

unsigned worker_count(uint64_t work) {
  unsigned hw = std::thread::hardware_concurrency();
  if (hw == 0) hw = 4;
  if (hw > 16) hw = 16;  // the GPU host's CPU share
  uint64_t need = work / (1u << 16) + 1;
  return static_cast<unsigned>(std::min<uint64_t>(hw, need));
}

template <typename F> void parallel_for(uint64_t n, F f) {
  unsigned t = worker_count(n);
  if (t <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  uint64_t chunk = (n + t - 1) / t;
  for (unsigned i = 0; i < t; ++i) {
    uint64_t lo = i * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(f, lo, hi);
  }
  for (auto &x : th) x.join();
}

// Sort 64-bit keys (src << 32 | dst) ascending, drop duplicates.  MSD
// partition on the top 16 significant bits, then per-bucket std::sort, so it
// runs on all worker threads.  Returns the unique count.
uint64_t sort_unique(std::vector<uint64_t> &keys, int key_bits) {
  const uint64_t n = keys.size();
  if (n == 0) return 0;
  const int pbits = std::min(16, std::max(1, key_bits));
  const int shift = key_bits - pbits;
  const uint32_t nb = 1u << pbits;
  unsigned t = worker_count(n);
  std::vector<std::vector<uint64_t>> hist(t, std::vector<uint64_t>(nb, 0));
  uint64_t chunk = (n + t - 1) / t;
  {
    std::vector<std::thread> th;
    for (unsigned i = 0; i < t; ++i)
      th.emplace_back([&, i] {
        uint64_t lo = i * chunk, hi = std::min(n, lo + chunk);
        for (uint64_t j = lo; j < hi; ++j) hist[i][keys[j] >> shift]++;
      });
    for (auto &x : th) x.join();
  }
  std::vector<uint64_t> bstart(nb + 1, 0);
  std::vector<std::vector<uint64_t>> pos(t, std::vector<uint64_t>(nb, 0));
  uint64_t acc = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    bstart[b] = acc;
    for (unsigned i = 0; i < t; ++i) {
      pos[i][b] = acc;
      acc += hist[i][b];
    }
  }
  bstart[nb] = acc;
  std::vector<uint64_t> out(n);
  {
    std::vector<std::thread> th;
    for (unsigned i = 0; i < t; ++i)
      th.emplace_back([&, i] {
        uint64_t lo = i * chunk, hi = std::min(n, lo + chunk);
        for (uint64_t j = lo; j < hi; ++j) out[pos[i][keys[j] >> shift]++] = keys[j];
      });
    for (auto &x : th) x.join();
  }
  std::vector<uint64_t> ucount(nb, 0);
  std::atomic<uint32_t> next{0};
  {
    std::vector<std::thread> th;
    for (unsigned i = 0; i < t; ++i)
      th.emplace_back([&] {
        for (uint32_t b; (b = next.fetch_add(1)) < nb;) {
          auto lo = out.begin() + bstart[b], hi = out.begin() + bstart[b + 1];
          std::sort(lo, hi);
          ucount[b] = std::unique(lo, hi) - lo;
        }
      });
    for (auto &x : th) x.join();
  }
  uint64_t w = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    if (w != bstart[b])
      std::memmove(&out[w], &out[bstart[b]], ucount[b] * sizeof(uint64_t));
    w += ucount[b];
  }
  out.resize(w);
  keys.swap(out);
  return w;
}

int bits_for(uint64_t n) {
  int b = 1;
  while ((1ull << b) < n) ++b;
  return b;
}

// Convert unique compact keys (src << vbits | dst) to output arrays with
// optional weights in [1,100] that depend only on (seed, src, dst).
void emit(const std::vector<uint64_t> &keys, int vbits, uint32_t *src, uint32_t *dst,
          uint32_t *w, uint64_t wseed) {
  const uint64_t mask = (1ull << vbits) - 1;
  parallel_for(keys.size(), [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      uint32_t s = static_cast<uint32_t>(keys[i] >> vbits);
      uint32_t d = static_cast<uint32_t>(keys[i] & mask);
      src[i] = s;
      dst[i] = d;
      if (w) w[i] = 1 + static_cast<uint32_t>(draw(wseed, (uint64_t(s) << 32) | d, 0) % 100);
    }
  });
}

thread_local std::vector<uint64_t> g_last_keys;  // result of the most recent generate call
thread_local int g_last_vbits = 0;

}  // namespace

extern "C" {

// Generate an RMAT edge list (Chakrabarti et al. recursive matrix).  Edge i
// picks one quadrant per level, most significant bit first, from draw
// (seed, i, level): a -> (0,0), b -> (0,1), c -> (1,0), d -> (1,1).  No vertex
// permutation, so vertex 0 is the hub.  Self-loops and duplicates are dropped;
// `undirected` adds the reverse of every edge before deduplication.  The
// result is held in a per-thread buffer; read it with mk_gen_fetch.
int mk_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                uint64_t seed, int undirected, uint64_t *out_edges) {
  if (scale < 1 || scale > 31 || !out_edges || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return MK_ERR_BAD_ARG;
  const uint64_t m = (uint64_t(edge_factor) << scale);
  const uint64_t ta = static_cast<uint64_t>(a * 4294967296.0);
  const uint64_t tb = static_cast<uint64_t>((a + b) * 4294967296.0);
  const uint64_t tc = static_cast<uint64_t>(std::min(1.0, a + b + c) * 4294967296.0);
  const int vbits = static_cast<int>(scale);
  // (n-1, n-1) is a self-loop, never an edge: it marks dropped draws and sorts last
  const uint64_t sentinel = (((1ull << vbits) - 1) << vbits) | ((1ull << vbits) - 1);
  std::vector<uint64_t> keys(undirected ? 2 * m : m);
  parallel_for(m, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t s = 0, d = 0;
      uint64_t r = 0;
      for (uint32_t l = 0; l < scale; ++l) {
        if ((l & 1) == 0) r = draw(seed, i, l >> 1);
        uint64_t u = (l & 1) ? (r >> 32) : (r & 0xFFFFFFFFull);
        uint32_t sb = u >= tb ? 1 : 0;           // quadrants c, d
        uint32_t db = (u >= ta && u < tb) || u >= tc ? 1 : 0;  // quadrants b, d
        s = (s << 1) | sb;
        d = (d << 1) | db;
      }
      if (s == d) {
        keys[undirected ? 2 * i : i] = sentinel;
        if (undirected) keys[2 * i + 1] = sentinel;
        continue;
      }
      if (undirected) {
        keys[2 * i] = (s << vbits) | d;
        keys[2 * i + 1] = (d << vbits) | s;
      } else {
        keys[i] = (s << vbits) | d;
      }
    }
  });
  uint64_t u = sort_unique(keys, 2 * vbits);
  if (u && keys[u - 1] == sentinel) keys.pop_back();
  g_last_keys.swap(keys);
  g_last_vbits = vbits;
  *out_edges = g_last_keys.size();
  return MK_OK;
}

// Directed Erdos-Renyi G(n, m): m uniform (u, v) pairs with u != v from
// draw(seed, i, *), duplicates dropped.
int mk_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint64_t *out_edges) {
  if (n < 2 || !out_edges) return MK_ERR_BAD_ARG;
  const int vbits = bits_for(n);
  std::vector<uint64_t> keys(m);
  parallel_for(m, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t k = 0;
      for (uint64_t t = 0;; ++t) {
        uint64_t r = draw(seed, i, t);
        uint64_t s = (uint64_t(static_cast<uint32_t>(r)) * n) >> 32;
        uint64_t d = ((r >> 32) * n) >> 32;
        if (s != d) {
          k = (s << vbits) | d;
          break;
        }
      }
      keys[i] = k;
    }
  });
  sort_unique(keys, 2 * vbits);
  g_last_keys.swap(keys);
  g_last_vbits = vbits;
  *out_edges = g_last_keys.size();
  return MK_OK;
}

// Copy the last generated edge list (sorted by (src, dst)) into caller
// buffers of at least `*out_edges` entries.  `w_host` may be NULL; otherwise
// it receives weights 1 + (draw(wseed, src<<32|dst, 0) mod 100).
int mk_gen_fetch(uint32_t *src_host, uint32_t *dst_host, uint32_t *w_host, uint64_t wseed) {
  if (!src_host || !dst_host) return MK_ERR_BAD_ARG;
  emit(g_last_keys, g_last_vbits, src_host, dst_host, w_host, wseed);
  return MK_OK;
}

// A batch of `count` edges that are absent from the current edge set and from
// each other: uniform (u, v), u != v, over n vertices, rejection-sampled from
// draw(seed, attempt, 0).  `sorted_keys` holds the current edges as
// (src << 32 | dst), ascending.  Weights as in mk_gen_fetch when w_host given.
int mk_gen_insert_batch(const uint64_t *sorted_keys, uint64_t nkeys, uint32_t n, uint32_t count,
                        uint64_t seed, uint32_t *src_host, uint32_t *dst_host, uint32_t *w_host,
                        uint64_t wseed) {
  if (n < 2 || (!src_host && count) || (!dst_host && count)) return MK_ERR_BAD_ARG;
  std::unordered_set<uint64_t> taken;
  taken.reserve(count * 2 + 16);
  uint64_t attempt = 0;
  const uint64_t limit = 64ull * count + 1024;
  for (uint32_t i = 0; i < count;) {
    if (attempt > limit) return MK_ERR_RANGE;  // graph too dense for rejection
    uint64_t r = draw(seed, attempt++, 0);
    uint32_t s = static_cast<uint32_t>((uint64_t(static_cast<uint32_t>(r)) * n) >> 32);
    uint32_t d = static_cast<uint32_t>(((r >> 32) * n) >> 32);
    if (s == d) continue;
    uint64_t k = (uint64_t(s) << 32) | d;
    if (sorted_keys && std::binary_search(sorted_keys, sorted_keys + nkeys, k)) continue;
    if (!taken.insert(k).second) continue;
    src_host[i] = s;
    dst_host[i] = d;
    if (w_host) w_host[i] = 1 + static_cast<uint32_t>(draw(wseed, k, 0) % 100);
    ++i;
  }
  return MK_OK;
}

// A batch of `count` distinct edges sampled uniformly without replacement
// from the current edge set (`sorted_keys`, ascending src<<32|dst).
int mk_gen_delete_batch(const uint64_t *sorted_keys, uint64_t nkeys, uint32_t count, uint64_t seed,
                        uint32_t *src_host, uint32_t *dst_host) {
  if (count > nkeys || (count && (!sorted_keys || !src_host || !dst_host))) return MK_ERR_BAD_ARG;
  std::unordered_set<uint64_t> chosen;
  chosen.reserve(count * 2 + 16);
  uint64_t attempt = 0;
  for (uint32_t i = 0; i < count;) {
    uint64_t idx = draw(seed, attempt++, 1) % nkeys;
    if (!chosen.insert(idx).second) continue;
    src_host[i] = static_cast<uint32_t>(sorted_keys[idx] >> 32);
    dst_host[i] = static_cast<uint32_t>(sorted_keys[idx]);
    ++i;
  }
  return MK_OK;
}

}  // extern "C"
