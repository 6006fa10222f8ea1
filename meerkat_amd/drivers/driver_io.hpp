// driver_io.hpp - file formats and device buffers shared by the driver
// programs (the command-line counterparts of the reference's src/BFS.cu,
// src/SSSP.cu, src/PageRank.cu and src/TC.cu).
//
// Formats, as the reference drivers read them:
//  - base graph: text "V E" then E lines "u v" (src/BFS.cu:27-36);
//  - update batch i: file "<prefix>.batch.<i>", text "n" then n lines
//    "u v" (src/BFS.cu:76-91);
//  - output: one CSV line "<graph file>,<static ms>,<batch 0 ms>,..."
//    (src/BFS.cu:12,70,126,130).
// An optional "--dump=<path>" argument (anywhere after the program name)
// writes the final result as raw little-endian binary for checking.
#ifndef MEERKAT_DRIVER_IO_HPP_
#define MEERKAT_DRIVER_IO_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace mkdrv {

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// A device copy of a host uint32 array.
class DeviceU32 {
 public:
  DeviceU32() = default;
  explicit DeviceU32(const std::vector<uint32_t> &h) : n_(h.size()) {
    if (n_) {
      hip_check(hipMalloc(&p_, n_ * 4), "hipMalloc");
      hip_check(hipMemcpy(p_, h.data(), n_ * 4, hipMemcpyHostToDevice), "hipMemcpy");
    }
  }
  ~DeviceU32() {
    if (p_) (void)hipFree(p_);
  }
  DeviceU32(const DeviceU32 &) = delete;
  DeviceU32 &operator=(const DeviceU32 &) = delete;
  uint32_t *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  uint32_t *p_ = nullptr;
  size_t n_ = 0;
};

struct EdgeList {
  uint32_t n = 0;
  std::vector<uint32_t> src, dst;
};

inline EdgeList read_graph(const std::string &path) {
  std::FILE *f = std::fopen(path.c_str(), "r");
  if (!f) throw std::runtime_error("cannot open " + path);
  EdgeList g;
  unsigned long long m = 0;
  if (std::fscanf(f, "%u%llu", &g.n, &m) != 2) {
    std::fclose(f);
    throw std::runtime_error("bad header in " + path);
  }
  g.src.resize(m);
  g.dst.resize(m);
  for (unsigned long long i = 0; i < m; ++i) {
    if (std::fscanf(f, "%u%u", &g.src[i], &g.dst[i]) != 2) {
      std::fclose(f);
      throw std::runtime_error("short edge list in " + path);
    }
  }
  std::fclose(f);
  return g;
}

inline EdgeList read_batch(const std::string &prefix, int i) {
  const std::string path = prefix + ".batch." + std::to_string(i);
  std::FILE *f = std::fopen(path.c_str(), "r");
  if (!f) throw std::runtime_error("cannot open " + path);
  EdgeList b;
  unsigned m = 0;
  if (std::fscanf(f, "%u", &m) != 1) {
    std::fclose(f);
    throw std::runtime_error("bad header in " + path);
  }
  b.src.resize(m);
  b.dst.resize(m);
  for (unsigned k = 0; k < m; ++k) {
    if (std::fscanf(f, "%u%u", &b.src[k], &b.dst[k]) != 2) {
      std::fclose(f);
      throw std::runtime_error("short batch " + path);
    }
  }
  std::fclose(f);
  return b;
}

// Positional arguments (flags removed) and the --dump path, if any.
struct Args {
  std::vector<std::string> pos;
  std::string dump;
};

inline Args parse(int argc, char **argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    if (std::strncmp(argv[i], "--dump=", 7) == 0)
      a.dump = argv[i] + 7;
    else
      a.pos.emplace_back(argv[i]);
  }
  return a;
}

template <typename T>
void dump(const std::string &path, const std::vector<T> &v) {
  if (path.empty()) return;
  std::FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  const size_t w = std::fwrite(v.data(), sizeof(T), v.size(), f);
  std::fclose(f);
  if (w != v.size()) throw std::runtime_error("short write to " + path);
}

}  // namespace mkdrv

#endif  // MEERKAT_DRIVER_IO_HPP_
