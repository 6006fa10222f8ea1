// mk_pagerank - command-line PageRank driver (the counterpart of the
// reference's src/PageRank.cu): out-degrees counted from the base file, the
// edges inserted into a Rev-policy graph (stored reversed), each batch file
// de-duplicated through an ordered set (src/PageRank.cu:90-108) before it is
// inserted (op 0) or deleted (op 1) and the ranks updated.
//
//   mk_pagerank <graph> <op 0|1> <update-prefix> <num-batches> <damping> <margin>
//               [--dump=<path>]
//
// --dump writes the final ranks (float32).
#include <cstdlib>
#include <iostream>
#include <memory>
#include <set>
#include <utility>

#include "driver_io.hpp"
#include "meerkat/Algorithms.hpp"

int main(int argc, char **argv) {
  try {
    const mkdrv::Args a = mkdrv::parse(argc, argv);
    if (a.pos.size() < 6) {
      std::fprintf(stderr, "usage: %s <graph> <op 0|1> <update-prefix> <num-batches> "
                   "<damping> <margin> [--dump=<path>]\n", argv[0]);
      return 1;
    }
    const std::string &file = a.pos[0];
    const bool insertion = a.pos[1] == "0";
    const int batches = std::stoi(a.pos[3]);
    const float damping = float(std::atof(a.pos[4].c_str()));
    const float margin = float(std::atof(a.pos[5].c_str()));
    std::cout << file << std::flush;

    mkdrv::EdgeList base = mkdrv::read_graph(file);
    std::vector<uint32_t> outdeg(base.n, 0), in_hints(base.n, 0);
    for (size_t i = 0; i < base.src.size(); ++i) {
      ++outdeg[base.src[i]];
      ++in_hints[base.dst[i]];
    }

    using AllocPolicy = meerkat::LightAllocatorPolicy<8, 32, 1>;
    using Allocator = AllocPolicy::DynamicAllocatorT;
    using Graph = meerkat::PageRank<AllocPolicy>::GraphT;
    std::unique_ptr<Allocator> alloc{new Allocator};
    // the reference passes an all-zero in-degree array here
    // (src/PageRank.cu:42-45,70); the real in-degrees size the buckets
    Graph g{base.n, *alloc, 0.7, in_hints.data(), 0};
    {
      mkdrv::DeviceU32 s(base.src), d(base.dst);
      g.InsertEdges(s.get(), d.get(), uint32_t(base.src.size()));
    }
    meerkat::PageRank<AllocPolicy> pr(g, damping, margin, base.n, outdeg.data(), 100);
    std::cout << "," << pr.Static() << std::flush;

    for (int i = 0; i < batches; ++i) {
      mkdrv::EdgeList b = mkdrv::read_batch(a.pos[2], i);
      std::set<std::pair<uint32_t, uint32_t>> uniq;
      for (size_t k = 0; k < b.src.size(); ++k) uniq.emplace(b.src[k], b.dst[k]);
      std::vector<uint32_t> us, ud;
      us.reserve(uniq.size());
      ud.reserve(uniq.size());
      for (const auto &e : uniq) {
        us.push_back(e.first);
        ud.push_back(e.second);
      }
      mkdrv::DeviceU32 s(us), d(ud);
      const uint32_t n = uint32_t(us.size());
      float ms;
      if (insertion) {
        g.InsertEdges(s.get(), d.get(), n);
        ms = pr.Incremental(s.get(), d.get(), n);
      } else {
        g.DeleteEdges(s.get(), d.get(), n);
        ms = pr.Decremental(s.get(), d.get(), n);
      }
      std::cout << "," << ms << std::flush;
    }
    std::cout << std::endl;
    mkdrv::dump(a.dump, pr.Values());
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "\nmk_pagerank: " << e.what() << std::endl;
    return 2;
  }
}
