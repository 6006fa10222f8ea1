// mk_bfs - command-line BFS driver (the counterpart of the reference's
// src/BFS.cu): builds the graph from a text edge list, runs static BFS from
// vertex 0 (2 for a file named *wiki.txt*, as src/BFS.cu:20-21), then for
// each batch file inserts (op 0) or deletes (op 1) it and updates the
// levels incrementally / decrementally.  Prints one CSV line of milliseconds.
//
//   mk_bfs <graph> <op 0|1> <update-prefix> <num-batches> [--dump=<path>]
//
// --dump writes the final levels (uint32, UINT32_MAX = unreachable).
#include <cstring>
#include <iostream>
#include <memory>

#include "driver_io.hpp"
#include "meerkat/Algorithms.hpp"

int main(int argc, char **argv) {
  try {
    const mkdrv::Args a = mkdrv::parse(argc, argv);
    if (a.pos.size() < 4) {
      std::fprintf(stderr, "usage: %s <graph> <op 0|1> <update-prefix> <num-batches> "
                   "[--dump=<path>]\n", argv[0]);
      return 1;
    }
    const std::string &file = a.pos[0];
    const bool insertion = a.pos[1] == "0";
    const int batches = std::stoi(a.pos[3]);
    const uint32_t source = std::strstr(file.c_str(), "wiki.txt") ? 2 : 0;
    std::cout << file << std::flush;

    mkdrv::EdgeList base = mkdrv::read_graph(file);
    std::vector<uint32_t> hints(base.n, 1);  // src/BFS.cu:29-30

    using AllocPolicy = meerkat::FullAllocatorPolicy<9, 48, 1>;
    using Allocator = AllocPolicy::DynamicAllocatorT;
    using Graph = meerkat::BFS<AllocPolicy>::GraphT;
    std::unique_ptr<Allocator> alloc{new Allocator};
    Graph g{base.n, *alloc, 0.7, hints.data(), 0};
    {
      mkdrv::DeviceU32 s(base.src), d(base.dst);
      g.InsertEdges(s.get(), d.get(), uint32_t(base.src.size()));
    }
    meerkat::BFS<AllocPolicy> bfs(g, source, base.n);
    std::cout << "," << bfs.Static(uint32_t(base.src.size())) << std::flush;

    for (int i = 0; i < batches; ++i) {
      mkdrv::EdgeList b = mkdrv::read_batch(a.pos[2], i);
      mkdrv::DeviceU32 s(b.src), d(b.dst);
      const uint32_t n = uint32_t(b.src.size());
      float ms;
      if (insertion) {
        g.InsertEdges(s.get(), d.get(), n);
        ms = bfs.Incremental(s.get(), d.get(), n);
      } else {
        g.DeleteEdges(s.get(), d.get(), n);
        ms = bfs.Decremental(s.get(), d.get(), n);
      }
      std::cout << "," << ms << std::flush;
    }
    std::cout << std::endl;
    mkdrv::dump(a.dump, bfs.Distances());
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "\nmk_bfs: " << e.what() << std::endl;
    return 2;
  }
}
