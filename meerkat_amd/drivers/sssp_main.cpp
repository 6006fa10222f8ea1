// mk_sssp - command-line SSSP driver (the counterpart of the reference's
// src/SSSP.cu).  Weights are drawn as the reference draws them: srand(seed),
// then 1 + rand() % 100 per base edge in file order, then per batch edge
// (src/SSSP.cu:17-19,50-57,110-116); all are drawn before the device is
// initialised (the batch files are read up front).  For a file named *orkut.txt* the
// degree hints count the base edges (src/SSSP.cu:45-49), else they are 1.
//
//   mk_sssp <graph> <op 0|1> <update-prefix> <num-batches> <seed> [--dump=<path>]
//
// --dump writes the final distances (uint32, UINT32_MAX = unreachable).
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>

#include "driver_io.hpp"
#include "meerkat/Algorithms.hpp"

int main(int argc, char **argv) {
  try {
    const mkdrv::Args a = mkdrv::parse(argc, argv);
    if (a.pos.size() < 5) {
      std::fprintf(stderr, "usage: %s <graph> <op 0|1> <update-prefix> <num-batches> <seed> "
                   "[--dump=<path>]\n", argv[0]);
      return 1;
    }
    const std::string &file = a.pos[0];
    const bool insertion = a.pos[1] == "0";
    const int batches = std::stoi(a.pos[3]);
    std::srand(unsigned(std::atoi(a.pos[4].c_str())));
    const uint32_t source = std::strstr(file.c_str(), "wiki.txt") ? 2 : 0;
    const bool orkut = std::strstr(file.c_str(), "orkut.txt") != nullptr;
    std::cout << file << std::flush;

    // every file is read and every weight drawn before the first device
    // call, so nothing the runtime does can consume rand() values in between
    mkdrv::EdgeList base = mkdrv::read_graph(file);
    std::vector<uint32_t> hints(base.n, 1), w(base.src.size());
    for (size_t i = 0; i < base.src.size(); ++i) {
      w[i] = 1 + uint32_t(std::rand() % 100);
      if (orkut) ++hints[base.src[i]];
    }
    std::vector<mkdrv::EdgeList> updates;
    std::vector<std::vector<uint32_t>> update_w;
    for (int i = 0; i < batches; ++i) {
      updates.push_back(mkdrv::read_batch(a.pos[2], i));
      std::vector<uint32_t> bw(updates.back().src.size());
      for (auto &x : bw) x = 1 + uint32_t(std::rand() % 100);
      update_w.push_back(std::move(bw));
    }

    using AllocPolicy = meerkat::FullAllocatorPolicy<9, 48, 1>;
    using Allocator = AllocPolicy::DynamicAllocatorT;
    using Graph = meerkat::SSSP<AllocPolicy>::GraphT;
    std::unique_ptr<Allocator> alloc{new Allocator};
    Graph g{base.n, *alloc, 0.7, hints.data(), 0};
    {
      mkdrv::DeviceU32 s(base.src), d(base.dst), wd(w);
      g.InsertEdges(s.get(), d.get(), uint32_t(base.src.size()), wd.get());
    }
    meerkat::SSSP<AllocPolicy> sssp(g, source, base.n);
    std::cout << "," << sssp.Static(uint32_t(base.src.size())) << std::flush;

    for (int i = 0; i < batches; ++i) {
      const mkdrv::EdgeList &b = updates[size_t(i)];
      mkdrv::DeviceU32 s(b.src), d(b.dst), wd(update_w[size_t(i)]);
      const uint32_t n = uint32_t(b.src.size());
      float ms;
      if (insertion) {
        g.InsertEdges(s.get(), d.get(), n, wd.get());
        ms = sssp.Incremental(s.get(), d.get(), wd.get(), n);
      } else {
        g.DeleteEdges(s.get(), d.get(), n);
        ms = sssp.Decremental(s.get(), d.get(), wd.get(), n);
      }
      std::cout << "," << ms << std::flush;
    }
    std::cout << std::endl;
    mkdrv::dump(a.dump, sssp.Distances());
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "\nmk_sssp: " << e.what() << std::endl;
    return 2;
  }
}
