// mk_tc - command-line triangle-count driver (the counterpart of the
// reference's src/TC.cu).  The base file and every batch file must hold
// each undirected edge in both directions, as the reference assumes.  Degree
// hints count the base edges (src/TC.cu:43-46).  Also runs connected
// components after each batch (edges taken as undirected) and prints their
// times on a second CSV line.
//
//   mk_tc <graph> <op 0|1> <update-prefix> <num-batches> [--dump=<path>]
//
// --dump writes the final triangle count (uint64) followed by the
// component labels (uint32, smallest vertex id of each component).
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
    std::cout << file << std::flush;

    mkdrv::EdgeList base = mkdrv::read_graph(file);
    std::vector<uint32_t> hints(base.n, 1);
    for (uint32_t s : base.src) ++hints[s];

    using AllocPolicy = meerkat::LightAllocatorPolicy<8, 32, 1>;
    using Allocator = AllocPolicy::DynamicAllocatorT;
    using Graph = meerkat::TC<AllocPolicy>::GraphT;
    std::unique_ptr<Allocator> alloc{new Allocator};
    Graph g{base.n, *alloc, 0.7, hints.data(), 0};
    {
      mkdrv::DeviceU32 s(base.src), d(base.dst);
      g.InsertEdges(s.get(), d.get(), uint32_t(base.src.size()));
    }
    meerkat::TC<AllocPolicy> tc(g, base.n);
    meerkat::CC<AllocPolicy> cc(g, base.n);
    std::cout << "," << tc.Static(uint32_t(base.src.size())) << std::flush;
    std::string cc_line = "cc," + std::to_string(cc.Static());

    for (int i = 0; i < batches; ++i) {
      mkdrv::EdgeList b = mkdrv::read_batch(a.pos[2], i);
      mkdrv::DeviceU32 s(b.src), d(b.dst);
      const uint32_t n = uint32_t(b.src.size());
      float ms, cms;
      if (insertion) {
        g.InsertEdges(s.get(), d.get(), n);
        ms = tc.Incremental(s.get(), d.get(), n, alloc);
        cms = cc.Incremental(s.get(), d.get(), n);
      } else {
        g.DeleteEdges(s.get(), d.get(), n);
        ms = tc.Decremental(s.get(), d.get(), n, alloc);
        cms = cc.Decremental(s.get(), d.get(), n);
      }
      std::cout << "," << ms << std::flush;
      cc_line += "," + std::to_string(cms);
    }
    std::cout << std::endl << cc_line << std::endl;
    if (!a.dump.empty()) {
      const std::vector<uint32_t> labels = cc.Labels();
      const unsigned long long count = tc.GetCount();
      std::vector<uint32_t> out(2 + labels.size());
      std::memcpy(out.data(), &count, 8);
      std::memcpy(out.data() + 2, labels.data(), labels.size() * 4);
      mkdrv::dump(a.dump, out);
    }
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "\nmk_tc: " << e.what() << std::endl;
    return 2;
  }
}
