// n-step degrees driver with the flags of example/nstepdegrees.cc:42-54 (--input --output
// --is_directed --step --bits --type --actives), running nstepdegrees_t on the device
// engine; writes `v,in,out` lines for the target vertices to <output>/0000_0.csv.
// --actives ALL (or no --actives) targets every vertex; otherwise a file of vertex ids,
// one per line.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <string>

#include "args.hpp"
#include "plato/algo/nstepdegrees/nstepdegrees.hpp"
#include "plato/util/foutput.h"

using namespace plato_examples;
using bcsr_spec_t = plato::bcsr_t<plato::empty_t>;
using dcsc_spec_t = plato::dcsc_t<plato::empty_t>;
using GRAPH_T = std::pair<bcsr_spec_t, dcsc_spec_t>;

template <uint32_t BitWidth>
static void work_flow(plato::dualmode_engine_t<dcsc_spec_t, bcsr_spec_t>* engine, const plato::graph_info_t& graph_info,
                      GRAPH_T& graph, const plato::bitmap_t<>& actives_v, const plato::algo::nstepdegree_opts_t& opts,
                      const std::string& output) {
  plato::algo::nstepdegrees_t<dcsc_spec_t, bcsr_spec_t, BitWidth> nstepdegrees(engine, graph_info, actives_v, opts);
  nstepdegrees.compute(graph.second, graph.first);
  plato::fs_mt_omp_output_t os(output, "0000_", false);
  nstepdegrees.save(os);
  std::printf("nstepdegrees done, device %.3fs\n", nstepdegrees.seconds());
}

int main(int argc, char** argv) {
  auto& cluster_info = plato::cluster_info_t::get_instance();
  cluster_info.initialize(&argc, &argv);
  const std::string input = flag(argc, argv, "input", "");
  const std::string output = flag(argc, argv, "output", "");
  if (input.empty() || output.empty()) plato::fatal("flags", -1, "--input and --output are required");
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", false));
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, input);
  plato::algo::nstepdegree_opts_t opts;
  opts.step = (uint32_t)std::atoll(flag(argc, argv, "step", "-1"));
  opts.type = flag(argc, argv, "type", "both");
  opts.is_directed = graph_info.is_directed_;
  plato::dualmode_engine_t<dcsc_spec_t, bcsr_spec_t> engine(
      std::shared_ptr<dcsc_spec_t>(&graph.second, [](dcsc_spec_t*) {}),
      std::shared_ptr<bcsr_spec_t>(&graph.first, [](bcsr_spec_t*) {}), graph_info);
  auto actives_v = engine.alloc_v_subset();
  const std::string actives = flag(argc, argv, "actives", "ALL");
  if (actives != "ALL") {
    std::ifstream in(actives);
    if (!in) plato::fatal("open", -1, actives.c_str());
    unsigned long long v = 0;
    while (in >> v) actives_v.set_bit(v);
  } else {
    actives_v.fill();
  }
  switch (std::atoi(flag(argc, argv, "bits", "6"))) {
    case 6: work_flow<6>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 7: work_flow<7>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 8: work_flow<8>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 9: work_flow<9>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 10: work_flow<10>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 11: work_flow<11>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 12: work_flow<12>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 13: work_flow<13>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 14: work_flow<14>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 15: work_flow<15>(&engine, graph_info, graph, actives_v, opts, output); break;
    case 16: work_flow<16>(&engine, graph_info, graph, actives_v, opts, output); break;
    default: plato::fatal("flags", -1, "--bits must be in [6, 16]");
  }
  return 0;
}
