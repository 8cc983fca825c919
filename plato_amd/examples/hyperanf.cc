// HyperANF driver with the flags of example/hyperanf.cc:34-39 (--input --is_directed
// --iterations --bits), running on the device engine; prints the average distance.
#include <chrono>
#include <cstdio>
#include <string>

#include "args.hpp"
#include "plato/algo/hyperanf/hyperanf.hpp"

using namespace plato_examples;

template <uint32_t B, typename IN, typename OUT>
static double run(IN& in, OUT& out, const plato::graph_info_t& info, const plato::algo::hyperanf_opts_t& o) {
  return plato::algo::hyperanf<IN, OUT, B>(in, out, info, o);
}

int main(int argc, char** argv) {
  auto& cluster_info = plato::cluster_info_t::get_instance();
  cluster_info.initialize(&argc, &argv);
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", false));
  const std::string input = flag(argc, argv, "input", "");
  if (input.empty()) plato::fatal("flags", -1, "--input is required");
  plato::algo::hyperanf_opts_t opts;
  opts.iteration_ = (uint32_t)std::atoi(flag(argc, argv, "iterations", "20"));
  opts.bits = (uint32_t)std::atoi(flag(argc, argv, "bits", "6"));
  auto t0 = std::chrono::steady_clock::now();
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, input);
  auto& in = graph.second;
  auto& out = graph.first;
  auto t1 = std::chrono::steady_clock::now();
  double avg = 0.0;
  switch (opts.bits) {
    case 6: avg = run<6>(in, out, graph_info, opts); break;
    case 7: avg = run<7>(in, out, graph_info, opts); break;
    case 8: avg = run<8>(in, out, graph_info, opts); break;
    case 9: avg = run<9>(in, out, graph_info, opts); break;
    case 10: avg = run<10>(in, out, graph_info, opts); break;
    case 11: avg = run<11>(in, out, graph_info, opts); break;
    case 12: avg = run<12>(in, out, graph_info, opts); break;
    case 13: avg = run<13>(in, out, graph_info, opts); break;
    case 14: avg = run<14>(in, out, graph_info, opts); break;
    case 15: avg = run<15>(in, out, graph_info, opts); break;
    case 16: avg = run<16>(in, out, graph_info, opts); break;
    default: plato::fatal("flags", -1, "--bits must be in [6, 16]");
  }
  auto t2 = std::chrono::steady_clock::now();
  std::printf("hyperanf done, avg_distance: %.17g, load %.3fs, cost %.3fs\n", avg,
              std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
  return 0;
}
