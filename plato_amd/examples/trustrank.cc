// TrustRank driver with the flags of example/trustrank.cc:15-35 (--input_edges
// --input_goods --output --is_directed --iterations --damping --seed_num --eps
// --select_method), running on the device engine.  --input_goods holds one vertex id
// per line.  Output: `v,rank` lines in <output>/0000_0.csv when --output is given.
// --random_seed fixes select_method 2's draw (the reference seeds it from
// std::random_device).
#include <chrono>
#include <cstdio>
#include <fstream>
#include <string>

#include "args.hpp"
#include "plato/algo/trustrank/trustrank.hpp"

using namespace plato_examples;

int main(int argc, char** argv) {
  auto& cluster_info = plato::cluster_info_t::get_instance();
  cluster_info.initialize(&argc, &argv);
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", false));
  const std::string input = flag(argc, argv, "input_edges", "");
  const std::string goods = flag(argc, argv, "input_goods", "");
  const unsigned method = (unsigned)std::atoi(flag(argc, argv, "select_method", "0"));
  if (input.empty() || goods.empty()) plato::fatal("flags", -1, "--input_edges and --input_goods are required");
  if (method > 2) plato::fatal("flags", -1, "--select_method must be 0, 1 or 2");
  auto t0 = std::chrono::steady_clock::now();
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, input);
  plato::bitmap_t<> good_vertices(graph_info.max_v_i_ + 1);
  {
    std::ifstream in(goods);
    if (!in) plato::fatal("open", -1, goods.c_str());
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty() || line[0] < '0' || line[0] > '9') continue;
      const unsigned long v = std::strtoul(line.c_str(), nullptr, 10);
      if (v <= graph_info.max_v_i_) good_vertices.set_bit(v);
    }
  }
  if (good_vertices.count() == 0) {
    std::fprintf(stderr, "trustrank terminated: load 0 good vertices from input file\n");
    return 0;
  }
  plato::algo::trustrank_opts_t opts;
  opts.iteration_ = (uint32_t)std::atoi(flag(argc, argv, "iterations", "100"));
  opts.damping_ = std::atof(flag(argc, argv, "damping", "0.85"));
  opts.seed_num_ = (uint32_t)std::atoi(flag(argc, argv, "seed_num", "100"));
  opts.eps_ = std::atof(flag(argc, argv, "eps", "0.001"));
  opts.select_method_ = method;
  opts.random_seed_ = std::strtoull(flag(argc, argv, "random_seed", "0"), nullptr, 10);
  auto t1 = std::chrono::steady_clock::now();
  auto ranks = plato::algo::trustrank(graph.second, graph.first, graph_info, good_vertices, opts);
  auto t2 = std::chrono::steady_clock::now();
  std::printf("vertices %u edges %llu good %zu load %.3fs trustrank %.3fs\n", graph_info.vertices_,
              (unsigned long long)graph_info.edges_, good_vertices.count(),
              std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
  const std::string output = flag(argc, argv, "output", "");
  if (!output.empty()) {
    const std::string path = output + "/0000_0.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) plato::fatal("fopen", -1, path.c_str());
    for (plato::vid_t v = 0; v <= graph_info.max_v_i_; ++v) std::fprintf(f, "%u,%.17g\n", v, ranks[v]);
    std::fclose(f);
  }
  return 0;
}
