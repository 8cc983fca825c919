// PageRank driver with the flags of example/pagerank_simple.cc:36-51
// (--input --output --is_directed --iterations --damping --eps), running on the
// device engine.  Output: `v,rank` lines in <output>/0000_0.csv when --output is given.
// --pull_layout binned selects the partition-centric pull (FP32 runs).
#include <chrono>
#include <cstdio>
#include <string>

#include "args.hpp"
#include "plato/algo/pagerank/pagerank.hpp"

using namespace plato_examples;

int main(int argc, char** argv) {
  auto& cluster_info = plato::cluster_info_t::get_instance();
  cluster_info.initialize(&argc, &argv);
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", false));
  const std::string input = flag(argc, argv, "input", "");
  auto t0 = std::chrono::steady_clock::now();
  auto pdcsc = plato::create_dcsc_seqs_from_path<plato::empty_t>(&graph_info, input);
  auto t1 = std::chrono::steady_clock::now();
  plato::algo::pagerank_opts_t opts;
  opts.iteration_ = std::atoi(flag(argc, argv, "iterations", "100"));
  opts.damping_ = std::atof(flag(argc, argv, "damping", "0.85"));
  opts.eps_ = std::atof(flag(argc, argv, "eps", "0.001"));
  opts.precision_ = std::atoi(flag(argc, argv, "precision", "32"));
  if (std::string(flag(argc, argv, "pull_layout", "merge_path")) == "binned") pdcsc->set_pull_layout(PE_PULL_BINNED);
  auto ranks = plato::algo::pagerank(*pdcsc, graph_info, opts);
  auto t2 = std::chrono::steady_clock::now();
  std::printf("vertices %u edges %llu max_v_i %u load %.3fs pagerank %.3fs\n", graph_info.vertices_,
              (unsigned long long)graph_info.edges_, graph_info.max_v_i_,
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
