// PPR driver with the flags of example/personalized_pagerank.cc:15-30 and
// example/pushpull_ppr.cc (--input --is_directed --src --iterations --damping
// --eps; --type pull|pushpull), on the device engine.
#include <cstdio>
#include <string>

#include "args.hpp"
#include "plato/algo/ppr/personalized_pagerank.hpp"

using namespace plato_examples;

int main(int argc, char** argv) {
  plato::cluster_info_t::get_instance().initialize(&argc, &argv);
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", false));
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, flag(argc, argv, "input", ""));
  plato::algo::personalized_pagerank_opts_t opts;
  opts.iteration_ = std::atoi(flag(argc, argv, "iterations", "100"));
  opts.damping_ = std::atof(flag(argc, argv, "damping", "0.85"));
  opts.eps_ = std::atof(flag(argc, argv, "eps", "0.001"));
  opts.src_ = (plato::vid_t)std::atoll(flag(argc, argv, "src", "0"));
  opts.precision_ = std::atoi(flag(argc, argv, "precision", "32"));
  const std::string type = flag(argc, argv, "type", "pull");
  auto ranks = type == "pushpull"
                   ? plato::algo::personalized_pagerank(graph.second, graph.first, graph_info, opts)
                   : plato::algo::personalized_pagerank(graph.second, graph_info, opts);
  const std::string output = flag(argc, argv, "output", "");
  if (!output.empty()) {
    const std::string path = output + "/0000_0.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) plato::fatal("fopen", -1, path.c_str());
    for (plato::vid_t v = 0; v <= graph_info.max_v_i_; ++v) std::fprintf(f, "%u,%.17g\n", v, ranks[v]);
    std::fclose(f);
  }
  double s = 0;
  for (plato::vid_t v = 0; v <= graph_info.max_v_i_; ++v) s += ranks[v];
  std::printf("sum of ranks %.12f\n", s);
  return 0;
}
