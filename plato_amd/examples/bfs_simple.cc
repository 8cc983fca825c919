// BFS driver with the flags of example/bfs_simple.cc (--input --is_directed
// --root), running plato::algo::breadth_first_search on the device engine.
// Prints the per-epoch frontier sizes and the visited count; --output writes
// `v,level` lines to <output>/0000_0.csv.
#include <cstdio>
#include <string>
#include <vector>

#include "args.hpp"
#include "plato/algo/bfs/bfs.hpp"

using namespace plato_examples;

int main(int argc, char** argv) {
  plato::cluster_info_t::get_instance().initialize(&argc, &argv);
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", false));
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, flag(argc, argv, "input", ""));
  plato::algo::bfs_opts_t opts;
  opts.root_ = (plato::vid_t)std::atoll(flag(argc, argv, "root", "0"));
  std::vector<int32_t> lv;
  const plato::vid_t visited =
      plato::algo::breadth_first_search(graph.second, graph.first, graph_info, opts, &lv, true);
  std::printf("visited: %u\n", visited);
  const std::string output = flag(argc, argv, "output", "");
  if (!output.empty()) {
    const std::string path = output + "/0000_0.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) plato::fatal("fopen", -1, path.c_str());
    for (plato::vid_t v = 0; v <= graph_info.max_v_i_; ++v) std::fprintf(f, "%u,%d\n", v, lv[v]);
    std::fclose(f);
  }
  return 0;
}
