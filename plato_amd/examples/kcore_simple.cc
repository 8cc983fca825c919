// k-core driver with the flags of example/kcore_simple.cc:36-48 (--input --output --type
// --is_directed), running on the device engine.  The graph is read as the reference
// driver reads it (a BCSR over incoming edges).  --type vertex writes `v,coreness` lines to
// <output>/0000_0.csv; --type subgraph (the default) writes, for k = 0, 1, ... until
// every vertex is saved, the adjacency `v,u` of the vertices of coreness k to
// <output>/<k>_core/0000_0.csv (kcore_simple.cc:127-160).
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "args.hpp"
#include "plato/algo/kcore/kcore.hpp"

using namespace plato_examples;

int main(int argc, char** argv) {
  auto& cluster_info = plato::cluster_info_t::get_instance();
  cluster_info.initialize(&argc, &argv);
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", true));
  const std::string input = flag(argc, argv, "input", "");
  const std::string output = flag(argc, argv, "output", "");
  const std::string type = flag(argc, argv, "type", "subgraph");
  if (input.empty()) plato::fatal("flags", -1, "--input is required");
  if (type != "subgraph" && type != "vertex") plato::fatal("flags", -1, "--type must be subgraph or vertex");
  auto t0 = std::chrono::steady_clock::now();
  auto graph = plato::create_bcsr_seqs_from_path<plato::empty_t>(&graph_info, input, plato::edge_format_t::CSV,
                                                                 nullptr, -1, false, nullptr, false);
  auto t1 = std::chrono::steady_clock::now();
  auto coreness = plato::algo::kcore_algo_t::compute_shell_index(graph_info, *graph, [](plato::vid_t, plato::vid_t,
                                                                                          uint32_t) {});
  auto t2 = std::chrono::steady_clock::now();
  std::printf("vertices %u edges %llu max_v_i %u load %.3fs kcore %.3fs\n", graph_info.vertices_,
              (unsigned long long)graph_info.edges_, graph_info.max_v_i_,
              std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
  if (output.empty()) return 0;
  const plato::vid_t n = graph_info.max_v_i_ + 1;
  if (type == "vertex") {
    const std::string path = output + "/0000_0.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) plato::fatal("fopen", -1, path.c_str());
    for (plato::vid_t v = 0; v < n; ++v) std::fprintf(f, "%u,%u\n", v, coreness[v]);
    std::fclose(f);
    return 0;
  }
  // subgraph: the incoming adjacency of each vertex, grouped by coreness
  const plato::vid_t V = graph->info().V;
  std::vector<uint64_t> ptr((size_t)V + 1);
  std::vector<uint32_t> idx(graph->info().nnz ? graph->info().nnz : 1);
  plato::check_status(pe_graph_download_csc(graph->handle(), ptr.data(), idx.data(), nullptr), "pe_graph_download_csc");
  std::vector<bool> left(n, true);
  plato::vid_t saved = 0;
  for (uint32_t k = 0; saved < n; ++k) {
    const std::string dir = output + "/" + std::to_string(k) + "_core";
    ::mkdir(dir.c_str(), 0755);
    const std::string path = dir + "/0000_0.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) plato::fatal("fopen", -1, path.c_str());
    for (plato::vid_t v = 0; v < n; ++v) {
      if (!left[v] || coreness[v] != k) continue;
      for (uint64_t e = ptr[v]; e < ptr[v + 1]; ++e) std::fprintf(f, "%u,%u\n", v, idx[e]);
      left[v] = false;
      ++saved;
    }
    std::fclose(f);
  }
  return 0;
}
