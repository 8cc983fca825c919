// Connected components driver with the flags of example/cgm_simple.cc:36-44 (--input
// --output --is_directed --label --output_method), running on the device engine.
// --output_method: sub_graph_by_label (default; the component of --label, -1 = the
// largest), all_vertices (`v,label`) or all_edges (`label,src,dst`), written to
// <output>/part_0000_0.csv; the summary goes to stdout.
#include <chrono>
#include <cstdio>
#include <string>

#include "args.hpp"
#include "plato/algo/cgm/connected_component.hpp"

using namespace plato_examples;

int main(int argc, char** argv) {
  auto& cluster_info = plato::cluster_info_t::get_instance();
  cluster_info.initialize(&argc, &argv);
  plato::graph_info_t graph_info(flag_bool(argc, argv, "is_directed", false));
  const std::string input = flag(argc, argv, "input", "");
  const std::string output = flag(argc, argv, "output", "");
  const std::string method = flag(argc, argv, "output_method", "sub_graph_by_label");
  if (input.empty()) plato::fatal("flags", -1, "--input is required");
  auto t0 = std::chrono::steady_clock::now();
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, input);
  using dcsc_spec_t = plato::dcsc_t<plato::empty_t>;
  using bcsr_spec_t = plato::bcsr_t<plato::empty_t>;
  plato::dualmode_engine_t<dcsc_spec_t, bcsr_spec_t> engine(
      std::shared_ptr<dcsc_spec_t>(&graph.second, [](dcsc_spec_t*) {}),
      std::shared_ptr<bcsr_spec_t>(&graph.first, [](bcsr_spec_t*) {}), graph_info);
  plato::algo::connected_component_t<dcsc_spec_t, bcsr_spec_t> cc(&engine, graph_info);
  auto t1 = std::chrono::steady_clock::now();
  cc.compute();
  auto t2 = std::chrono::steady_clock::now();
  std::printf("vertices %u edges %llu load %.3fs cc %.3fs\n%s", graph_info.vertices_,
              (unsigned long long)graph_info.edges_, std::chrono::duration<double>(t1 - t0).count(),
              std::chrono::duration<double>(t2 - t1).count(), cc.get_summary().c_str());
  if (output.empty()) return 0;
  if (method == "sub_graph_by_label") cc.write_component(output, std::atoll(flag(argc, argv, "label", "-1")));
  else if (method == "all_vertices") cc.write_all_vertices(output);
  else if (method == "all_edges") cc.write_all_edges(output);
  else plato::fatal("flags", -1, "--output_method must be sub_graph_by_label, all_vertices or all_edges");
  return 0;
}
