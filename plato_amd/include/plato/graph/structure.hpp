// Graph structures and factories on the device.
//
// dcsc_t / bcsr_t (plato/graph/structure/dcsc.hpp:58-185, bcsr.hpp:67-208)
// become handles onto one device graph (CSC over all ids, optional CSR);
// sequence_balanced_by_* (plato/graph/partition/sequence.hpp) is the
// one-partition identity.  The factories keep the names and arguments of
// create_dcsc_seqs_from_path / create_bcsr_seqd_from_path /
// create_dualmode_seq_from_path (plato/graph/structure.hpp:473-555, 987-1086)
// over Plato's CSV edge format `src,dst[,w]` (parsers.hpp:225-293), parsed on the device.
#pragma once

#include <algorithm>
#include <fstream>
#include <iterator>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "plato/graph/base.hpp"
#include "plato/graph/state/dense_state.hpp"

namespace plato {

enum class edge_format_t { CSV = 0 };

struct sequence_v_view {
  vid_t start_, end_;
  vid_t start() const { return start_; }
  vid_t end() const { return end_; }
};

// one partition covering every vertex id
struct sequence_balanced_by_destination_t {
  vid_t vertices_ = 0;
  sequence_v_view self_v_view() const { return {0, vertices_}; }
  int get_partition_id(vid_t) const { return 0; }
};
using sequence_balanced_by_source_t = sequence_balanced_by_destination_t;

// Shared owner of one device graph.
struct device_graph {
  pe_graph_t* h = nullptr;
  pe_graph_info_t info{};
  ~device_graph() {
    if (h) pe_graph_destroy(h);
  }
};

namespace detail {
template <typename EDATA>
class graph_facade {
 public:
  using partition_t = sequence_balanced_by_destination_t;
  using edata_t = EDATA;
  struct adj_unit_list_spec_t {};  // adjacency spans live on the device

  graph_facade() = default;
  explicit graph_facade(std::shared_ptr<device_graph> g) : g_(std::move(g)) {
    part_ = std::make_shared<partition_t>();
    part_->vertices_ = g_->info.V;
  }
  std::shared_ptr<partition_t> partitioner() const { return part_; }
  pe_graph_t* handle() const { return g_->h; }
  // Pull layout of the FP32 rank iterations on this graph (PE_PULL_MERGE_PATH or
  // PE_PULL_BINNED; include/plato_hip.h).  The binned layout is built on the first call.
  void set_pull_layout(int layout, const pe_binned_opts_t* opts = nullptr) {
    check_status(pe_graph_set_pull_layout(g_->h, layout, opts), "pe_graph_set_pull_layout");
    check_status(pe_graph_get_info(g_->h, &g_->info), "pe_graph_get_info");
  }
  const pe_graph_info_t& info() const { return g_->info; }
  std::shared_ptr<device_graph> shared() const { return g_; }
  // true when the structure's adjacency lists are out-edges (a BCSR built with
  // is_outgoing), false for in-edges (a DCSC, or a BCSR built over incoming edges)
  bool outgoing() const { return outgoing_; }
  void set_outgoing(bool o) { outgoing_ = o; }

 private:
  std::shared_ptr<device_graph> g_;
  std::shared_ptr<partition_t> part_;
  bool outgoing_ = false;
};
}  // namespace detail

template <typename EDATA, typename PART = sequence_balanced_by_destination_t>
class dcsc_t : public detail::graph_facade<EDATA> {
  using detail::graph_facade<EDATA>::graph_facade;
};

template <typename EDATA, typename PART = sequence_balanced_by_source_t>
class bcsr_t : public detail::graph_facade<EDATA> {
  using detail::graph_facade<EDATA>::graph_facade;
};

// decoders: float weights or none
inline float float_decoder(const std::string& s) { return std::stof(s); }
template <typename T>
inline T dummy_decoder(const std::string&) { return T(); }

namespace detail {
inline std::string read_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) fatal("open", -1, path.c_str());
  return std::string((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

// The file's text goes to the device, which parses it with csv_parser's rules
// (parsers.hpp:225-293) and builds the CSC/CSR (pe_graph_create_csv).
inline std::shared_ptr<device_graph> build(graph_info_t* info, const std::string& path, bool weighted,
                                           bool build_csr, bool relabel) {
  const std::string text = read_file(path);
  uint32_t flags = info->is_directed_ ? 0u : PE_GRAPH_SYMMETRIZE;
  if (build_csr) flags |= PE_GRAPH_BUILD_CSR;
  if (weighted) flags |= PE_GRAPH_WEIGHTED;
  if (relabel) flags |= PE_GRAPH_RELABEL_DEGREE;
  auto g = std::make_shared<device_graph>();
  check_status(pe_graph_create_csv(cluster_info_t::get_instance().ctx(), text.data(), text.size(), flags, &g->h),
               "pe_graph_create_csv");
  check_status(pe_graph_get_info(g->h, &g->info), "pe_graph_get_info");
  info->vertices_ = g->info.vertices;
  info->edges_ = g->info.edges;
  info->max_v_i_ = g->info.max_v_i;
  return g;
}
}  // namespace detail

// The reference's loaders stage edges in host caches (edge_block_cache_t,
// edge_file_cache_t, graph/structure/edge_cache.hpp); here the device parses the file,
// so these types only keep the template arguments of the reference's call sites.
template <typename EDATA, typename VID_T = vid_t>
struct edge_block_cache_t {};
template <typename EDATA, typename VID_T = vid_t>
struct edge_file_cache_t {};

// create_dcsc_seqs_from_path (structure.hpp:544-555): same template and argument lists;
// alpha and use_in_degree shape the reference's MPI partitions and have no meaning on
// one device; a vid encoder is not supported (ids are used as given).
template <typename EDATA, typename VID_T = vid_t, template <typename, typename> class CACHE = edge_block_cache_t,
          typename DECODER = void*, typename ENCODER = std::nullptr_t>
std::shared_ptr<dcsc_t<EDATA>> create_dcsc_seqs_from_path(graph_info_t* info, const std::string& path,
                                                          edge_format_t = edge_format_t::CSV,
                                                          DECODER = nullptr, int /*alpha*/ = -1,
                                                          bool /*use_in_degree*/ = false, ENCODER = nullptr) {
  const bool weighted = !std::is_same<EDATA, empty_t>::value;
  auto g = detail::build(info, path, weighted, false, false);
  return std::make_shared<dcsc_t<EDATA>>(g);
}

template <typename EDATA, typename DECODER = void*>
std::shared_ptr<bcsr_t<EDATA>> create_bcsr_seqd_from_path(graph_info_t* info, const std::string& path,
                                                          edge_format_t = edge_format_t::CSV,
                                                          DECODER = nullptr) {
  const bool weighted = !std::is_same<EDATA, empty_t>::value;
  auto g = detail::build(info, path, weighted, true, false);
  auto b = std::make_shared<bcsr_t<EDATA>>(g);
  b->set_outgoing(true);
  return b;
}

// create_bcsr_seqs_from_path (structure.hpp:727-739): is_outgoing = false keeps the
// incoming edges in the structure (as example/kcore_simple.cc asks for).
template <typename EDATA, typename VID_T = vid_t, template <typename, typename> class CACHE = edge_block_cache_t,
          typename DECODER = void*, typename ENCODER = std::nullptr_t>
std::shared_ptr<bcsr_t<EDATA>> create_bcsr_seqs_from_path(graph_info_t* info, const std::string& path,
                                                          edge_format_t = edge_format_t::CSV,
                                                          DECODER = nullptr, int /*alpha*/ = -1,
                                                          bool /*use_in_degree*/ = false,
                                                          ENCODER = nullptr, bool is_outgoing = true) {
  const bool weighted = !std::is_same<EDATA, empty_t>::value;
  auto g = detail::build(info, path, weighted, is_outgoing, false);
  auto b = std::make_shared<bcsr_t<EDATA>>(g);
  b->set_outgoing(is_outgoing);
  return b;
}

// create_dualmode_seq_from_path (structure.hpp:978-1086): same template and argument lists.
template <typename EDATA, typename VID_T = vid_t, template <typename, typename> class CACHE = edge_block_cache_t,
          typename DECODER = void*, typename ENCODER = std::nullptr_t>
std::pair<bcsr_t<EDATA>, dcsc_t<EDATA>> create_dualmode_seq_from_path(
    graph_info_t* info, const std::string& path, edge_format_t = edge_format_t::CSV, DECODER = nullptr,
    int /*alpha*/ = -1, bool /*use_in_degree*/ = false, ENCODER = nullptr) {
  const bool weighted = !std::is_same<EDATA, empty_t>::value;
  auto g = detail::build(info, path, weighted, true, false);
  bcsr_t<EDATA> out(g);
  out.set_outgoing(true);
  return {out, dcsc_t<EDATA>(g)};
}

// out-degrees counted from the in-edges (structure.hpp:292-346)
template <typename T, typename GRAPH>
host_state_t<T, typename GRAPH::partition_t> generate_dense_out_degrees_fg(const graph_info_t&, GRAPH& graph,
                                                                           bool /*is_out_edge*/) {
  std::vector<uint32_t> od(graph.info().V);
  check_status(pe_graph_out_degrees(graph.handle(), od.data()), "pe_graph_out_degrees");
  return host_state_t<T, typename GRAPH::partition_t>(std::vector<T>(od.begin(), od.end()));
}

}  // namespace plato
