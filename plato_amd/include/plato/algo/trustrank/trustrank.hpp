// plato::algo::trustrank on the device engine (plato/algo/trustrank/trustrank.hpp:23-259):
// generate_rank_pairs, select_good_seeds and trustrank with the reference's options and
// result shapes.  Seed ranking and the iteration run on the device (pe_trustrank_*);
// inverse PageRank pulls over a reversed copy of the graph built for the call.
#pragma once

#include <vector>

#include "plato/algo/pagerank/pagerank.hpp"
#include "plato/util/bitmap.hpp"

namespace plato {
namespace algo {

struct trustrank_opts_t {
  uint32_t iteration_ = 100;     // number of iterations
  double damping_ = 0.85;        // the damping factor
  uint32_t seed_num_ = 100;      // number of seeds
  uint32_t select_method_ = 0;   // 0 inverse PageRank, 1 PageRank, 2 random
  double eps_ = 0.001;           // stop when delta < eps (then one final iteration)
  uint64_t random_seed_ = 0;     // method 2 (the reference draws it from std::random_device)
  int precision_ = 64;
};

struct rank_pair_t {
  vid_t v_i;
  double rank;
};

// each vertex's id and PageRank (default pagerank_opts_t), ids 0..max_v_i
template <typename GRAPH>
std::vector<rank_pair_t> generate_rank_pairs(GRAPH& graph, const graph_info_t& graph_info) {
  pagerank_opts_t o;
  o.precision_ = 64;
  auto ranks = pagerank(graph, graph_info, o);
  std::vector<rank_pair_t> pairs(graph_info.max_v_i_ + 1);
  for (vid_t v = 0; v <= graph_info.max_v_i_; ++v) pairs[v] = rank_pair_t{v, ranks[v]};
  return pairs;
}

namespace detail {
inline pe_trustrank_opts_t to_c(const trustrank_opts_t& opts) {
  pe_trustrank_opts_t o{};
  o.iterations = opts.iteration_;
  o.damping = opts.damping_;
  o.seed_num = opts.seed_num_;
  o.select_method = opts.select_method_;
  o.eps = opts.eps_;
  o.random_seed = opts.random_seed_;
  o.precision = opts.precision_;
  return o;
}

inline std::vector<uint64_t> good_words(const bitmap_t<>& good, uint32_t V) {
  std::vector<uint64_t> w((V >> 6) + 1, 0);
  const uint64_t have = (good.size_ >> 6) + 1;
  std::vector<uint64_t> tmp(have);
  check_status(pe_bitmap_download(good.handle(), tmp.data(), have), "pe_bitmap_download");
  for (uint64_t i = 0; i < have && i < w.size(); ++i) w[i] = tmp[i];
  return w;
}

// the reversed graph, when inverse PageRank on a directed graph needs it
struct reverse_guard {
  pe_graph_t* h = nullptr;
  reverse_guard(pe_graph_t* g, const graph_info_t& info, uint32_t method) {
    if (method == 0 && info.is_directed_) check_status(pe_graph_create_reverse(g, 0, &h), "pe_graph_create_reverse");
  }
  ~reverse_guard() {
    if (h) pe_graph_destroy(h);
  }
};
}  // namespace detail

// select_good_seeds (trustrank.hpp:80-145)
template <typename INCOMING, typename OUTGOING>
bitmap_t<> select_good_seeds(INCOMING& in_edges, OUTGOING& /*out_edges*/, const graph_info_t& graph_info,
                             const bitmap_t<>& good_vertices, const trustrank_opts_t& opts) {
  const uint32_t V = in_edges.info().V;
  auto good = detail::good_words(good_vertices, V);
  detail::reverse_guard rev(in_edges.handle(), graph_info, opts.select_method_);
  const pe_trustrank_opts_t o = detail::to_c(opts);
  std::vector<uint64_t> seeds((V >> 6) + 1, 0);
  uint32_t n = 0;
  check_status(pe_trustrank_select(in_edges.handle(), rev.h, good.data(), &o, seeds.data(), &n),
               "pe_trustrank_select");
  bitmap_t<> out(V);
  check_status(pe_bitmap_upload(out.handle(), seeds.data(), seeds.size()), "pe_bitmap_upload");
  return out;
}

// trustrank (trustrank.hpp:163-259): each vertex's rank, dense
template <typename INCOMING, typename OUTGOING>
dense_state_t<double, typename INCOMING::partition_t> trustrank(INCOMING& in_edges, OUTGOING& /*out_edges*/,
                                                             const graph_info_t& graph_info,
                                                             const bitmap_t<>& good_vertices,
                                                             const trustrank_opts_t& opts = trustrank_opts_t()) {
  const uint32_t V = in_edges.info().V;
  auto good = detail::good_words(good_vertices, V);
  detail::reverse_guard rev(in_edges.handle(), graph_info, opts.select_method_);
  const pe_trustrank_opts_t o = detail::to_c(opts);
  std::vector<double> out(V);
  check_status(pe_trustrank(in_edges.handle(), rev.h, good.data(), &o, out.data(), nullptr, nullptr),
               "pe_trustrank");
  return dense_state_t<double, typename INCOMING::partition_t>::from_host(std::move(out));
}

}  // namespace algo
}  // namespace plato
