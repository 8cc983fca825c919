// plato::algo::trustrank on the device engine (plato/algo/trustrank/trustrank.hpp:23-259):
// generate_rank_pairs, select_good_seeds and trustrank with the reference's options and
// result shapes.  Seed ranking runs on the device (pe_trustrank_select; inverse PageRank
// pulls over a reversed copy of the graph built for the call), and trustrank follows the
// reference's loop over the engine's operators.
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

namespace detail {
template <typename T, typename INCOMING, typename OUTGOING>
dense_state_t<double, typename INCOMING::partition_t> trustrank_loop(INCOMING& in_edges, OUTGOING& out_edges,
                                                                     const graph_info_t& graph_info,
                                                                     const bitmap_t<>& good_vertices,
                                                                     const trustrank_opts_t& opts) {
  using rank_state_t = dense_state_t<T, typename INCOMING::partition_t>;
  dualmode_engine_t<INCOMING, OUTGOING> engine(std::shared_ptr<INCOMING>(&in_edges, [](INCOMING*) {}),
                                               std::shared_ptr<OUTGOING>(&out_edges, [](OUTGOING*) {}), graph_info);
  pe_graph_t* g = engine.walk_graph();
  // select good seeds by opts.select_method
  bitmap_t<> good_seeds = select_good_seeds(in_edges, out_edges, graph_info, good_vertices, opts);
  if (good_seeds.count() == 0) fatal("trustrank", PE_ERR_STATE, "no good seeds were selected");

  rank_state_t curt_rank = engine.template alloc_v_state<T>();
  rank_state_t next_rank = engine.template alloc_v_state<T>();
  const double init_dist = 1. / good_seeds.count();
  // [v in seeds] * init_dist / od (trustrank.hpp:197-206)
  double delta = curt_rank.template foreach<double>(ops::rank_init{g, init_dist, &good_seeds, true});

  for (uint32_t epoch_i = 0; epoch_i < opts.iteration_; ++epoch_i) {
    next_rank.fill(0.0);
    engine.template foreach_edges<T, int>(ops::gather_sum<T>{curt_rank}, ops::write_add<T>{next_rank});
    const bool final_iter = opts.iteration_ - 1 == epoch_i;
    // use good seeds to compute: ([v in seeds] ? (1 - d) * init_dist : 0) + d * sum
    delta = next_rank.template foreach<double>(ops::rank_apply<rank_state_t>{
        g, &curt_rank, opts.damping_, 0.0, (1.0 - opts.damping_) * init_dist, &good_seeds, true, final_iter});
    if (!final_iter && opts.eps_ > 0.0 && delta < opts.eps_) epoch_i = opts.iteration_ - 2;
    std::swap(curt_rank, next_rank);
  }
  return to_fp64(curt_rank);
}
}  // namespace detail

// trustrank (trustrank.hpp:163-259): each vertex's rank, dense.  The reference's control
// flow over the engine's operators: seeds chosen by select_good_seeds, the state started
// on them, then per epoch a pull-only foreach_edges and a restart-apply foreach on the
// seed set, with the eps jump.
template <typename INCOMING, typename OUTGOING>
dense_state_t<double, typename INCOMING::partition_t> trustrank(INCOMING& in_edges, OUTGOING& out_edges,
                                                             const graph_info_t& graph_info,
                                                             const bitmap_t<>& good_vertices,
                                                             const trustrank_opts_t& opts = trustrank_opts_t()) {
  if (opts.precision_ == 32) return detail::trustrank_loop<float>(in_edges, out_edges, graph_info, good_vertices, opts);
  if (opts.precision_ != 64 && opts.precision_ != 0)
    fatal("trustrank", PE_ERR_INVALID, "precision_ must be 32 or 64");
  return detail::trustrank_loop<double>(in_edges, out_edges, graph_info, good_vertices, opts);
}

}  // namespace algo
}  // namespace plato
