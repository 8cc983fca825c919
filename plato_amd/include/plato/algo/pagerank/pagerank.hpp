// plato::algo::pagerank and weighted_pagerank on the device engine.
//
// Same names, options, signatures and control flow as plato/algo/pagerank/pagerank.hpp:60-149
// and weighted_pagerank.hpp:60-157 (one pagerank_opts_t for both, avoiding the reference's
// duplicate definition): a pull-only engine over the graph, the state initialised to
// 1 / out-degree, then per epoch a pull-only foreach_edges that gathers curt_rank into
// next_rank, and a foreach that applies restart and damping, measures delta and takes the
// eps jump to one final iteration.  The reference's lambdas are replaced by the engine's
// operators (ops::gather_sum / write_add / rank_init / rank_apply); the pull and the apply
// of an epoch run on the device as one fused pass.  The state is FP32 (compensated or
// exact sums) or FP64 by precision_; the result is returned in FP64.
#pragma once

#include <utility>
#include <vector>

#include "plato/engine/dualmode.hpp"
#include "plato/graph/structure.hpp"

namespace plato {
namespace algo {

struct pagerank_opts_t {
  uint32_t iteration_ = 100;  // number of iterations
  double damping_ = 0.85;     // the damping factor
  double eps_ = 0.001;        // stop when delta < eps (then one final iteration); 0: run all
  int precision_ = 32;        // device state: 32 (FP32, compensated sums) or 64
};

namespace detail {

// FP64 copy of a state (the reference's rank_state_t is dense_state_t<double>)
template <typename T, typename PART>
dense_state_t<double, PART> to_fp64(dense_state_t<T, PART>& st) {
  if constexpr (std::is_same<T, double>::value) {
    return std::move(st);
  } else {
    const std::vector<T>& h = st.to_host();
    return dense_state_t<double, PART>::from_host(std::vector<double>(h.begin(), h.end()));
  }
}

template <typename T, typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> pagerank_loop(GRAPH& graph, const graph_info_t& graph_info,
                                                                 const pagerank_opts_t& opts, bool weighted) {
  using rank_state_t = dense_state_t<T, typename GRAPH::partition_t>;
  // init pull-only engine
  dualmode_engine_t<GRAPH, std::nullptr_t> engine(std::shared_ptr<GRAPH>(&graph, [](GRAPH*) {}), nullptr,
                                                  graph_info);
  rank_state_t curt_rank = engine.template alloc_v_state<T>();
  rank_state_t next_rank = engine.template alloc_v_state<T>();
  pe_graph_t* g = engine.walk_graph();

  // 1 / od (pagerank.hpp:92-100); the weighted variant keeps 1 (weighted_pagerank.hpp:100-105)
  double delta = curt_rank.template foreach<double>(ops::rank_init{g, 1.0, nullptr, !weighted});

  for (uint32_t epoch_i = 0; epoch_i < opts.iteration_; ++epoch_i) {
    next_rank.fill(0.0);
    engine.template foreach_edges<T, int>(ops::gather_sum<T>{curt_rank, weighted}, ops::write_add<T>{next_rank});

    if (opts.iteration_ - 1 == epoch_i) {
      delta = next_rank.template foreach<double>(
          ops::rank_apply<rank_state_t>{g, &curt_rank, opts.damping_, 1.0 - opts.damping_, 0.0, nullptr, !weighted,
                                        true});
    } else {
      delta = next_rank.template foreach<double>(
          ops::rank_apply<rank_state_t>{g, &curt_rank, opts.damping_, 1.0 - opts.damping_, 0.0, nullptr, !weighted,
                                        false});
      if (opts.eps_ > 0.0 && delta < opts.eps_) {
        epoch_i = opts.iteration_ - 2;
      }
    }
    std::swap(curt_rank, next_rank);
  }
  return to_fp64(curt_rank);
}

template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> run_pagerank(GRAPH& graph, const graph_info_t& graph_info,
                                                               const pagerank_opts_t& opts, bool weighted) {
  if (opts.precision_ == 64) return pagerank_loop<double>(graph, graph_info, opts, weighted);
  if (opts.precision_ != 32) fatal("pagerank", PE_ERR_INVALID, "precision_ must be 32 or 64");
  return pagerank_loop<float>(graph, graph_info, opts, weighted);
}

}  // namespace detail

template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> pagerank(GRAPH& graph, const graph_info_t& graph_info,
                                                           const pagerank_opts_t& opts = pagerank_opts_t()) {
  return detail::run_pagerank(graph, graph_info, opts, false);
}

template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> weighted_pagerank(GRAPH& graph, const graph_info_t& graph_info,
                                                                    const pagerank_opts_t& opts = pagerank_opts_t()) {
  return detail::run_pagerank(graph, graph_info, opts, true);
}

}  // namespace algo
}  // namespace plato
