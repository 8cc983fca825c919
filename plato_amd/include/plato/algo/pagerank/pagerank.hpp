// plato::algo::pagerank and weighted_pagerank on the device engine.
//
// Same names, options and return shape as plato/algo/pagerank/pagerank.hpp:39-149
// and weighted_pagerank.hpp:60-157 (one pagerank_opts_t for both, avoiding the
// reference's duplicate definition).  The whole loop of the reference -- init,
// pull aggregate + apply per epoch, the eps jump to one final iteration -- runs
// on the device through pe_pagerank; the result is the undivided rank in FP64.
#pragma once

#include <vector>

#include "plato/engine/dualmode.hpp"

namespace plato {
namespace algo {

struct pagerank_opts_t {
  uint32_t iteration_ = 100;  // number of iterations
  double damping_ = 0.85;     // the damping factor
  double eps_ = 0.001;        // stop when delta < eps (then one final iteration); 0: run all
  int precision_ = 32;        // device state: 32 (FP32, compensated sums) or 64
};

namespace detail {
template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> run_pagerank(GRAPH& graph, const pagerank_opts_t& opts,
                                                               bool weighted) {
  pe_rank_opts_t o{};
  o.iterations = opts.iteration_;
  o.damping = opts.damping_;
  o.eps = opts.eps_;
  o.precision = opts.precision_;
  o.weighted = weighted ? 1 : 0;
  std::vector<double> out(graph.info().V);
  pe_rank_stats_t st{};
  check_status(pe_pagerank(graph.handle(), &o, out.data(), nullptr, &st), "pe_pagerank");
  return dense_state_t<double, typename GRAPH::partition_t>::from_host(std::move(out));
}
}  // namespace detail

template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> pagerank(GRAPH& graph, const graph_info_t& /*graph_info*/,
                                                           const pagerank_opts_t& opts = pagerank_opts_t()) {
  return detail::run_pagerank(graph, opts, false);
}

template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> weighted_pagerank(GRAPH& graph, const graph_info_t& /*graph_info*/,
                                                                    const pagerank_opts_t& opts = pagerank_opts_t()) {
  return detail::run_pagerank(graph, opts, true);
}

}  // namespace algo
}  // namespace plato
