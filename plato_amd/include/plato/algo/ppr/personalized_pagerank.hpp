// plato::algo::personalized_pagerank on the device engine: the pull-only form
// (plato/algo/ppr/personalized_pagerank.hpp:40-134) and the push/pull form
// with the sticky switch (137-275).  Same options and result shape.
#pragma once

#include <vector>

#include "plato/engine/dualmode.hpp"

namespace plato {
namespace algo {

struct personalized_pagerank_opts_t {
  uint32_t iteration_ = 100;  // number of iterations
  double damping_ = 0.85;     // the damping factor
  vid_t src_ = 0;             // the source vertex
  double eps_ = 0.001;        // stop when delta < eps (then one final iteration)
  int precision_ = 32;
};

template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> personalized_pagerank(
    GRAPH& graph, const graph_info_t& /*graph_info*/,
    const personalized_pagerank_opts_t& opts = personalized_pagerank_opts_t()) {
  pe_rank_opts_t o{};
  o.iterations = opts.iteration_;
  o.damping = opts.damping_;
  o.eps = opts.eps_;
  o.precision = opts.precision_;
  std::vector<double> out(graph.info().V);
  const uint32_t src = opts.src_;
  check_status(pe_ppr(graph.handle(), &src, 1, &o, out.data(), nullptr, nullptr), "pe_ppr");
  return dense_state_t<double, typename GRAPH::partition_t>::from_host(std::move(out));
}

template <typename INCOMING, typename OUTGOING>
dense_state_t<double, typename INCOMING::partition_t> personalized_pagerank(
    INCOMING& in_edges, OUTGOING& /*out_edges*/, const graph_info_t& /*graph_info*/,
    const personalized_pagerank_opts_t& opts = personalized_pagerank_opts_t()) {
  pe_rank_opts_t o{};
  o.iterations = opts.iteration_;
  o.damping = opts.damping_;
  o.eps = opts.eps_;
  o.precision = opts.precision_;
  std::vector<double> out(in_edges.info().V);
  check_status(pe_ppr_pushpull(in_edges.handle(), opts.src_, &o, out.data(), nullptr, nullptr, nullptr),
               "pe_ppr_pushpull");
  return dense_state_t<double, typename INCOMING::partition_t>::from_host(std::move(out));
}

}  // namespace algo
}  // namespace plato
