// plato::algo::personalized_pagerank on the device engine: the pull-only form
// (plato/algo/ppr/personalized_pagerank.hpp:40-134) and the push/pull form with the sticky
// switch (137-275).  Same options, signatures and control flow: the state starts at
// 1 / od(src) on the source alone; each epoch gathers (pull) or scatters from the active
// set (push) into next_rank, then applies the restart on the source and damping, measures
// delta and takes the eps jump.  The push/pull form switches to pull for good once the
// active set's out-edges exceed E'/20 (integer division, 188-205).  The reference's
// lambdas are replaced by the engine's operators; an epoch's pull and apply run as one
// fused pass.  The result is returned in FP64.
#pragma once

#include <utility>
#include <vector>

#include "plato/algo/pagerank/pagerank.hpp"
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

namespace detail {

// PPR state of the source alone: 1 / od(src) (personalized_pagerank.hpp:73-83)
template <typename STATE, typename ENGINE>
void ppr_init(STATE& curt_rank, ENGINE& engine, const personalized_pagerank_opts_t& opts, bitmap_t<>& src_bit) {
  src_bit.set_bit(opts.src_);
  curt_rank.template foreach<double>(ops::rank_init{engine.walk_graph(), 1.0, &src_bit, true});
}

template <typename T, typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> ppr_pull(GRAPH& graph, const graph_info_t& graph_info,
                                                            const personalized_pagerank_opts_t& opts) {
  using rank_state_t = dense_state_t<T, typename GRAPH::partition_t>;
  // init pull-only engine
  dualmode_engine_t<GRAPH, std::nullptr_t> engine(std::shared_ptr<GRAPH>(&graph, [](GRAPH*) {}), nullptr,
                                                  graph_info);
  rank_state_t curt_rank = engine.template alloc_v_state<T>();
  rank_state_t next_rank = engine.template alloc_v_state<T>();
  bitmap_t<> src_bit = engine.alloc_v_subset();
  pe_graph_t* g = engine.walk_graph();
  ppr_init(curt_rank, engine, opts, src_bit);

  double delta = 0.0;
  for (uint32_t epoch_i = 0; epoch_i < opts.iteration_; ++epoch_i) {
    next_rank.fill(0.0);
    engine.template foreach_edges<T, int>(ops::gather_sum<T>{curt_rank}, ops::write_add<T>{next_rank});
    const bool final_iter = opts.iteration_ - 1 == epoch_i;
    // personalized: (v == src ? 1 - d : 0) + d * sum
    delta = next_rank.template foreach<double>(ops::rank_apply<rank_state_t>{
        g, &curt_rank, opts.damping_, 0.0, 1.0 - opts.damping_, &src_bit, true, final_iter});
    if (!final_iter && opts.eps_ > 0.0 && delta < opts.eps_) epoch_i = opts.iteration_ - 2;
    std::swap(curt_rank, next_rank);
  }
  return to_fp64(curt_rank);
}

template <typename T, typename INCOMING, typename OUTGOING>
dense_state_t<double, typename INCOMING::partition_t> ppr_pushpull(INCOMING& in_edges, OUTGOING& out_edges,
                                                                   const graph_info_t& graph_info,
                                                                   const personalized_pagerank_opts_t& opts) {
  using rank_state_t = dense_state_t<T, typename INCOMING::partition_t>;
  dualmode_engine_t<INCOMING, OUTGOING> engine(std::shared_ptr<INCOMING>(&in_edges, [](INCOMING*) {}),
                                               std::shared_ptr<OUTGOING>(&out_edges, [](OUTGOING*) {}), graph_info);
  rank_state_t curt_rank = engine.template alloc_v_state<T>();
  rank_state_t next_rank = engine.template alloc_v_state<T>();
  bitmap_t<> actives = engine.alloc_v_subset();
  bitmap_t<> src_bit = engine.alloc_v_subset();
  pe_graph_t* g = engine.walk_graph();
  eid_t edges = graph_info.edges_;
  if (false == graph_info.is_directed_) edges = edges * 2;

  ppr_init(curt_rank, engine, opts, src_bit);
  actives.set_bit(opts.src_);

  bool do_pull = false;
  double delta = 0.0;
  for (uint32_t epoch_i = 0; epoch_i < opts.iteration_; ++epoch_i) {
    if (!do_pull) do_pull = (engine.active_edges(actives) > edges / 20);
    next_rank.fill(0.0);
    if (do_pull) {  // pull
      engine.template foreach_edges<T, int>(ops::gather_sum<T>{curt_rank}, ops::write_add<T>{next_rank});
    } else {  // push: every active vertex adds its rank to its out-neighbours and activates them
      engine.template foreach_edges<T, int>(ops::push_signal<T>{curt_rank}, ops::push_add_slot<T>{next_rank},
                                            actives);
    }
    const bool final_iter = opts.iteration_ - 1 == epoch_i;
    delta = next_rank.template foreach<double>(ops::rank_apply<rank_state_t>{
        g, &curt_rank, opts.damping_, 0.0, 1.0 - opts.damping_, &src_bit, true, final_iter});
    if (!final_iter && opts.eps_ > 0.0 && delta < opts.eps_) epoch_i = opts.iteration_ - 2;
    std::swap(curt_rank, next_rank);
  }
  return to_fp64(curt_rank);
}

}  // namespace detail

template <typename GRAPH>
dense_state_t<double, typename GRAPH::partition_t> personalized_pagerank(
    GRAPH& graph, const graph_info_t& graph_info,
    const personalized_pagerank_opts_t& opts = personalized_pagerank_opts_t()) {
  if (opts.precision_ == 64) return detail::ppr_pull<double>(graph, graph_info, opts);
  if (opts.precision_ != 32) fatal("personalized_pagerank", PE_ERR_INVALID, "precision_ must be 32 or 64");
  return detail::ppr_pull<float>(graph, graph_info, opts);
}

template <typename INCOMING, typename OUTGOING>
dense_state_t<double, typename INCOMING::partition_t> personalized_pagerank(
    INCOMING& in_edges, OUTGOING& out_edges, const graph_info_t& graph_info,
    const personalized_pagerank_opts_t& opts = personalized_pagerank_opts_t()) {
  if (opts.precision_ == 64) return detail::ppr_pushpull<double>(in_edges, out_edges, graph_info, opts);
  if (opts.precision_ != 32) fatal("personalized_pagerank", PE_ERR_INVALID, "precision_ must be 32 or 64");
  return detail::ppr_pushpull<float>(in_edges, out_edges, graph_info, opts);
}

}  // namespace algo
}  // namespace plato
