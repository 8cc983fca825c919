// plato::algo::breadth_first_search (plato/algo/bfs/bfs.hpp:37-155) on the
// device engine.  The loop is the reference's: per epoch, clear the next
// frontier, one 5-argument foreach_edges (push or pull by the active-edge rule),
// swap, stop when nothing was activated; the return value is |visited|.
// The signal/slot lambdas become the engine's BFS operators; the device state
// keeps levels instead of parents (parents depend on claim order, levels do not).
#pragma once

#include <cstdio>
#include <utility>
#include <vector>

#include "plato/engine/dualmode.hpp"

namespace plato {
namespace algo {

struct bfs_opts_t {
  vid_t root_ = 0;
};

// Returns |visited| as the reference does.  When `levels` is given it receives this
// search's levels (root 0, unreached -1) in input id order; every call owns its own
// device state, so searches on different contexts or threads do not share anything.
template <typename INCOMING, typename OUTGOING>
vid_t breadth_first_search(INCOMING& in_edges, OUTGOING& out_edges, const graph_info_t& graph_info,
                           const bfs_opts_t& opts, std::vector<int32_t>* levels = nullptr,
                           bool log_epochs = false) {
  dualmode_engine_t<INCOMING, OUTGOING> engine(std::shared_ptr<INCOMING>(&in_edges, [](INCOMING*) {}),
                                               std::shared_ptr<OUTGOING>(&out_edges, [](OUTGOING*) {}),
                                               graph_info);
  vid_t actives = 1;
  auto active_current = engine.alloc_v_subset();
  auto active_next = engine.alloc_v_subset();
  ops::bfs_state state;
  check_status(pe_bfs_create(in_edges.handle(), &state.h), "pe_bfs_create");
  check_status(pe_bfs_reset(state.h, opts.root_, active_current.handle()), "pe_bfs_reset");
  vid_t visited = 1;
  for (int epoch_i = 0; 0 != actives; ++epoch_i) {
    state.next = &active_next;  // the step overwrites the next frontier (bfs.hpp:95)
    actives = engine.template foreach_edges<vid_t, vid_t>(ops::bfs_push_signal{&state}, ops::bfs_push_slot{&state},
                                                          ops::bfs_pull_signal{&state}, ops::bfs_pull_slot{&state},
                                                          active_current);
    visited += actives;
    if (log_epochs)
      std::printf("active_v[%d] = %u (%s, active edges %llu)\n", epoch_i, actives,
                  engine.last_was_push() ? "push" : "pull", (unsigned long long)engine.last_active_edges());
    std::swap(active_current, active_next);
  }
  if (levels) {
    levels->assign(in_edges.info().V, -1);
    check_status(pe_bfs_get_levels(state.h, levels->data()), "pe_bfs_get_levels");
  }
  pe_bfs_destroy(state.h);
  return visited;
}

}  // namespace algo
}  // namespace plato
