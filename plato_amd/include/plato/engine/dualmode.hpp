// The dual-mode engine (plato/engine/dualmode.hpp:75-322) over one device graph.
//
// The reference's foreach_edges takes host lambdas as signal/slot functors; a
// GPU cannot run them, so the four functor slots take engine operator objects
// from a closed set (plato::ops).  Each combination the reference's algorithms
// use maps onto one C-ABI call:
//   pull  gather_sum + write_add    aggregate_message with the PageRank
//                                   signal/slot (pagerank.hpp:104-116) -> pe_pull_sum
//   pull  bfs_pull_signal/slot      bfs.hpp:119-136                 -> pe_bfs_step(pull)
//   push  bfs_push_signal/slot      bfs.hpp:98-118                  -> pe_bfs_step(push)
// The 5-argument form keeps the reference's rule: A = sum of out-degrees over the
// active set, push iff (double)A / E' < push_threshold_ (dualmode.hpp:261-279).
// After reverse() pulls sum over out-edges and pushes follow in-edges
// (dualmode.hpp:299-322): a directed graph's reverse is built on the device once, over
// the same ids (pe_graph_create_reverse with PE_GRAPH_SAME_IDS), and every operator
// then walks it; the push/pull choice keeps the graph's own out-degrees, as the
// reference's does (dualmode.hpp:250-278).
// Any other functor type fails to compile with a message: there is no host
// fallback.
#pragma once

#include <memory>
#include <type_traits>
#include <utility>

#include "plato/graph/base.hpp"
#include "plato/graph/state/dense_state.hpp"
#include "plato/graph/structure.hpp"
#include "plato/util/bitmap.hpp"

namespace plato {

struct dualmode_engine_opts_t {
  double push_threshold_ = 0.05;  // active_edges < 1/20 of the edges: push mode
};

namespace ops {

template <typename T>
struct gather_sum {
  dense_state_t<T, sequence_balanced_by_destination_t>& src;
  bool weighted = false;
};
template <typename T>
struct write_add {
  dense_state_t<T, sequence_balanced_by_destination_t>& dst;
};

// Push operators of push/pull PPR (personalized_pagerank.hpp:222-239): every active v
// sends src[v]; the slot adds it to dst[w] for each out-neighbour w and activates w (the
// active set grows after the step).
template <typename T>
struct push_signal {
  dense_state_t<T, sequence_balanced_by_destination_t>& src;
};
template <typename T>
struct push_add_slot {
  dense_state_t<T, sequence_balanced_by_destination_t>& dst;
};

// BFS operators share one device BFS state; `next` receives the next frontier.
struct bfs_state {
  pe_bfs_t* h = nullptr;
  bitmap_t<>* next = nullptr;
  eid_t next_active_edges = 0;
  pe_graph_t* walked = nullptr;  // graph the search last stepped over
};
struct bfs_push_signal { bfs_state* s; };
struct bfs_push_slot { bfs_state* s; };
struct bfs_pull_signal { bfs_state* s; };
struct bfs_pull_slot { bfs_state* s; };

template <typename...>
struct unsupported : std::false_type {};

}  // namespace ops

template <typename INCOMING, typename OUTGOING>
class dualmode_engine_t {
 public:
  using incoming_graph_t = INCOMING;
  using outgoing_graph_t = OUTGOING;
  using partition_t = sequence_balanced_by_destination_t;
  template <typename T>
  using v_state_t = dense_state_t<T, partition_t>;
  using v_subset_t = bitmap_t<>;

  dualmode_engine_t(std::shared_ptr<INCOMING> in_edges, std::shared_ptr<OUTGOING> out_edges,
                    const graph_info_t& graph_info, const dualmode_engine_opts_t opts = dualmode_engine_opts_t())
      : in_edges_(std::move(in_edges)), out_edges_(std::move(out_edges)), graph_info_(graph_info), opts_(opts) {}

  dualmode_engine_t(dualmode_engine_t&&) = default;
  dualmode_engine_t& operator=(dualmode_engine_t&&) = default;
  dualmode_engine_t(const dualmode_engine_t&) = delete;
  dualmode_engine_t& operator=(const dualmode_engine_t&) = delete;

  template <typename T>
  v_state_t<T> alloc_v_state() {
    return v_state_t<T>(graph_info_.max_v_i_, partition_t{});
  }
  v_subset_t alloc_v_subset() { return v_subset_t((size_t)device_graph_handle_info().V); }

  // 5-argument form: push or pull by the active-edge rule
  template <typename MSG, typename R, typename PUSH_SIGNAL, typename PUSH_SLOT, typename PULL_SIGNAL,
            typename PULL_SLOT>
  R foreach_edges(PUSH_SIGNAL&& push_signal, PUSH_SLOT&& push_slot, PULL_SIGNAL&& pull_signal,
                  PULL_SLOT&& pull_slot, v_subset_t& actives) {
    int is_push = 0;
    uint64_t A = 0;
    check_status(pe_choose_mode(graph(), actives.handle(), opts_.push_threshold_, &is_push, &A), "pe_choose_mode");
    last_active_edges_ = A;
    last_was_push_ = is_push != 0;
    if (is_push)
      return foreach_edges<MSG, R>(std::forward<PUSH_SIGNAL>(push_signal), std::forward<PUSH_SLOT>(push_slot), actives);
    // the pull operators of BFS read the active set of this call (bfs.hpp:123)
    current_frontier_ = &actives;
    const R r = foreach_edges<MSG, R>(std::forward<PULL_SIGNAL>(pull_signal), std::forward<PULL_SLOT>(pull_slot));
    current_frontier_ = nullptr;
    return r;
  }

  // push only
  template <typename MSG, typename R, typename SIGNAL, typename SLOT>
  R foreach_edges(SIGNAL&& signal, SLOT&& slot, v_subset_t& actives) {
    return push_dispatch<R>(signal, slot, actives);
  }

  // pull only
  template <typename MSG, typename R, typename SIGNAL, typename SLOT>
  R foreach_edges(SIGNAL&& signal, SLOT&& slot) {
    return pull_dispatch<R>(signal, slot);
  }

  std::shared_ptr<INCOMING> in_edges() { return in_edges_; }
  std::shared_ptr<OUTGOING> out_edges() { return out_edges_; }
  void reverse() { reversed_ = !reversed_; }
  bool is_reversed() { return reversed_; }

  // the engine's out-degrees (generate_dense_out_degrees_fg, kept by the engine)
  eid_t active_edges(v_subset_t& actives) {
    uint64_t a = 0;
    check_status(pe_active_edges(graph(), actives.handle(), &a), "pe_active_edges");
    return (eid_t)a;
  }

  // the graph whose in-edges a pull and whose out-edges a push walk
  pe_graph_t* walk_graph() {
    if (!reversed_) return graph();
    pe_graph_info_t gi = device_graph_handle_info();
    if (!gi.is_directed) return graph();  // a symmetrized graph is its own reverse
    if (!rev_) {
      pe_graph_t* r = nullptr;
      check_status(pe_graph_create_reverse(graph(), PE_GRAPH_BUILD_CSR | PE_GRAPH_SAME_IDS, &r),
                   "pe_graph_create_reverse");
      rev_.reset(r);
    }
    return rev_.get();
  }

  // decision of the last 5-argument call (the reference logs it under __DUALMODE_DEBUG__)
  bool last_was_push() const { return last_was_push_; }
  eid_t last_active_edges() const { return last_active_edges_; }

 private:
  pe_graph_t* graph() const {
    if (in_edges_) return in_graph(in_edges_);
    return out_graph(out_edges_);
  }
  template <typename G>
  static pe_graph_t* in_graph(const std::shared_ptr<G>& g) { return g->handle(); }
  static pe_graph_t* in_graph(const std::shared_ptr<std::nullptr_t>&) { return nullptr; }
  template <typename G>
  static pe_graph_t* out_graph(const std::shared_ptr<G>& g) { return g->handle(); }
  static pe_graph_t* out_graph(const std::shared_ptr<std::nullptr_t>&) { return nullptr; }
  pe_graph_info_t device_graph_handle_info() const {
    pe_graph_info_t gi{};
    check_status(pe_graph_get_info(graph(), &gi), "pe_graph_get_info");
    return gi;
  }

  template <typename R, typename T>
  R pull_dispatch(ops::gather_sum<T>& sig, ops::write_add<T>& slot) {
    // runs when dst is next used: fused with a restart-apply that follows (pe_rank_iterate),
    // else alone (pe_pull_sum)
    slot.dst.defer_pull(walk_graph(), sig.src, sig.weighted);
    return R();  // the PageRank slot returns 0 for every message
  }
  template <typename R>
  R pull_dispatch(ops::bfs_pull_signal& sig, ops::bfs_pull_slot&) {
    return bfs_step<R>(sig.s, 0, current_frontier_);
  }
  template <typename R, typename A, typename B>
  R pull_dispatch(A&, B&) {
    static_assert(ops::unsupported<A, B>::value,
                  "foreach_edges (pull): use plato::ops::gather_sum/write_add or bfs_pull_signal/bfs_pull_slot");
    return R();
  }
  template <typename R>
  R push_dispatch(ops::bfs_push_signal& sig, ops::bfs_push_slot&, v_subset_t& actives) {
    return bfs_step<R>(sig.s, 1, &actives);
  }
  template <typename R, typename T>
  R push_dispatch(ops::push_signal<T>& sig, ops::push_add_slot<T>& slot, v_subset_t& actives) {
    check_status(pe_push_add(walk_graph(), actives.handle(), sig.src.handle(), slot.dst.handle()), "pe_push_add");
    slot.dst.invalidate();
    return R();
  }
  template <typename R, typename A, typename B>
  R push_dispatch(A&, B&, v_subset_t&) {
    static_assert(ops::unsupported<A, B>::value, "foreach_edges (push): use plato::ops::bfs_push_signal/bfs_push_slot");
    return R();
  }
  template <typename R>
  R bfs_step(ops::bfs_state* s, int push, v_subset_t* current) {
    if (!current) fatal("foreach_edges", PE_ERR_STATE, "BFS pull needs the active set of the 5-argument call");
    pe_graph_t* wg = walk_graph();
    if (s->walked != wg) {
      check_status(pe_bfs_bind(s->h, wg), "pe_bfs_bind");
      s->walked = wg;
    }
    uint64_t activated = 0, next_a = 0;
    check_status(pe_bfs_step(s->h, push, current->handle(), s->next->handle(), &activated, &next_a), "pe_bfs_step");
    s->next_active_edges = next_a;
    return (R)activated;
  }

 private:
  struct graph_deleter {
    void operator()(pe_graph_t* g) const { pe_graph_destroy(g); }
  };

  std::shared_ptr<INCOMING> in_edges_;
  std::shared_ptr<OUTGOING> out_edges_;
  bool reversed_ = false;
  std::unique_ptr<pe_graph_t, graph_deleter> rev_;  // reverse of a directed graph, same ids
  graph_info_t graph_info_;
  dualmode_engine_opts_t opts_;
  v_subset_t* current_frontier_ = nullptr;
  bool last_was_push_ = false;
  eid_t last_active_edges_ = 0;
};

}  // namespace plato
