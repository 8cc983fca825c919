// plato::algo::kcore_algo_t on the device engine (plato/algo/kcore/kcore.hpp:50-268):
// compute_shell_index returns each vertex's coreness, computed by pe_kcore with the
// reference's synchronous h-index iteration over the structure's adjacency lists
// (in-edges for a DCSC or an incoming BCSR, out-edges for an outgoing BCSR).
#pragma once

#include <vector>

#include "plato/engine/dualmode.hpp"

namespace plato {
namespace algo {

struct kcore_info_t {
  vid_t v_count_;
  eid_t e_count_;
};

enum class kcore_calc_type_t { SUBGRAPH = 1, VERTEX = 2 };

class kcore_algo_t {
 public:
  // The callback is part of the reference's signature; compute_shell_index there does
  // not call it either (its drivers use it to write results).
  template <typename Graph, typename Callback>
  static host_state_t<vid_t, typename Graph::partition_t> compute_shell_index(const graph_info_t& graph_info,
                                                                            Graph& incomings,
                                                                            Callback&& /*callback*/) {
    std::vector<uint32_t> core(incomings.info().V);
    pe_kcore_stats_t st{};
    check_status(pe_kcore(incomings.handle(), incomings.outgoing() ? 1 : 0, core.data(), nullptr, nullptr, 0, &st),
                 "pe_kcore");
    (void)graph_info;
    return host_state_t<vid_t, typename Graph::partition_t>(std::vector<vid_t>(core.begin(), core.end()));
  }
};

}  // namespace algo
}  // namespace plato
