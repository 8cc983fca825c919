// plato::algo::hyperanf on the device engine (plato/algo/hyperanf/hyperanf.hpp:39-197):
// same options, template width and return value (the average distance estimate);
// the iteration runs on the device (pe_hyperanf) over the in-edges.
#pragma once

#include "plato/engine/dualmode.hpp"

namespace plato {
namespace algo {

struct hyperanf_opts_t {
  uint32_t iteration_ = 20;  // number of iterations
  uint32_t bits = 12;        // kept for the reference's layout; the width is BitWidth
};

template <typename INCOMING, typename OUTGOING, uint32_t BitWidth = 12>
double hyperanf(INCOMING& in_edges, OUTGOING& /*out_edges*/, const graph_info_t& /*graph_info*/,
                const hyperanf_opts_t& opts = hyperanf_opts_t()) {
  static_assert(BitWidth >= 6 && BitWidth <= 16, "hyperloglog bit width must be in [6, 16]");
  pe_hyperanf_opts_t o{};
  o.iterations = opts.iteration_;
  o.bits = BitWidth;
  double anfs[64];
  uint32_t n = 0;
  double avg = 0.0;
  check_status(pe_hyperanf(in_edges.handle(), &o, anfs, 64, &n, &avg, nullptr, nullptr), "pe_hyperanf");
  return avg;
}

}  // namespace algo
}  // namespace plato
