// plato::algo::nstepdegrees_t (plato/algo/nstepdegrees/nstepdegrees.hpp:46-489) on the
// device engine: the same class, options and result accessors.  compute() runs both
// passes on the device (pe_nstep_degrees): every vertex keeps a 2^BitWidth-register
// HyperLogLog of the vertices within `step` hops (against the edges for in-degrees,
// along them for out-degrees); a target's degree is its estimate minus one.  The
// reference restricts the propagation to the vertices within `step` hops of the
// targets (spread); propagating over every vertex gives the targets the same counters.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "plato/engine/dualmode.hpp"

namespace plato {
namespace algo {

struct nstepdegree_opts_t {
  uint32_t step = 20;   // steps; 0xFFFFFFFF (the driver's -1) until nothing grows
  std::string type = "both";
  bool is_directed = false;
};

enum class count_type_t { OUT_DEGREE = 1, IN_DEGREE = 2 };

struct degrees_record_t {
  vid_t in_;
  vid_t out_;
};

template <typename INCOMING, typename OUTGOING, uint32_t BitWidth = 12>
class nstepdegrees_t {
 public:
  using v_subset_t = bitmap_t<>;

  nstepdegrees_t(dualmode_engine_t<INCOMING, OUTGOING>* engine, const graph_info_t& graph_info,
                 const v_subset_t& active_v, const nstepdegree_opts_t& opts = nstepdegree_opts_t())
      : engine_(engine), graph_info_(graph_info), opts_(opts) {
    static_assert(BitWidth >= 6 && BitWidth <= 16, "hyperloglog bit width must be in [6, 16]");
    const uint32_t V = engine_->in_edges()->info().V;
    targets_.assign(((uint64_t)V >> 6) + 1, 0ull);
    check_status(pe_bitmap_download(active_v.handle(), targets_.data(), targets_.size()), "pe_bitmap_download");
  }

  void compute(INCOMING& in_edges, OUTGOING& /*out_edges*/) {
    const uint32_t V = in_edges.info().V;
    int type = 0;
    if (opts_.type == "in" || opts_.type == "both") type |= PE_NSTEP_IN;
    if (opts_.type == "out" || opts_.type == "both") type |= PE_NSTEP_OUT;
    if (!type) fatal("nstepdegrees_t::compute", PE_ERR_INVALID, "type must be in, out or both");
    in_.assign(V, 0);
    out_.assign(V, 0);
    check_status(pe_nstep_degrees(in_edges.handle(), targets_.data(), opts_.step, BitWidth, type,
                                  (type & PE_NSTEP_IN) ? in_.data() : nullptr,
                                  (type & PE_NSTEP_OUT) ? out_.data() : nullptr, &seconds_),
                 "pe_nstep_degrees");
  }

  // `v,in,out` per target vertex (nstepdegrees.hpp:468-489)
  template <typename STREAM>
  void save(STREAM& ss) {
    for (uint32_t v = 0; v < (uint32_t)in_.size(); ++v)
      if (is_target(v)) ss.ostream() << v << "," << in_[v] << "," << out_[v] << "\n";
  }

  void view_degrees() {
    for (uint32_t v = 0; v < (uint32_t)in_.size(); ++v)
      if (is_target(v)) std::printf("vertexId:%u---->%u,%u\n", v, in_[v], out_[v]);
  }

  vid_t get_degree(vid_t pos, bool in) { return in ? in_[pos] : out_[pos]; }
  double seconds() const { return seconds_; }

 private:
  bool is_target(uint32_t v) const { return (targets_[v >> 6] >> (v & 63)) & 1ull; }

  dualmode_engine_t<INCOMING, OUTGOING>* engine_;
  graph_info_t graph_info_;
  nstepdegree_opts_t opts_;
  std::vector<uint64_t> targets_;
  std::vector<vid_t> in_, out_;
  double seconds_ = 0.0;
};

}  // namespace algo
}  // namespace plato
