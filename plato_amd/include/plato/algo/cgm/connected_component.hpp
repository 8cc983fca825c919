// plato::algo::connected_component_t on the device engine
// (plato/algo/cgm/connected_component.hpp:38-543): compute() runs the label
// propagation on the device (pe_connected_components); the component table (vertices
// and edges per label, the major component) and the writers are host code over the
// downloaded labels, with the reference's output formats.
#pragma once

#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "plato/engine/dualmode.hpp"

namespace plato {
namespace algo {

template <typename INCOMING, typename OUTGOING>
class connected_component_t {
 public:
  struct comp_info_t {
    vid_t vertices_;
    eid_t edges_;
    comp_info_t() : vertices_(0), edges_(0) {}
  };
  using partition_t = typename INCOMING::partition_t;
  using cgm_state_t = host_state_t<vid_t, partition_t>;

  connected_component_t(dualmode_engine_t<INCOMING, OUTGOING>* engine, const graph_info_t& graph_info)
      : engine_(engine), graph_info_(graph_info) {}

  // labels, then the component table (connected_component.hpp:170-339)
  void compute() {
    pe_graph_t* g = engine_->in_edges()->handle();
    const vid_t V = engine_->in_edges()->info().V;
    std::vector<uint32_t> lab(V);
    pe_cc_stats_t st{};
    check_status(pe_connected_components(g, lab.data(), nullptr, nullptr, 0, &st), "pe_connected_components");
    labels_ = cgm_state_t(std::vector<vid_t>(lab.begin(), lab.end()));
    // od(u) = in-edge entries with source u, so a component's entries are the sum of
    // its vertices' od (the reference counts comp_edges[labels_[src]] per entry)
    std::vector<uint32_t> od(V);
    check_status(pe_graph_out_degrees(g, od.data()), "pe_graph_out_degrees");
    global_label_info_.clear();
    for (vid_t v = 0; v < V; ++v) {
      comp_info_t& c = global_label_info_[lab[v]];
      ++c.vertices_;
      c.edges_ += od[v];
    }
    vid_t major_cnt = 0;
    major_label_ = 0;
    for (auto& e : global_label_info_) {
      if (!graph_info_.is_directed_) e.second.edges_ /= 2;
      // ties go to the smaller label (the reference takes the first in hash order)
      if (e.second.vertices_ > major_cnt || (e.second.vertices_ == major_cnt && e.first < major_label_)) {
        major_cnt = e.second.vertices_;
        major_label_ = e.first;
      }
    }
  }

  int get_num_components() const { return (int)global_label_info_.size(); }
  std::unordered_map<vid_t, comp_info_t>* get_components_info() { return &global_label_info_; }
  vid_t get_major_label() const { return major_label_; }
  vid_t get_major_vertices() const { return global_label_info_.at(major_label_).vertices_; }
  cgm_state_t* get_labels() { return &labels_; }

  // components by vertices desc (label asc on ties), the reference's text
  std::string get_summary() const {
    std::vector<std::pair<vid_t, comp_info_t>> comps(global_label_info_.begin(), global_label_info_.end());
    std::sort(comps.begin(), comps.end(), [](const auto& a, const auto& b) {
      return a.second.vertices_ != b.second.vertices_ ? a.second.vertices_ > b.second.vertices_ : a.first < b.first;
    });
    std::stringstream ss;
    ss << "Connected component summary: \n";
    ss << "There are " << get_num_components() << " components.\n";
    int count = 0;
    for (const auto& c : comps)
      ss << "Component #" << ++count << ", label=" << c.first << ", vertices=" << c.second.vertices_
         << ", edges=" << c.second.edges_ << "\n";
    return ss.str();
  }

  // `v,label` for every vertex into <prefix>/part_0000_0.csv (:469-489)
  void write_all_vertices(const std::string& prefix) const {
    FILE* f = open_part(prefix);
    for (vid_t v = 0; v < (vid_t)labels_.size(); ++v) std::fprintf(f, "%u,%u\n", v, labels_[v]);
    std::fclose(f);
  }

  // `label,src,dst` per in-edge entry (undirected: src < dst only) (:491-530)
  void write_all_edges(const std::string& prefix) const {
    FILE* f = open_part(prefix);
    for_each_entry([&](vid_t src, vid_t dst) { std::fprintf(f, "%u,%u,%u\n", labels_[src], src, dst); });
    std::fclose(f);
  }

  // `src,dst` per in-edge entry of the component of target_label (-1: the major one;
  // a vertex id that is not a label stands for its component) (:392-466)
  void write_component(const std::string& prefix, int64_t target_label) const {
    vid_t label = target_label < 0 ? major_label_ : (vid_t)target_label;
    if (label >= (vid_t)labels_.size()) return;
    if (!global_label_info_.count(label)) label = labels_[label];
    FILE* f = open_part(prefix);
    for_each_entry([&](vid_t src, vid_t dst) {
      if (labels_[src] == label) std::fprintf(f, "%u,%u\n", src, dst);
    });
    std::fclose(f);
  }

 private:
  static FILE* open_part(const std::string& prefix) {
    ::mkdir(prefix.c_str(), 0755);
    const std::string path = prefix + "/part_0000_0.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) fatal("fopen", -1, path.c_str());
    return f;
  }
  template <typename F>
  void for_each_entry(F&& f) const {
    const vid_t V = engine_->in_edges()->info().V;
    std::vector<uint64_t> ptr((size_t)V + 1);
    std::vector<uint32_t> idx(std::max<uint64_t>(1, engine_->in_edges()->info().nnz));
    std::vector<uint32_t> rank_of(V), perm(V);
    pe_graph_t* h = engine_->in_edges()->handle();
    check_status(pe_graph_download_csc(h, ptr.data(), idx.data(), nullptr), "pe_graph_download_csc");
    check_status(pe_graph_permutation(h, rank_of.data()), "pe_graph_permutation");  // internal ids -> input ids
    for (vid_t i = 0; i < V; ++i) perm[rank_of[i]] = i;
    for (vid_t c = 0; c < V; ++c)
      for (uint64_t e = ptr[c]; e < ptr[c + 1]; ++e) {
        const vid_t src = perm[idx[e]], dst = perm[c];
        if (graph_info_.is_directed_ || src < dst) f(src, dst);
      }
  }

  std::unordered_map<vid_t, comp_info_t> global_label_info_;
  vid_t major_label_ = 0;
  dualmode_engine_t<INCOMING, OUTGOING>* engine_;
  graph_info_t graph_info_;
  cgm_state_t labels_;
};

}  // namespace algo
}  // namespace plato
