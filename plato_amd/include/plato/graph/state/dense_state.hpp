// Device-resident dense_state_t (plato/graph/state/dense_state.hpp:43-175):
// one value per vertex id in [0, max_v_i + 1), held in HBM.  Element access
// from the host copies single values; whole-state reads use to_host().
#pragma once

#include <type_traits>
#include <vector>

#include "plato/graph/base.hpp"
#include "plato/util/bitmap.hpp"

namespace plato {

namespace detail {
template <typename T> struct pe_dtype;
template <> struct pe_dtype<float> { static constexpr int value = PE_F32; };
template <> struct pe_dtype<double> { static constexpr int value = PE_F64; };
template <> struct pe_dtype<uint32_t> { static constexpr int value = PE_U32; };
}  // namespace detail

template <typename T, typename PART_IMPL>
class dense_state_t;

// Device operators that dense_state_t::foreach runs in place of the reference's host
// lambdas (the closed set the rank algorithms use):
//   rank_init    the initialisers of pagerank.hpp:92-100 / personalized_pagerank.hpp:73-83:
//                value (on every vertex, or on the seeds), divided by the out-degree
//   rank_apply   the restart-apply of pagerank.hpp:118-140 and its relatives: for every v,
//                r = base + [v in seeds] * seed_mass + damping * (*pval); store r, or r / od
//                with delta sum |new - curt| * od (|new - curt| without out-edges)
// Both act on the graph the engine walks (dualmode_engine_t::walk_graph()).
namespace ops {
struct rank_init {
  pe_graph_t* g;
  double value = 1.0;
  const bitmap_t<>* seeds = nullptr;
  bool divide_by_od = true;
};
template <typename STATE>
struct rank_apply {
  pe_graph_t* g;
  STATE* curt;  // the state the delta is measured against
  double damping;
  double base;
  double seed_mass = 0.0;
  const bitmap_t<>* seeds = nullptr;
  bool divide_by_od = true;
  bool final_iter = false;
};
template <typename F>
struct is_rank_apply : std::false_type {};
template <typename STATE>
struct is_rank_apply<rank_apply<STATE>> : std::true_type {};
}  // namespace ops

template <typename T, typename PART_IMPL>
class dense_state_t {
 public:
  using value_t = T;
  using partition_t = PART_IMPL;


  dense_state_t() = default;
  // (max_v_i, partitioner) as the reference's constructor (dense_state.hpp:133-137)
  template <typename PART>
  dense_state_t(vid_t max_v_i, const PART&) : n_((uint64_t)max_v_i + 1) { alloc(); }
  explicit dense_state_t(uint64_t n) : n_(n) { alloc(); }
  dense_state_t(dense_state_t&& o) noexcept
      : n_(o.n_), h_(o.h_), host_(std::move(o.host_)), pending_(o.pending_) {
    o.h_ = nullptr;
    o.pending_ = pending_pull{};
  }
  dense_state_t& operator=(dense_state_t&& o) noexcept {
    if (this != &o) {
      release();
      n_ = o.n_;
      h_ = o.h_;
      host_ = std::move(o.host_);
      pending_ = o.pending_;
      o.h_ = nullptr;
      o.pending_ = pending_pull{};
    }
    return *this;
  }
  dense_state_t(const dense_state_t&) = delete;
  ~dense_state_t() { release(); }

  void fill(const T& v) {
    pending_ = pending_pull{};  // every value is overwritten
    host_.clear();
    check_status(pe_vec_fill(h_, (double)v), "pe_vec_fill");
  }
  void set(vid_t i, const T& v) {
    flush();
    check_status(pe_vec_set(h_, i, (double)v), "pe_vec_set");
    if (host_.size() == n_) host_[i] = v;
  }
  T get(vid_t i) {
    if (host_.size() != n_) to_host();
    return host_[i];
  }
  // host view of the whole state (downloaded once, refreshed by to_host())
  const std::vector<T>& to_host() {
    flush();
    host_.resize(n_);
    check_status(pe_vec_download(h_, host_.data(), n_), "pe_vec_download");
    return host_;
  }
  // a new state holding the given host values (one per vertex)
  static dense_state_t from_host(std::vector<T> v) {
    dense_state_t st((uint64_t)v.size());
    st.upload(v);
    st.host_ = std::move(v);
    return st;
  }
  void upload(const std::vector<T>& v) {
    flush();
    check_status(pe_vec_upload(h_, v.data(), v.size()), "pe_vec_upload");
  }
  // element read from the cached host view (writes: set())
  T operator[](vid_t i) { return get(i); }
  // foreach (dense_state.hpp:119,422): with a device operator (ops::rank_init,
  // ops::rank_apply) on the device; with a host function process(v, T* value) for every
  // vertex in [0, size) -- or only those set in *pactives -- on the host, values written
  // through the pointer uploaded afterwards.  Returns the sum of the returns (the delta).
  template <typename R, typename PROCESS>
  R foreach(PROCESS&& process, bitmap_t<>* pactives = nullptr, size_t /*chunk_size*/ = 0) {
    using P = typename std::decay<PROCESS>::type;
    if constexpr (std::is_same<P, ops::rank_init>::value) {
      pending_ = pending_pull{};
      host_.clear();
      check_status(pe_rank_init(process.g, h_, process.seeds ? process.seeds->handle() : nullptr, process.value,
                                process.divide_by_od ? 1 : 0),
                   "pe_rank_init");
      return (R)(double)n_;
    } else if constexpr (ops::is_rank_apply<P>::value) {
      return (R)apply(process);
    } else {
      return host_foreach<R>(process, pactives);
    }
  }
  double sum() const {
    flush();
    double s = 0;
    check_status(pe_vec_sum(h_, &s), "pe_vec_sum");
    return s;
  }
  uint64_t size() const { return n_; }
  // the device vector, with any deferred pull done first
  pe_vec_t* handle() const {
    flush();
    return h_;
  }
  // drop the cached host view after device-side updates
  void invalidate() { host_.clear(); }

  // The engine's pull (gather_sum into this state's write_add slot) is deferred until the
  // state is next used, so that a restart-apply right after it runs as one fused pass
  // (pe_rank_iterate); any other use first runs the pull alone (pe_pull_sum).
  void defer_pull(pe_graph_t* g, const dense_state_t& src, bool weighted) {
    src.flush();
    flush();
    host_.clear();
    pending_ = pending_pull{g, src.h_, weighted};
  }

 private:
  struct pending_pull {
    pe_graph_t* g = nullptr;
    pe_vec_t* src = nullptr;
    bool weighted = false;
  };

  void flush() const {
    if (!pending_.g) return;
    const pending_pull p = pending_;
    pending_ = pending_pull{};
    check_status(pe_pull_sum(p.g, p.src, h_, p.weighted ? 1 : 0), "pe_pull_sum");
  }

  template <typename R, typename PROCESS>
  R host_foreach(PROCESS& process, bitmap_t<>* pactives) {
    to_host();
    std::vector<uint64_t> words;
    if (pactives) {
      words.resize((pactives->size_ >> 6) + 1);
      check_status(pe_bitmap_download(pactives->handle(), words.data(), words.size()), "pe_bitmap_download");
    }
    R r = R();
    for (uint64_t v = 0; v < n_; ++v)
      if (!pactives || (v < pactives->size_ && ((words[v >> 6] >> (v & 63)) & 1u)))
        r += process((vid_t)v, &host_[v]);
    upload(host_);
    return r;
  }

  // ops::rank_apply on this state: fused with a deferred pull from op.curt when there is
  // one (pe_rank_iterate), else on the values as they stand (pe_rank_apply)
  template <typename OP>
  double apply(const OP& op) {
    pe_apply_params_t p{};
    p.damping = op.damping;
    p.base = op.base;
    p.seed_mass = op.seed_mass;
    p.seeds = op.seeds ? op.seeds->handle() : nullptr;
    p.divide_by_od = op.divide_by_od ? 1 : 0;
    p.final_iter = op.final_iter ? 1 : 0;
    double delta = 0.0;
    host_.clear();
    pe_vec_t* curt = op.curt->handle();
    if (pending_.g && pending_.src == curt && pending_.g == op.g) {
      const pending_pull pp = pending_;
      pending_ = pending_pull{};
      check_status(pe_rank_iterate(op.g, curt, h_, pp.weighted ? 1 : 0, &p, &delta), "pe_rank_iterate");
    } else {
      flush();
      check_status(pe_rank_apply(op.g, curt, h_, &p, &delta), "pe_rank_apply");
    }
    return op.final_iter ? 0.0 : delta;
  }

  void alloc() {
    check_status(pe_vec_create(cluster_info_t::get_instance().ctx(), detail::pe_dtype<T>::value, n_, &h_),
                 "pe_vec_create");
  }
  void release() {
    if (h_) pe_vec_destroy(h_);
    h_ = nullptr;
  }
  uint64_t n_ = 0;
  pe_vec_t* h_ = nullptr;
  std::vector<T> host_;
  mutable pending_pull pending_;
};

// Results returned by the algorithms: host values with the reference's
// accessors (a dense_state_t<double> downloaded once).
template <typename T, typename PART_IMPL>
class host_state_t {
 public:
  host_state_t() = default;
  explicit host_state_t(std::vector<T> v) : v_(std::move(v)) {}
  T& operator[](vid_t i) { return v_[i]; }
  const T& operator[](vid_t i) const { return v_[i]; }
  size_t size() const { return v_.size(); }
  const std::vector<T>& values() const { return v_; }
  // dense_state_t::foreach-style read-only traversal with a reduction
  template <typename R, typename F>
  R foreach(F&& f) const {
    R r = R();
    for (vid_t i = 0; i < (vid_t)v_.size(); ++i) r += f(i, &v_[i]);
    return r;
  }

 private:
  std::vector<T> v_;
};

}  // namespace plato
