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
class dense_state_t {
 public:
  using value_t = T;
  using partition_t = PART_IMPL;

  dense_state_t() = default;
  // (max_v_i, partitioner) as the reference's constructor (dense_state.hpp:133-137)
  template <typename PART>
  dense_state_t(vid_t max_v_i, const PART&) : n_((uint64_t)max_v_i + 1) { alloc(); }
  explicit dense_state_t(uint64_t n) : n_(n) { alloc(); }
  dense_state_t(dense_state_t&& o) noexcept : n_(o.n_), h_(o.h_), host_(std::move(o.host_)) { o.h_ = nullptr; }
  dense_state_t& operator=(dense_state_t&& o) noexcept {
    if (this != &o) {
      release();
      n_ = o.n_;
      h_ = o.h_;
      host_ = std::move(o.host_);
      o.h_ = nullptr;
    }
    return *this;
  }
  dense_state_t(const dense_state_t&) = delete;
  ~dense_state_t() { release(); }

  void fill(const T& v) { check_status(pe_vec_fill(h_, (double)v), "pe_vec_fill"); }
  void set(vid_t i, const T& v) { check_status(pe_vec_set(h_, i, (double)v), "pe_vec_set"); }
  // host view of the whole state (downloaded once, refreshed by to_host())
  const std::vector<T>& to_host() {
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
  void upload(const std::vector<T>& v) { check_status(pe_vec_upload(h_, v.data(), v.size()), "pe_vec_upload"); }
  T operator[](vid_t i) {
    if (host_.size() != n_) to_host();
    return host_[i];
  }
  // foreach (dense_state.hpp:119,422): process(v, T* value) for every vertex in
  // [0, size) -- or only those set in *pactives -- on the host; the sum of the returns
  // comes back, and values written through the pointer are uploaded to the device.
  template <typename R, typename PROCESS>
  R foreach(PROCESS&& process, bitmap_t<>* pactives = nullptr, size_t /*chunk_size*/ = 0) {
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
  double sum() const {
    double s = 0;
    check_status(pe_vec_sum(h_, &s), "pe_vec_sum");
    return s;
  }
  uint64_t size() const { return n_; }
  pe_vec_t* handle() const { return h_; }
  // drop the cached host view after device-side updates
  void invalidate() { host_.clear(); }

 private:
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
