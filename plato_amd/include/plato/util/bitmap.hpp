// Device-resident bitmap_t (plato/util/bitmap.hpp:43-350): same word layout,
// (size >> 6) + 1 uint64 words, bit i in word i >> 6 at position i & 63.
#pragma once

#include <utility>
#include <vector>

#include "plato/graph/base.hpp"

namespace plato {

template <typename ALLOC = void>
class bitmap_t {
 public:
  bitmap_t() = default;
  explicit bitmap_t(size_t size) : size_(size) {
    check_status(pe_bitmap_create(cluster_info_t::get_instance().ctx(), size, &h_), "pe_bitmap_create");
  }
  bitmap_t(bitmap_t&& o) noexcept : size_(o.size_), h_(o.h_) { o.h_ = nullptr; }
  bitmap_t& operator=(bitmap_t&& o) noexcept {
    if (this != &o) {
      release();
      size_ = o.size_;
      h_ = o.h_;
      o.h_ = nullptr;
    }
    return *this;
  }
  bitmap_t(const bitmap_t&) = delete;
  ~bitmap_t() { release(); }

  int copy_from(const bitmap_t& other) {
    check_status(pe_bitmap_copy_from(h_, other.h_), "pe_bitmap_copy_from");
    return 0;
  }
  void clear() { check_status(pe_bitmap_clear(h_), "pe_bitmap_clear"); }
  void fill() { check_status(pe_bitmap_fill(h_), "pe_bitmap_fill"); }
  void set_bit(size_t i) { check_status(pe_bitmap_set_bit(h_, i), "pe_bitmap_set_bit"); }
  void clr_bit(size_t i) { check_status(pe_bitmap_clr_bit(h_, i), "pe_bitmap_clr_bit"); }
  uint64_t get_bit(size_t i) const {
    int b = 0;
    check_status(pe_bitmap_get_bit(h_, i, &b), "pe_bitmap_get_bit");
    return (uint64_t)b;
  }
  size_t msb() const {
    uint64_t m = 0;
    check_status(pe_bitmap_msb(h_, &m), "pe_bitmap_msb");
    return m;
  }
  size_t count() const {
    uint64_t c = 0;
    check_status(pe_bitmap_count(h_, &c), "pe_bitmap_count");
    return c;
  }
  void sync() {}  // one GPU: nothing to allreduce (bitmap.hpp:240-243)
  std::vector<size_t> to_vector() {
    uint64_t n = 0;
    check_status(pe_bitmap_to_vector(h_, nullptr, 0, &n), "pe_bitmap_to_vector");
    std::vector<uint32_t> ids(n ? n : 1);
    check_status(pe_bitmap_to_vector(h_, ids.data(), n, &n), "pe_bitmap_to_vector");
    return std::vector<size_t>(ids.begin(), ids.begin() + n);
  }

  pe_bitmap_t* handle() const { return h_; }
  size_t size_ = 0;

 private:
  void release() {
    if (h_) pe_bitmap_destroy(h_);
    h_ = nullptr;
  }
  pe_bitmap_t* h_ = nullptr;
};

}  // namespace plato
