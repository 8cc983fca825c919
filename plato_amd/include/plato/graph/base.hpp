// Basic types and the one-GPU cluster context of the engine.
//
// Names follow plato/graph/base.hpp (vid_t/eid_t 58-59, empty_t, graph_info_t
// 241-257, cluster_info_t 158-233).  cluster_info_t here owns the device
// context (libplato_hip.so) instead of MPI: one process, one GPU,
// partitions_ = 1.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>

#include "plato_hip.h"

namespace plato {

using vid_t = uint32_t;
using eid_t = uint64_t;

struct empty_t {};

struct graph_info_t {
  bool is_directed_ = false;  // input params
  vid_t vertices_ = 0;        // output params
  eid_t edges_ = 0;
  vid_t max_v_i_ = 0;

  graph_info_t() = default;
  explicit graph_info_t(bool is_directed) : is_directed_(is_directed) {}
};

enum class traverse_mode_t { ORIGIN = 0, RANDOM = 1, CIRCLE = 2 };

struct traverse_opts_t {
  traverse_mode_t mode_ = traverse_mode_t::ORIGIN;
  bool auto_release_ = false;
};

// Failures abort with a message, as the reference's CHECK does.
[[noreturn]] inline void fatal(const char* what, int rc, const char* detail) {
  std::fprintf(stderr, "plato: %s failed (status %d): %s\n", what, rc, detail ? detail : "");
  std::abort();
}

class cluster_info_t {
 public:
  static cluster_info_t& get_instance() {
    static cluster_info_t instance;
    return instance;
  }

  // argc/argv accepted for source compatibility; the device is
  // $PLATO_HIP_DEVICE or 0.
  void initialize(int* = nullptr, char*** = nullptr) { (void)ctx(); }

  pe_ctx_t* ctx() {
    if (!ctx_) {
      const char* d = std::getenv("PLATO_HIP_DEVICE");
      const int rc = pe_ctx_create(d ? std::atoi(d) : 0, &ctx_);
      if (rc != PE_OK) fatal("pe_ctx_create", rc, "no usable GPU");
    }
    return ctx_;
  }

  void check(int rc, const char* what) {
    if (rc != PE_OK) fatal(what, rc, ctx_ ? pe_ctx_last_error(ctx_) : "");
  }

  int partitions_ = 1;
  int partition_id_ = 0;
  int threads_ = 1;
  int sockets_ = 1;

  ~cluster_info_t() {
    if (ctx_) pe_ctx_destroy(ctx_);
  }

 private:
  cluster_info_t() = default;
  pe_ctx_t* ctx_ = nullptr;
};

inline void check_status(int rc, const char* what) { cluster_info_t::get_instance().check(rc, what); }

}  // namespace plato
