// Batched PPR: device-side structures shared by ppr_batch.hip and serve.cpp.
#pragma once

#include "pe_internal.hpp"

namespace pe {

// column (query slot) states
enum : uint32_t { COL_EMPTY = 0, COL_LOADING = 1, COL_ACTIVE = 2, COL_FINAL = 3, COL_DONE = 4 };

// one wave's work: columns [v0, v1) (chunk == UINT32_MAX), or chunk `chunk` of
// column v0 covering edges [e0, e1)
struct bt_task {
  uint32_t v0, v1;
  uint64_t e0, e1;
  uint32_t chunk;
  uint32_t pad;
};

// a long column split into nchunks partials starting at chunk slot c0
struct bt_long {
  uint32_t v;
  uint32_t c0;
  uint32_t nchunks;
};

struct batch_ctl {
  uint32_t cur;        // 0: state in X0, 1: in X1
  uint32_t iteration;  // passes run since creation
};

}  // namespace pe
