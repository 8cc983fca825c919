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
  // the pass plan (k_batch_plan): the pass kernel to run (1, 2, 4: edges per row load
  // over the wide block; BATCH_MODE_PACKED: the packed block), and the columns it writes
  uint32_t mode;
  uint32_t ncols;
  uint8_t cols[64];   // covered columns, ascending
  uint8_t qof[64];    // position of a column in cols (0xFF: not covered)
  uint32_t pend[64];  // 1: the column turned DONE last pass and is not yet in both buffers
  // packed block: its column list as of the last packed pass (pk_n = 0xFFFFFFFF: the last
  // pass was not packed) and whether this pass may read it
  uint32_t pk_n;
  uint32_t pk_valid;
  uint8_t pk_cols[64];
  // the plan's limits (pe_ppr_batch_opts_t, resolved): largest covered count per kernel, 0 = off
  uint32_t max_packed, max_edges4, max_edges2;
};

static constexpr uint32_t BATCH_MODE_PACKED = 5u;
static constexpr int BATCH_PACKED_W = 8;  // columns in a packed row (32 B)

}  // namespace pe
