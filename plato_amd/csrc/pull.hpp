// Pull-mode rank iteration: shared declarations (kernels in pull.hip).
#pragma once

#include "pe_internal.hpp"

namespace pe {

enum : int32_t { MODE_APPLY = 0, MODE_RAW_SUM = 1 };

// Restart-apply parameters: r(v) = base + [v in seeds] * seed_mass + d * sum(v).
struct rank_args {
  double damping;
  double base;
  double seed_mass;
  const uint64_t* seeds;  // device bitmap words or nullptr
  int32_t divide_by_od;   // store r/od (PageRank, PPR) or r (weighted PageRank)
  int32_t final_iter;     // used when no controller is attached
  int32_t mode;           // MODE_APPLY or MODE_RAW_SUM (gather only)
};

// Device-resident loop controller: lets iterations be enqueued (or replayed
// from a captured graph) without host round-trips.  Kernels return at once
// when done is set; the finish kernel advances t and flips cur.
struct iter_ctl {
  uint32_t iters;       // pagerank_opts_t::iteration_
  uint32_t t;           // the reference's epoch_i
  uint32_t executed;    // iterations run
  uint32_t done;
  uint32_t cur;         // 0: state in buf0, 1: state in buf1
  uint32_t cap;         // capacity of the per-iteration delta log
  double eps;           // pagerank_opts_t::eps_
  double l1tol;         // normalised tolerance on delta / sum(rank)
  double last_delta;
};

struct rank_work {
  unsigned tile_grid = 0, fix_grid = 0;
  dbuf<uint8_t> tile_carry, tile_head;  // compensated pair per tile
  dbuf<double> part_delta, part_rsum;   // per-block partials
  dbuf<double> scalars;                 // delta, rank sum (engine-level calls)
  void init(pe_graph_s* g, int precision);
};

// One pull iteration (tiles + fix-up + finish) on the context stream.
// With ctl: buffers picked by ctl->cur, deltas/rsums are the per-iteration logs.
// Without ctl: x = b0, y = b1, a.final_iter decides, deltas[0]/rsums[0] get the result.
void rank_iteration(pe_graph_s* g, int precision, void* b0, void* b1, bool weighted,
                    const rank_args& a, iter_ctl* ctl, rank_work& wk, double* deltas,
                    double* rsums);
void rank_apply_only(pe_graph_s* g, int precision, const void* x, void* y, const rank_args& a,
                     rank_work& wk);

}  // namespace pe
