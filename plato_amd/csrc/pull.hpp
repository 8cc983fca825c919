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
  double mass;            // bound on the total of all contributions sum_v s(v) (the rank
                          // mass that flows along edges); sets the binned gather's
                          // fixed-point scale
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

// ---- apply --------------------------------------------------------------------------
__device__ __forceinline__ bool seed_bit(const uint64_t* seeds, uint32_t v) {
  return seeds && ((seeds[v >> 6] >> (v & 63)) & 1ull);
}

// Finishes row v from its gathered sum; returns its delta contribution and adds
// the (undivided) rank to rsum.
template <typename T>
__device__ __forceinline__ double apply_row(const rank_args& a, bool final_iter, uint32_t v,
                                            double s, const T* __restrict__ x,
                                            T* __restrict__ y, const uint32_t* __restrict__ od,
                                            double& rsum) {
  if (a.mode == MODE_RAW_SUM) {
    y[v] = (T)s;
    return 0.0;
  }
  const double restart = a.base + (seed_bit(a.seeds, v) ? a.seed_mass : 0.0);
  const double r = restart + a.damping * s;
  rsum += r;
  if (final_iter) {
    y[v] = (T)r;
    return 0.0;
  }
  if (a.divide_by_od) {
    const uint32_t d = od[v];
    if (d > 0) {
      const T xn = (T)(r / (double)d);
      y[v] = xn;
      return fabs((double)xn - (double)x[v]) * (double)d;
    }
  }
  const T xn = (T)r;
  y[v] = xn;
  return fabs((double)xn - (double)x[v]);
}

// block reduce of two doubles; result valid in thread 0
template <int BLOCK>
__device__ __forceinline__ void block_reduce2(double& a, double& b, double* sm) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) {
    sm[wave] = a;
    sm[BLOCK / 64 + wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tb = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) {
      ta += sm[w];
      tb += sm[BLOCK / 64 + w];
    }
    a = ta;
    b = tb;
  }
}

struct rank_work {
  unsigned tile_grid = 0, fix_grid = 0;
  unsigned pb_parts = 0;                     // partials of a binned iteration (0: not binned)
  uint8_t* tile_carry = nullptr;             // compensated pair per tile
  uint8_t* tile_head = nullptr;
  double* part_delta = nullptr;              // per-block partials
  double* part_rsum = nullptr;
  double* scalars = nullptr;                 // delta, rank sum (engine-level calls)
  bool timed = true;                         // per-kernel event timing allowed
  dbuf<uint8_t> own_tiles;                   // storage when init() allocates it
  dbuf<double> own_parts;
  void init(pe_graph_s* g, int precision);
  // The same buffers carved from caller-provided device memory (cap bytes, 256-byte
  // aligned) with no allocation and no kernel timing; fails with PE_ERR_RANGE when the
  // graph needs more than cap.
  void init_in(pe_graph_s* g, int precision, uint8_t* arena, uint64_t cap);
  // bytes init_in needs for this graph and precision
  static uint64_t arena_bytes(pe_graph_s* g, int precision);
};

// Address of this GPU's static scratch array of pe_pull_sum_noalloc (looked up once per
// device, with the device current, by pe_ctx_create).
void* pull_scratch_setup(int device);

// One pull iteration (tiles + fix-up + finish) on the context stream.
// With ctl: buffers picked by ctl->cur, deltas/rsums are the per-iteration logs.
// Without ctl: x = b0, y = b1, a.final_iter decides, deltas[0]/rsums[0] get the result.
void rank_iteration(pe_graph_s* g, int precision, void* b0, void* b1, bool weighted,
                    const rank_args& a, iter_ctl* ctl, rank_work& wk, double* deltas,
                    double* rsums);
// Binned pull (pull_binned.hip): grids, and one FP32 iteration writing its
// per-block partials (gather grid, then split-apply grid).
unsigned binned_grid_gather(pe_graph_s* g);
unsigned binned_grid_split(pe_graph_s* g);
void binned_iteration(pe_graph_s* g, float* b0, float* b1, bool weighted, const rank_args& a, iter_ctl* ctl,
                      double* part_delta, double* part_rsum);
// Bound on the total contribution of state x (rank_args::mass) for a binned iteration;
// 0 when x has a negative or non-finite source value (such states take the merge-path
// gather).  Synchronizes the stream.
double binned_mass_bound(pe_graph_s* g, const float* x, bool weighted);
void rank_apply_only(pe_graph_s* g, int precision, const void* x, void* y, const rank_args& a,
                     rank_work& wk);

}  // namespace pe
