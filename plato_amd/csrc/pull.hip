// Pull mode: CSC gather-sum with the rank apply fused into the same pass.
//
// Replaces aggregate_message (plato/graph/message_passing.hpp:88-144) driven by
// the gather-sum signal / write_add slot of pagerank.hpp:104-116, and the
// dense_state_t::foreach apply of pagerank.hpp:118-140 (and its variants in
// weighted_pagerank.hpp, personalized_pagerank.hpp, trustrank.hpp).
//
// Work decomposition: merge-path tiles over the sequence of (edge, row-end)
// items (Merrill & Garland, SC16).  Each tile holds a fixed number of items,
// so every block streams the same number of row_idx entries whatever the
// degree skew, and long runs of empty columns cost one item each.
//   1. stage: the block reads its row_idx slice coalesced, gathers x[u]
//      (times the normalised weight) into LDS, and stages the row ends;
//   2. walk: each thread consumes ITEMS items of the merge path from LDS,
//      summing edges and closing rows;
//   3. rows that span threads are combined by a segmented scan of
//      compensated (TwoSum) pairs across the block; rows that span tiles leave a
//      partial per tile and are finished by the fix-up kernel, in tile order;
//   4. the apply (restart + damping, divide by out-degree, delta) runs in FP64
//      on the finished sum and stores the new FP32 state.
// Per-column order is fixed by the tiling, so results are bit-reproducible.
#include <algorithm>

#include <mutex>

#include "pe_internal.hpp"
#include "pull.hpp"

#ifndef PE_PULL_NT
#define PE_PULL_NT 0
#endif
#ifndef PE_GATHER_NT
#define PE_GATHER_NT 0
#endif

namespace pe {

// Streamed structure loads (row_idx, col_ptr) may bypass cache allocation so
// the gathered vector keeps the caches.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
#if PE_PULL_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

template <typename T>
__device__ __forceinline__ T ld_gather(const T* p) {
#if PE_GATHER_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

// ---- compensated pairs ----------------------------------------------------------
template <typename T>
struct cpair {
  T hi, lo;
};

template <typename T>
__device__ __forceinline__ cpair<T> cadd(cpair<T> a, cpair<T> b) {
  // TwoSum (Knuth, TAOCP vol. 2, 4.2.2) of the high parts; low parts carried.
  const T s = a.hi + b.hi;
  const T bp = s - a.hi;
  const T e = (a.hi - (s - bp)) + (b.hi - bp);
  return cpair<T>{s, e + (a.lo + b.lo)};
}

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int off) {
  return __shfl_up(v, off, 64);
}

// Apply with the row's out-degree and previous state already in registers.
template <typename T>
__device__ __forceinline__ double apply_vals(const rank_args& a, bool final_iter, uint32_t v,
                                             double s, T xold, uint32_t d, T* __restrict__ y,
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
  if (a.divide_by_od && d > 0) {
    const T xn = (T)(r / (double)d);
    y[v] = xn;
    return fabs((double)xn - (double)xold) * (double)d;
  }
  const T xn = (T)r;
  y[v] = xn;
  return fabs((double)xn - (double)xold);
}

// ---- tile kernel -----------------------------------------------------------------------
// Persistent: block b handles tiles b, b+G, ...  The next tile's descriptors and
// row_idx slice are loaded while the current one is summed (software pipeline);
// each row's out-degree and previous state are loaded with the row ends, so the
// apply at the end of the tile is one coalesced pass with no dependent loads.
template <typename T, bool WEIGHTED, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_pull_tiles(
    const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ row_idx,
    const float* __restrict__ wn, const uint32_t* __restrict__ tile_row,
    const uint64_t* __restrict__ tile_edge, uint64_t T_tiles, T* buf0, T* buf1,
    const uint32_t* __restrict__ od, rank_args a, const iter_ctl* __restrict__ ctl,
    cpair<T>* __restrict__ tile_carry, cpair<T>* __restrict__ tile_head,
    double* __restrict__ part_delta, double* __restrict__ part_rsum) {
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int WAVES = BLOCK / 64;
  __shared__ T s_val[TILE];
  __shared__ uint32_t s_end[TILE];
  __shared__ double s_sum[TILE];
  __shared__ cpair<T> s_wtot[WAVES];
  __shared__ int s_wflag[WAVES];
  __shared__ double s_red[2 * WAVES];

  bool final_iter = false;
  const T* x = buf0;
  T* y = buf1;
  if (ctl) {
    if (ctl->done) return;
    final_iter = (ctl->t + 1 == ctl->iters);
    if (ctl->cur) { x = buf1; y = buf0; }
  } else {
    final_iter = a.final_iter;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double delta = 0.0, rsum = 0.0;

  uint64_t t = blockIdx.x;
  uint32_t r0 = 0, r1 = 0;
  uint64_t e0 = 0, e1 = 0;
  uint32_t u[ITEMS];
  float w[ITEMS];
  if (t < T_tiles) {
    r0 = tile_row[t];
    r1 = tile_row[t + 1];
    e0 = tile_edge[t];
    e1 = tile_edge[t + 1];
    const int ne = (int)(e1 - e0);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int i = tid + k * BLOCK;
      u[k] = i < ne ? ld_stream(row_idx + e0 + i) : 0u;
      w[k] = (WEIGHTED && i < ne) ? ld_stream(wn + e0 + i) : 1.0f;
    }
  }

  for (; t < T_tiles; t += gridDim.x) {
    const int nr = (int)(r1 - r0), ne = (int)(e1 - e0);
    const int total = nr + ne;
    // 1. gathers of this tile, row ends, and the apply operands of its rows
    T val[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int i = tid + k * BLOCK;
      val[k] = i < ne ? ld_gather(x + u[k]) : (T)0;
    }
    uint32_t rod[ITEMS];
    T rx[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int i = tid + k * BLOCK;
      if (i < nr) {
        s_end[i] = (uint32_t)(ld_stream(col_ptr + r0 + i + 1) - e0);
        rod[k] = od[r0 + i];
        rx[k] = x[r0 + i];
      } else {
        rod[k] = 0u;
        rx[k] = (T)0;
      }
    }
    const bool continued = col_ptr[r0] < e0;  // first row began in an earlier tile
    // prefetch the next tile
    const uint64_t tn = t + gridDim.x;
    uint32_t nr0 = r0, nr1 = r1;
    uint64_t ne0 = e0, ne1 = e1;
    if (tn < T_tiles) {
      nr0 = tile_row[tn];
      nr1 = tile_row[tn + 1];
      ne0 = tile_edge[tn];
      ne1 = tile_edge[tn + 1];
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int i = tid + k * BLOCK;
      if (i < ne) {
        T v = val[k];
        if (WEIGHTED) v = v * (T)w[k];
        s_val[i] = v;
      }
    }
    if (tn < T_tiles) {
      const int nne = (int)(ne1 - ne0);
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int i = tid + k * BLOCK;
        u[k] = i < nne ? ld_stream(row_idx + ne0 + i) : 0u;
        w[k] = (WEIGHTED && i < nne) ? ld_stream(wn + ne0 + i) : 1.0f;
      }
    }
    __syncthreads();
    // 2. merge-path walk over this thread's ITEMS items
    const int d0 = min(tid * ITEMS, total), d1 = min(d0 + ITEMS, total);
    int lo = max(0, d0 - ne), hi = min(d0, nr);
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int)s_end[mid] + mid + 1 <= d0) lo = mid + 1;
      else hi = mid;
    }
    int xr = lo, ye = d0 - lo;
    T acc = (T)0;
    bool has_end = false;
    int head_row = -1;
    T head_acc = (T)0;
    for (int d = d0; d < d1; ++d) {
      if (xr < nr && (int)s_end[xr] <= ye) {
        if (!has_end) {
          has_end = true;
          head_row = xr;
          head_acc = acc;
        } else {
          s_sum[xr] = (double)acc;  // row started and ended inside this thread
        }
        acc = (T)0;
        ++xr;
      } else {
        acc += s_val[ye];
        ++ye;
      }
    }
    // 3. segmented inclusive scan of (has_end, carry) across the block
    cpair<T> sv{acc, (T)0};
    int sf = has_end ? 1 : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const T uh = shfl_up_t(sv.hi, off), ul = shfl_up_t(sv.lo, off);
      const int uf = __shfl_up(sf, off, 64);
      if (lane >= off && !sf) {
        sv = cadd(cpair<T>{uh, ul}, sv);
        sf = uf;
      }
    }
    if (lane == 63) {
      s_wtot[wave] = sv;
      s_wflag[wave] = sf;
    }
    __syncthreads();
    cpair<T> pv{(T)0, (T)0};
    for (int wv = 0; wv < wave; ++wv) {
      if (s_wflag[wv]) pv = s_wtot[wv];
      else pv = cadd(pv, s_wtot[wv]);
    }
    cpair<T> inc = sv;
    if (!sf) inc = cadd(pv, sv);
    cpair<T> ex{shfl_up_t(inc.hi, 1), shfl_up_t(inc.lo, 1)};
    if (lane == 0) ex = pv;
    if (tid == 0) ex = cpair<T>{(T)0, (T)0};
    if (has_end) {
      const cpair<T> full = cadd(ex, cpair<T>{head_acc, (T)0});
      if (head_row == 0 && continued) tile_head[t] = full;
      else s_sum[head_row] = (double)full.hi + (double)full.lo;
    }
    if (tid == BLOCK - 1) tile_carry[t] = inc;
    __syncthreads();
    // 4. apply, one coalesced pass over the tile's rows
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int i = tid + k * BLOCK;
      if (i < nr && !(i == 0 && continued))
        delta += apply_vals<T>(a, final_iter, r0 + i, s_sum[i], rx[k], rod[k], y, rsum);
    }
    r0 = nr0;
    r1 = nr1;
    e0 = ne0;
    e1 = ne1;
  }
  block_reduce2<BLOCK>(delta, rsum, s_red);
  if (threadIdx.x == 0) {
    part_delta[blockIdx.x] = delta;
    part_rsum[blockIdx.x] = rsum;
  }
}

// ---- fix-up of rows spanning tiles -------------------------------------------------------
// Short spans (<= 8 tiles): one thread per row, carries summed in tile order.
template <typename T>
__global__ __launch_bounds__(256) void k_pull_fixup_short(
    const uint32_t* __restrict__ fix_row, const uint32_t* __restrict__ fix_t0,
    const uint32_t* __restrict__ fix_t1, uint64_t n, const cpair<T>* __restrict__ tile_carry,
    const cpair<T>* __restrict__ tile_head, T* buf0, T* buf1, const uint32_t* __restrict__ od,
    rank_args a, const iter_ctl* __restrict__ ctl, double* __restrict__ part_delta,
    double* __restrict__ part_rsum) {
  __shared__ double s_red[8];
  bool final_iter = false;
  const T* x = buf0;
  T* y = buf1;
  if (ctl) {
    if (ctl->done) return;
    final_iter = (ctl->t + 1 == ctl->iters);
    if (ctl->cur) { x = buf1; y = buf0; }
  } else {
    final_iter = a.final_iter;
  }
  double delta = 0.0, rsum = 0.0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t0 = fix_t0[i], t1 = fix_t1[i];
    cpair<T> s{(T)0, (T)0};
    for (uint32_t t = t0; t < t1; ++t) s = cadd(s, tile_carry[t]);
    s = cadd(s, tile_head[t1]);
    delta += apply_row<T>(a, final_iter, fix_row[i], (double)s.hi + (double)s.lo, x, y, od, rsum);
  }
  block_reduce2<256>(delta, rsum, s_red);
  if (threadIdx.x == 0) {
    part_delta[blockIdx.x] = delta;
    part_rsum[blockIdx.x] = rsum;
  }
}

// Long spans (hub columns): one wave per row; lane j sums tiles t0+j, t0+j+64, ...
// then a fixed butterfly combines the lanes.
template <typename T>
__global__ __launch_bounds__(256) void k_pull_fixup_long(
    const uint32_t* __restrict__ fix_row, const uint32_t* __restrict__ fix_t0,
    const uint32_t* __restrict__ fix_t1, uint64_t n, const cpair<T>* __restrict__ tile_carry,
    const cpair<T>* __restrict__ tile_head, T* buf0, T* buf1, const uint32_t* __restrict__ od,
    rank_args a, const iter_ctl* __restrict__ ctl, double* __restrict__ part_delta,
    double* __restrict__ part_rsum) {
  __shared__ double s_red[8];
  bool final_iter = false;
  const T* x = buf0;
  T* y = buf1;
  if (ctl) {
    if (ctl->done) return;
    final_iter = (ctl->t + 1 == ctl->iters);
    if (ctl->cur) { x = buf1; y = buf0; }
  } else {
    final_iter = a.final_iter;
  }
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id();
  const uint64_t nwaves = grid_waves();
  double delta = 0.0, rsum = 0.0;
  for (uint64_t i = wave_id; i < n; i += nwaves) {
    const uint32_t t0 = fix_t0[i], t1 = fix_t1[i];
    cpair<T> s{(T)0, (T)0};
    for (uint32_t t = t0 + lane; t < t1; t += 64) s = cadd(s, tile_carry[t]);
    for (int off = 32; off > 0; off >>= 1) {
      cpair<T> o{__shfl_down(s.hi, off, 64), __shfl_down(s.lo, off, 64)};
      if (lane < off) s = cadd(s, o);
    }
    if (lane == 0) {
      s = cadd(s, tile_head[t1]);
      delta += apply_row<T>(a, final_iter, fix_row[i], (double)s.hi + (double)s.lo, x, y, od, rsum);
    }
  }
  block_reduce2<256>(delta, rsum, s_red);
  if (threadIdx.x == 0) {
    part_delta[blockIdx.x] = delta;
    part_rsum[blockIdx.x] = rsum;
  }
}

// ---- unfused apply (pe_rank_apply) -------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_rank_apply(uint32_t V, const T* __restrict__ x,
                                                    T* __restrict__ y, const uint32_t* __restrict__ od,
                                                    rank_args a, double* __restrict__ part_delta,
                                                    double* __restrict__ part_rsum) {
  __shared__ double s_red[8];
  double delta = 0.0, rsum = 0.0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    delta += apply_row<T>(a, a.final_iter, (uint32_t)v, (double)y[v], x, y, od, rsum);
  block_reduce2<256>(delta, rsum, s_red);
  if (threadIdx.x == 0) {
    part_delta[blockIdx.x] = delta;
    part_rsum[blockIdx.x] = rsum;
  }
}

// ---- end of an iteration: reduce the partials, advance the controller ----------------------
// Plato's loop control (pagerank.hpp:100-146): after a non-final iteration,
// delta < eps jumps to one more, final, iteration.  The optional normalised
// tolerance does the same on delta / sum(rank).
__global__ __launch_bounds__(256) void k_iter_finish(const double* __restrict__ part_delta,
                                                     const double* __restrict__ part_rsum,
                                                     unsigned n, iter_ctl* __restrict__ ctl,
                                                     double* __restrict__ deltas,
                                                     double* __restrict__ rsums) {
  __shared__ double sd[256], sr[256];
  if (ctl && ctl->done) return;
  double d = 0.0, r = 0.0;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    d += part_delta[i];
    r += part_rsum[i];
  }
  sd[threadIdx.x] = d;
  sr[threadIdx.x] = r;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sd[threadIdx.x] += sd[threadIdx.x + w];
      sr[threadIdx.x] += sr[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double delta = sd[0], rs = sr[0];
  if (!ctl) {
    deltas[0] = delta;
    rsums[0] = rs;
    return;
  }
  const bool final_iter = (ctl->t + 1 == ctl->iters);
  const uint32_t k = ctl->executed;
  if (k < ctl->cap) {
    deltas[k] = final_iter ? 0.0 : delta;
    rsums[k] = rs;
  }
  ctl->executed = k + 1;
  ctl->last_delta = final_iter ? 0.0 : delta;
  ctl->cur ^= 1u;
  if (final_iter) {
    ctl->done = 1;
    return;
  }
  uint32_t t = ctl->t;
  const bool hit_eps = ctl->eps > 0.0 && delta < ctl->eps;
  const bool hit_tol = ctl->l1tol > 0.0 && rs > 0.0 && (delta / rs) < ctl->l1tol;
  if ((hit_eps || hit_tol) && ctl->iters >= 2) t = ctl->iters - 2;
  ctl->t = t + 1;
}

// ---- host side ------------------------------------------------------------------------------

template <typename T, bool W>
static void launch_tiles_t(pe_graph_s* g, T* b0, T* b1, const rank_args& a, const iter_ctl* ctl,
                           rank_work& wk) {
  hipStream_t s = g->ctx->stream;
  const unsigned G = wk.tile_grid;
  auto* tc = reinterpret_cast<cpair<T>*>(wk.tile_carry);
  auto* th = reinterpret_cast<cpair<T>*>(wk.tile_head);
  const float* wn = W ? g->wn.p : nullptr;
  ktime_scope kt(wk.timed ? g->ctx : nullptr, PE_KT_PULL_TILE);
  switch (g->tile_items) {
    case 1024:
      k_pull_tiles<T, W, 256, 4><<<G, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, wn, g->tile_row.p,
                                                   g->tile_edge.p, g->num_tiles, b0, b1, g->od.p, a,
                                                   ctl, tc, th, wk.part_delta, wk.part_rsum);
      break;
    case 4096:
      k_pull_tiles<T, W, 256, 16><<<G, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, wn, g->tile_row.p,
                                                    g->tile_edge.p, g->num_tiles, b0, b1, g->od.p, a,
                                                    ctl, tc, th, wk.part_delta, wk.part_rsum);
      break;
    default:
      k_pull_tiles<T, W, 256, 8><<<G, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, wn, g->tile_row.p,
                                                   g->tile_edge.p, g->num_tiles, b0, b1, g->od.p, a,
                                                   ctl, tc, th, wk.part_delta, wk.part_rsum);
  }
  PE_HIP(hipGetLastError());
}

template <typename T>
static void launch_iteration_t(pe_graph_s* g, T* b0, T* b1, bool weighted, const rank_args& a,
                               iter_ctl* ctl, rank_work& wk, double* deltas, double* rsums) {
  hipStream_t s = g->ctx->stream;
  if (weighted) launch_tiles_t<T, true>(g, b0, b1, a, ctl, wk);
  else launch_tiles_t<T, false>(g, b0, b1, a, ctl, wk);
  const uint64_t nshort = g->num_fix - g->num_fix_long, nlong = g->num_fix_long;
  auto* tc = reinterpret_cast<cpair<T>*>(wk.tile_carry);
  auto* th = reinterpret_cast<cpair<T>*>(wk.tile_head);
  {
    ktime_scope kt(wk.timed ? g->ctx : nullptr, PE_KT_PULL_FIXUP);
    k_pull_fixup_short<T><<<wk.fix_grid, 256, 0, s>>>(
        g->fix_row.p, g->fix_t0.p, g->fix_t1.p, nshort, tc, th, b0, b1, g->od.p, a, ctl,
        wk.part_delta + wk.tile_grid, wk.part_rsum + wk.tile_grid);
    k_pull_fixup_long<T><<<wk.fix_grid, 256, 0, s>>>(
        g->fix_row.p + nshort, g->fix_t0.p + nshort, g->fix_t1.p + nshort, nlong, tc, th, b0, b1,
        g->od.p, a, ctl, wk.part_delta + wk.tile_grid + wk.fix_grid,
        wk.part_rsum + wk.tile_grid + wk.fix_grid);
  }
  PE_HIP(hipGetLastError());
  k_iter_finish<<<1, 256, 0, s>>>(wk.part_delta, wk.part_rsum, wk.tile_grid + 2 * wk.fix_grid, ctl,
                                  deltas, rsums);
  PE_HIP(hipGetLastError());
}

void rank_work::init(pe_graph_s* g, int precision) {
  pe_ctx_s* ctx = g->ctx;
  const size_t pair_bytes = precision == 64 ? 16 : 8;
  tile_grid = persistent_grid(ctx, g->num_tiles, 8);
  fix_grid = persistent_grid(ctx, std::max<uint64_t>(1, (g->num_fix + 255) / 256), 2);
  own_tiles.alloc(ctx, 2 * g->num_tiles * pair_bytes);
  tile_carry = own_tiles.p;
  tile_head = own_tiles.p + g->num_tiles * pair_bytes;
  pb_parts = 0;
  if (precision != 64 && g->pull_layout == PE_PULL_BINNED && g->pb)
    pb_parts = binned_grid_gather(g) + binned_grid_split(g);
  const size_t parts = std::max<size_t>(tile_grid + 2 * fix_grid, pb_parts);
  own_parts.alloc(ctx, 2 * parts + 2);
  part_delta = own_parts.p;
  part_rsum = own_parts.p + parts;
  scalars = own_parts.p + 2 * parts;
  timed = true;
}

static uint64_t align256(uint64_t b) { return (b + 255) & ~255ull; }

uint64_t rank_work::arena_bytes(pe_graph_s* g, int precision) {
  const uint64_t pair_bytes = precision == 64 ? 16 : 8;
  const uint64_t tg = persistent_grid(g->ctx, g->num_tiles, 8);
  const uint64_t fg = persistent_grid(g->ctx, std::max<uint64_t>(1, (g->num_fix + 255) / 256), 2);
  const uint64_t parts = tg + 2 * fg;
  return 2 * align256(g->num_tiles * pair_bytes) + 2 * align256(parts * 8) + align256(16);
}

void rank_work::init_in(pe_graph_s* g, int precision, uint8_t* arena, uint64_t cap) {
  const uint64_t need = arena_bytes(g, precision);
  if (need > cap) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "graph needs %llu bytes of scratch, above the %llu-byte static array",
                  (unsigned long long)need, (unsigned long long)cap);
    fail(PE_ERR_RANGE, msg);
  }
  const uint64_t pair_bytes = precision == 64 ? 16 : 8;
  tile_grid = persistent_grid(g->ctx, g->num_tiles, 8);
  fix_grid = persistent_grid(g->ctx, std::max<uint64_t>(1, (g->num_fix + 255) / 256), 2);
  const uint64_t parts = tile_grid + 2 * fix_grid;
  uint8_t* p = arena;
  tile_carry = p;
  p += align256(g->num_tiles * pair_bytes);
  tile_head = p;
  p += align256(g->num_tiles * pair_bytes);
  part_delta = reinterpret_cast<double*>(p);
  p += align256(parts * 8);
  part_rsum = reinterpret_cast<double*>(p);
  p += align256(parts * 8);
  scalars = reinterpret_cast<double*>(p);
  pb_parts = 0;  // raw sums and FP64 iterations never take the binned layout
  timed = false;
}

// The static scratch array of pe_pull_sum_noalloc: one copy per GPU, held until the
// process ends.  Host code never passes its name to a kernel: it is a host-side
// stand-in, not a device address.  Every user goes through the address that
// hipGetSymbolAddress returns for the device (pull_scratch_setup).
static __device__ __align__(256) uint8_t scratch[PE_PULL_SUM_NOALLOC_MAX_BYTES];

void* pull_scratch_setup(int device) {
  static std::mutex mu;
  static void* addr[256] = {};
  PE_REQUIRE(device >= 0 && device < 256, PE_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* ptr = nullptr;
    PE_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    PE_REQUIRE(ptr != nullptr, PE_ERR_HIP, "no device address for the scratch array");
    addr[device] = ptr;
  }
  return addr[device];
}

void rank_iteration(pe_graph_s* g, int precision, void* b0, void* b1, bool weighted,
                    const rank_args& a, iter_ctl* ctl, rank_work& wk, double* deltas,
                    double* rsums) {
  if (precision != 64 && a.mode == MODE_APPLY && a.mass > 0.0 && g->pull_layout == PE_PULL_BINNED && g->pb) {
    PE_REQUIRE(wk.pb_parts > 0, PE_ERR_STATE, "pull layout changed after the iteration was set up");
    binned_iteration(g, (float*)b0, (float*)b1, weighted, a, ctl, wk.part_delta, wk.part_rsum);
    k_iter_finish<<<1, 256, 0, g->ctx->stream>>>(wk.part_delta, wk.part_rsum, wk.pb_parts, ctl, deltas,
                                                 rsums);
    PE_HIP(hipGetLastError());
    return;
  }
  if (precision == 64)
    launch_iteration_t<double>(g, (double*)b0, (double*)b1, weighted, a, ctl, wk, deltas, rsums);
  else
    launch_iteration_t<float>(g, (float*)b0, (float*)b1, weighted, a, ctl, wk, deltas, rsums);
}

void rank_apply_only(pe_graph_s* g, int precision, const void* x, void* y, const rank_args& a,
                     rank_work& wk) {
  hipStream_t s = g->ctx->stream;
  const unsigned G = wk.tile_grid;
  if (precision == 64)
    k_rank_apply<double><<<G, 256, 0, s>>>(g->V, (const double*)x, (double*)y, g->od.p, a,
                                           wk.part_delta, wk.part_rsum);
  else
    k_rank_apply<float><<<G, 256, 0, s>>>(g->V, (const float*)x, (float*)y, g->od.p, a,
                                          wk.part_delta, wk.part_rsum);
  PE_HIP(hipGetLastError());
  k_iter_finish<<<1, 256, 0, s>>>(wk.part_delta, wk.part_rsum, G, nullptr, wk.scalars,
                                  wk.scalars + 1);
  PE_HIP(hipGetLastError());
}

}  // namespace pe
