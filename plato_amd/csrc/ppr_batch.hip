// Batched Personalized PageRank: k = 64 queries as the 64 columns of one
// FP32 rank block X[V][64] (row-major, 256 B per vertex).
//
// Each query is plato::algo::personalized_pagerank with its restart mass spread
// over a seed set (personalized_pagerank.hpp:40-134 for one seed, the TrustRank
// restart of trustrank.hpp:197-250 for several), with the reference's stopping
// rule per column: after a non-final iteration with delta < tol, one more, final,
// iteration stores the undivided rank (pagerank.hpp:137-139).
//
// Pull gather, one wave per task, lane j = column j: every in-edge u->v reads the
// 256 B row X[u][0..63] once for all 64 queries.  A task is either a run of
// consecutive columns (vertices) holding at most C edges, or one chunk of a long
// column; chunk partials are combined in chunk order by the fix-up kernel.
// Sums are Kahan-compensated per lane in ascending source order, so a column's
// arithmetic never depends on the other 63 columns: a query gives the same bits
// alone or in any mix.
//
// Columns move through EMPTY -> LOADING -> ACTIVE -> FINAL -> DONE on the device
// (k_batch_finish); a slot is refilled by writing its initial column during the
// same pass that updates the others, so admission costs no extra pass.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "pe_internal.hpp"
#include "ppr_batch.hpp"

namespace pe {

static constexpr int BK = 64;  // columns per block (one per lane)

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  // Kahan, "Further remarks on reducing truncation errors", CACM 8(1), 1965
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// TwoSum of two (sum, -compensation) accumulators, in order a then b
__device__ __forceinline__ void pair_add(float& s, float& c, float s2, float c2) {
  const float hi = s + s2;
  const float bp = hi - s;
  const float lo = (s - (hi - bp)) + (s2 - bp);
  s = hi;
  c = c + c2 - lo;
}

struct lane_cfg {
  uint32_t status;
  double init_dist;  // 1/|S|
  double restart;    // (1-d)/|S|
};

// Per-lane apply for vertex v given the gathered sum (s - c).
__device__ __forceinline__ float batch_apply(const lane_cfg& L, double damping, double sum, float xold,
                                             uint32_t d, bool seed, double& delta) {
  switch (L.status) {
    case COL_LOADING: {  // x0 = [v in S] / |S| / od (trustrank.hpp:197-206)
      double x0 = seed ? L.init_dist : 0.0;
      if (d > 0) x0 = x0 / (double)d;
      return (float)x0;
    }
    case COL_ACTIVE: {
      const double r = (seed ? L.restart : 0.0) + damping * sum;
      if (d > 0) {
        const float xn = (float)(r / (double)d);
        delta += fabs((double)xn - (double)xold) * (double)d;
        return xn;
      }
      const float xn = (float)r;
      delta += fabs((double)xn - (double)xold);
      return xn;
    }
    case COL_FINAL:  // final iteration stores the undivided rank
      return (float)((seed ? L.restart : 0.0) + damping * sum);
    case COL_DONE:
      return xold;
    default:
      return 0.0f;
  }
}

#ifndef PE_BATCH_UNROLL
#define PE_BATCH_UNROLL 4
#endif
static constexpr int BU = PE_BATCH_UNROLL;  // row gathers in flight per wave
#ifndef PE_BATCH_GROUP
#define PE_BATCH_GROUP 32  // columns per task at most (their pointers fit one load per lane)
#endif
static_assert(PE_BATCH_GROUP >= 1 && PE_BATCH_GROUP <= 63, "a task's column pointers take one lane each");
// Cache hints: the new block Y and the edge list stream through once per pass, so
// they are stored / loaded non-temporally and leave the caches to the hot rows of X
// (with degree relabeling the first 2^20 rows cover ~88% of the gathers at scale 24).
// Rows at or above PE_BATCH_COLD_ROW are loaded non-temporally too (0 = never).
#ifndef PE_BATCH_NT_STREAM
#define PE_BATCH_NT_STREAM 0
#endif
#ifndef PE_BATCH_COLD_ROW
#define PE_BATCH_COLD_ROW 0
#endif

__device__ __forceinline__ uint32_t ld_edge(const uint32_t* p) {
#if PE_BATCH_NT_STREAM
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

__device__ __forceinline__ void st_row(float* p, float v) {
#if PE_BATCH_NT_STREAM
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// entry c of row u of a block with `stride` floats per row (the wide block: 64, packed: 8)
__device__ __forceinline__ float ld_row(const float* X, uint32_t u, uint32_t stride, uint32_t c) {
  const float* p = X + (uint64_t)u * stride + c;
#if PE_BATCH_COLD_ROW
  if (u >= (uint32_t)PE_BATCH_COLD_ROW) return __builtin_nontemporal_load(p);
#endif
  return *p;
}

// A column's sum (or a long column's chunk sum) runs as BS interleaved Kahan sums: edge k
// of the column (counted from its first edge, or the chunk's) goes to sub-sum k % BS, and
// the sub-sums are combined in order 0..BS-1 with TwoSum.  Every pass kernel below, however
// many edges it puts in one row load, keeps exactly this arithmetic, so a column's bits do
// not depend on how many other columns are in the pass.
static constexpr int BS = 4;

// Edges per row load P: lane group p = lane / (64 / P) takes edge t * P + p of step t, so a
// wave-wide load covers P edges for up to 64 / P columns.  P = 1 is lane j = column j.
// Each lane keeps the BS / P sub-sums of its edge residue: sub-sum a * P + p in slot a.
template <int P>
struct col_sum {
  static constexpr int NA = BS / P;
  float s[NA], c[NA];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int a = 0; a < NA; ++a) s[a] = c[a] = 0.f;
  }
  // the combined sum (s - c) on the lanes of group 0
  __device__ __forceinline__ void combine(float& S, float& C) const {
    constexpr int W = BK / P;
    S = 0.f;
    C = 0.f;
#pragma unroll
    for (int k = 0; k < BS; ++k) {
      float vs = s[k / P], vc = c[k / P];
      if (P > 1 && (k % P) != 0) {
        vs = __shfl_down(vs, (k % P) * W, 64);
        vc = __shfl_down(vc, (k % P) * W, 64);
      }
      pair_add(S, C, vs, vc);
    }
  }
};

// Rows of steps [t0, t0 + NB) into r (0 where there is no edge or no sum is needed).
// The ids of up to 64 edges sit one per lane in idv: P = 1 reads them back with readlane,
// wider steps with a lane shuffle.
template <int P, int NB>
__device__ __forceinline__ void load_rows(const float* X, uint32_t stride, uint32_t idv, uint32_t t0, uint32_t m, int p,
                                          uint32_t col, bool needs, float (&r)[NB]) {
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const uint32_t t = t0 + (uint32_t)i;  // wave-uniform
    r[i] = 0.0f;
    if (t * P < m) {
      const uint32_t k = t * P + (uint32_t)p;
      uint32_t u;
      if (P == 1) u = __builtin_amdgcn_readlane(idv, (int)t);
      else u = (uint32_t)__shfl((int)idv, (int)(k & 63u), 64);
      if (needs && k < m) r[i] = ld_row(X, u, stride, col);
    }
  }
}

// The rows of BU steps starting at step t (a multiple of BU) into their sub-sums: step
// t + i carries sub-sums (t + i) % (BS / P) of the lane's residue, i.e. slot (OFF + i) % NA
// with OFF = t % NA (0, or BU when a lane keeps more sub-sums than a group has steps).
template <int P, int OFF>
__device__ __forceinline__ void add_rows(col_sum<P>& acc, const float (&r)[BU], uint32_t t, uint32_t m, int p) {
  constexpr int NA = col_sum<P>::NA;
#pragma unroll
  for (int i = 0; i < BU; ++i)
    if ((t + (uint32_t)i) * P + (uint32_t)p < m) kahan_add(acc.s[(OFF + i) % NA], acc.c[(OFF + i) % NA], r[i]);
}

// Sum over the in-edges [eb, ee) of one column in ascending source order per sub-sum.  The
// ids of up to 64 edges come in one vector load (lane i holds edge i), and the rows go in
// groups of BU steps, double-buffered: the next group's rows are in flight while the
// current group is added.  Lanes whose column needs no sum issue no row loads.
template <int P>
__device__ __forceinline__ void column_sum(const float* X, uint32_t stride, const uint32_t* __restrict__ row_idx,
                                           uint64_t eb, uint64_t ee, int lane, int p, uint32_t col, bool needs,
                                           col_sum<P>& acc) {
  constexpr int NA = col_sum<P>::NA;
  static_assert(NA <= BU ? BU % NA == 0 : NA == 2 * BU, "sub-sum slots must tile the step groups");
  for (uint64_t e = eb; e < ee; e += 64) {
    const uint32_t m = (uint32_t)std::min<uint64_t>(64, ee - e);  // wave-uniform
    const uint32_t idv = (uint32_t)lane < m ? ld_edge(row_idx + e + lane) : 0u;
    const uint32_t steps = (m + P - 1) / P;
    float cur[BU];
    load_rows<P, BU>(X, stride, idv, 0, m, p, col, needs, cur);
    for (uint32_t t = 0; t < steps; t += BU) {
      float nxt[BU];
      if (t + BU < steps) load_rows<P, BU>(X, stride, idv, t + BU, m, p, col, needs, nxt);
      if (NA > BU && ((t / BU) & 1u)) add_rows<P, (NA > BU ? BU : 0)>(acc, cur, t, m, p);
      else add_rows<P, 0>(acc, cur, t, m, p);
#pragma unroll
      for (int i = 0; i < BU; ++i) cur[i] = nxt[i];
    }
  }
}

// the value of lane + 1 (64-bit, as two 32-bit shuffles)
__device__ __forceinline__ uint64_t lane_pair_down(uint64_t v) {
  const uint32_t lo = (uint32_t)__shfl_down((int)(uint32_t)v, 1, 64);
  const uint32_t hi = (uint32_t)__shfl_down((int)(uint32_t)(v >> 32), 1, 64);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int i) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, i);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), i);
  return ((uint64_t)hi << 32) | lo;
}

// Which columns a pass must write, and which pass kernel runs.  A column is covered while
// LOADING, ACTIVE or FINAL, and for the first pass after it turns DONE (that pass copies its
// final value into the other buffer; after it both buffers hold it).  EMPTY and settled DONE
// columns are left alone: nothing reads an EMPTY column, and a settled one is the same in
// both buffers.  At most BATCH_PACKED_W covered columns run on the packed block (it may be
// read when the last pass was packed over the same columns: else the pass reads the wide
// block); at most 16 or 32 put 4 or 2 edges in a row load of the wide block; more run one
// edge per load, every column written.  The limits come from the batch options (these
// widths by default; a kernel can be turned off).
__global__ void k_batch_plan(batch_ctl* __restrict__ ctl, const uint32_t* __restrict__ status) {
  const int j = threadIdx.x;  // one wave, lane j = column j
  const uint32_t st = status[j];
  const bool covered = st == COL_LOADING || st == COL_ACTIVE || st == COL_FINAL || (st == COL_DONE && ctl->pend[j]);
  const uint64_t mask = __ballot(covered);
  const uint32_t n = (uint32_t)__popcll(mask);
  const uint32_t r = (uint32_t)__popcll(mask & ((1ull << j) - 1ull));
  const bool packed = n >= 1 && n <= ctl->max_packed;
  // the packed block is valid if the last pass was packed over exactly these columns
  const uint32_t last_n = ctl->pk_n;
  const bool same = !covered || (r < (uint32_t)BATCH_PACKED_W && ctl->pk_cols[r] == (uint8_t)j);
  const bool valid = packed && last_n == n && __ballot(!same) == 0ull;
  __syncthreads();  // every lane has read the last plan before it is overwritten
  ctl->qof[j] = covered ? (uint8_t)r : (uint8_t)0xFF;
  if (covered) {
    ctl->cols[r] = (uint8_t)j;
    if (packed) ctl->pk_cols[r] = (uint8_t)j;
  }
  if (j == 0) {
    ctl->ncols = n;
    ctl->mode = packed ? BATCH_MODE_PACKED : n <= ctl->max_edges4 ? 4u : n <= ctl->max_edges2 ? 2u : 1u;
    ctl->pk_valid = valid ? 1u : 0u;
    ctl->pk_n = packed ? n : 0xFFFFFFFFu;
  }
}

// One pass over the tasks with P edges per row load; runs only if the plan chose it.
// PK: the packed pass (P = 4, at most BATCH_PACKED_W columns): rows come from the packed
// block when the plan allows it, else from the wide one, and every new value goes to both.
template <int P, bool PK>
__global__ __launch_bounds__(256) void k_batch_tasks(
    const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ row_idx,
    const uint32_t* __restrict__ od, const uint64_t* __restrict__ seedmask,
    const bt_task* __restrict__ tasks, uint64_t ntasks, float* X0, float* X1, float* P0, float* P1,
    const batch_ctl* __restrict__ ctl, const uint32_t* __restrict__ status,
    const double* __restrict__ init_dist, double damping, float2* __restrict__ partial,
    double* __restrict__ part_delta) {
  static_assert(!PK || P == 4, "the packed pass puts 4 edges in a load");
  if (ctl->mode != (PK ? BATCH_MODE_PACKED : (uint32_t)P)) return;
  constexpr int W = BK / P;
  __shared__ double s_red[4][BK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = lane / W, q = lane % W;
  const uint32_t cur = ctl->cur;
  const float* X = cur ? X1 : X0;
  float* Y = cur ? X0 : X1;
  float* YP = cur ? P0 : P1;
  const bool has_col = P == 1 || (uint32_t)q < ctl->ncols;
  const uint32_t col = P == 1 ? (uint32_t)lane : (has_col ? (uint32_t)ctl->cols[q] : 0u);
  // where this pass reads rows: the packed block (entry q) or the wide one (entry col)
  const bool from_packed = PK && ctl->pk_valid;
  const float* XR = from_packed ? (cur ? P1 : P0) : X;
  const uint32_t rstride = from_packed ? (uint32_t)BATCH_PACKED_W : (uint32_t)BK;
  const uint32_t rc = from_packed ? (uint32_t)q : col;
  const bool writer = has_col && p == 0;  // the lane that applies and stores its column
  lane_cfg L;
  L.status = has_col ? status[col] : COL_EMPTY;
  L.init_dist = has_col ? init_dist[col] : 0.0;
  L.restart = (1.0 - damping) * L.init_dist;
  // only ACTIVE and FINAL columns use the gathered sum; when none does, no row is read
  const bool needs = L.status == COL_ACTIVE || L.status == COL_FINAL;
  const bool any_needs = __ballot(needs) != 0ull;
  double delta = 0.0;
  const uint64_t wave_id = grid_wave_id();
  const uint64_t nwaves = grid_waves();
  col_sum<P> acc;
  for (uint64_t ti = wave_id; ti < ntasks; ti += nwaves) {
    const bt_task t = tasks[ti];
    if (t.chunk == 0xFFFFFFFFu) {
      // the task's columns (at most 63) with their pointers, seed masks and out-degrees
      // in one load per lane, read back per column with readlane
      const uint32_t n = t.v1 - t.v0;
      const uint64_t cp = (uint32_t)lane <= n ? col_ptr[t.v0 + lane] : 0ull;
      const uint64_t smv = (uint32_t)lane < n ? seedmask[t.v0 + lane] : 0ull;
      const uint32_t odv = (uint32_t)lane < n ? od[t.v0 + lane] : 0u;
      // A vertex without in-edges that seeds no column holds 0 in every column and both
      // buffers, in every column state (its sum is 0 and it gets no restart), so it is
      // skipped: no read, no write.  Release zeroes a seed's entries when it stops being one.
      const uint64_t cp_next = lane_pair_down(cp);
      const bool live = (uint32_t)lane < n && (cp != cp_next || smv != 0ull);
      const uint64_t live_mask = __ballot(live);
      float xnext = 0.f;  // the old row of the first live vertex, ahead
      if (live_mask && writer) xnext = ld_row(XR, t.v0 + __builtin_ctzll(live_mask), rstride, rc);
      for (uint32_t i = 0; i < n; ++i) {
        if (!((live_mask >> i) & 1ull)) continue;  // wave-uniform
        const uint32_t v = t.v0 + i;
        const float xold = xnext;
        const uint64_t rest = live_mask & ~((2ull << i) - 1ull);
        if (rest && writer) xnext = ld_row(XR, t.v0 + __builtin_ctzll(rest), rstride, rc);  // the next live row
        float S = 0.f, C = 0.f;
        if (any_needs) {
          acc.zero();
          column_sum<P>(XR, rstride, row_idx, lane_u64(cp, i), lane_u64(cp, i + 1), lane, p, rc, needs, acc);
          acc.combine(S, C);
        }
        const uint64_t sm = lane_u64(smv, i);
        const uint32_t d = __builtin_amdgcn_readlane(odv, i);
        if (writer) {
          const float y = batch_apply(L, damping, (double)S - (double)C, xold, d, (sm >> col) & 1ull, delta);
          st_row(Y + (uint64_t)v * BK + col, y);
          if (PK) st_row(YP + (uint64_t)v * BATCH_PACKED_W + q, y);
        }
      }
    } else {
      float S = 0.f, C = 0.f;
      if (any_needs) {
        acc.zero();
        column_sum<P>(XR, rstride, row_idx, t.e0, t.e1, lane, p, rc, needs, acc);
        acc.combine(S, C);
      }
      if (writer) partial[(uint64_t)t.chunk * BK + col] = make_float2(S, C);
    }
  }
  s_red[wave][lane] = 0.0;
  __syncthreads();
  if (writer) s_red[wave][col] = delta;
  __syncthreads();
  if (threadIdx.x < BK) {
    double d = 0.0;
    for (int w = 0; w < 4; ++w) d += s_red[w][threadIdx.x];
    part_delta[(uint64_t)blockIdx.x * BK + threadIdx.x] = d;
  }
}

// Long columns: chunk partials combined in chunk order, then applied.
__global__ __launch_bounds__(256) void k_batch_fixup(
    const uint32_t* __restrict__ od, const uint64_t* __restrict__ seedmask,
    const bt_long* __restrict__ longs, uint64_t nlong, float* X0, float* X1, float* P0, float* P1,
    const batch_ctl* __restrict__ ctl, const uint32_t* __restrict__ status,
    const double* __restrict__ init_dist, double damping, const float2* __restrict__ partial,
    double* __restrict__ part_delta) {
  __shared__ double s_red[4][BK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t cur = ctl->cur;
  const float* X = cur ? X1 : X0;
  float* Y = cur ? X0 : X1;
  lane_cfg L;
  L.status = status[lane];
  L.init_dist = init_dist[lane];
  L.restart = (1.0 - damping) * L.init_dist;
  const bool packed = ctl->mode == BATCH_MODE_PACKED;
  const uint32_t qk = ctl->qof[lane];
  float* YP = cur ? P0 : P1;
  double delta = 0.0;
  const uint64_t wave_id = grid_wave_id();
  const uint64_t nwaves = grid_waves();
  for (uint64_t li = wave_id; li < nlong; li += nwaves) {
    const bt_long lc = longs[li];
    float s = 0.f, c = 0.f;
    uint32_t k = 0;
    for (; k + 8 <= lc.nchunks; k += 8) {
      float2 p[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) p[i] = partial[(uint64_t)(lc.c0 + k + i) * BK + lane];
#pragma unroll
      for (int i = 0; i < 8; ++i) pair_add(s, c, p[i].x, p[i].y);
    }
    for (; k < lc.nchunks; ++k) {
      const float2 p = partial[(uint64_t)(lc.c0 + k) * BK + lane];
      pair_add(s, c, p.x, p.y);
    }
    const uint32_t v = lc.v;
    const uint64_t sm = seedmask[v];
    const float xold = X[(uint64_t)v * BK + lane];  // the wide block holds every covered column
    const float y = batch_apply(L, damping, (double)s - (double)c, xold, od[v], (sm >> lane) & 1ull, delta);
    Y[(uint64_t)v * BK + lane] = y;
    if (packed && qk != 0xFFu) YP[(uint64_t)v * BATCH_PACKED_W + qk] = y;
  }
  s_red[wave][lane] = delta;
  __syncthreads();
  if (threadIdx.x < BK) {
    double d = 0.0;
    for (int w = 0; w < 4; ++w) d += s_red[w][threadIdx.x];
    part_delta[(uint64_t)blockIdx.x * BK + threadIdx.x] = d;
  }
}

// Per-column delta, state transitions, buffer flip.  One block of 64 threads.
static constexpr int FIN_GROUPS = 16;

__global__ __launch_bounds__(FIN_GROUPS * BK) void k_batch_finish(const double* __restrict__ part_delta, uint32_t nparts,
                               batch_ctl* __restrict__ ctl, uint32_t* __restrict__ status,
                               uint32_t* __restrict__ iters, double* __restrict__ last_delta,
                               double tol, uint32_t max_iters) {
  // FIN_GROUPS groups of 64 lanes each sum every FIN_GROUPS-th block's partials
  // (fixed order), then group sums are combined in group order: deterministic
  __shared__ double s_sum[FIN_GROUPS][BK];
  const int j = threadIdx.x & (BK - 1), grp = threadIdx.x / BK;
  {
    double d = 0.0;
    for (uint32_t b = grp; b < nparts; b += FIN_GROUPS) d += part_delta[(uint64_t)b * BK + j];
    s_sum[grp][j] = d;
  }
  __syncthreads();
  if (threadIdx.x < BK) {
    double d = 0.0;
    for (int q = 0; q < FIN_GROUPS; ++q) d += s_sum[q][j];
    const uint32_t st = status[j];
    if (st == COL_LOADING) {
      status[j] = COL_ACTIVE;
      iters[j] = 0;
    } else if (st == COL_ACTIVE) {
      const uint32_t it = iters[j] + 1;
      iters[j] = it;
      last_delta[j] = d;
      // delta < tol jumps to one more, final, iteration; the cap makes the
      // max_iters-th iteration the final one
      if (d < tol || it + 1 >= max_iters) status[j] = COL_FINAL;
    } else if (st == COL_FINAL) {
      iters[j] = iters[j] + 1;
      status[j] = COL_DONE;
      ctl->pend[j] = 1u;  // the next pass copies the final value into the other buffer
    } else if (st == COL_DONE) {
      ctl->pend[j] = 0u;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->cur ^= 1u;
    ctl->iteration += 1;
  }
}

__global__ void k_batch_seed_bits(uint64_t* __restrict__ seedmask, const uint32_t* __restrict__ ids,
                                  uint32_t n, uint32_t slot, int set) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (set) atomicOr((unsigned long long*)&seedmask[ids[i]], 1ull << slot);
    else atomicAnd((unsigned long long*)&seedmask[ids[i]], ~(1ull << slot));
  }
}

// A released slot's seeds stop being seeds: their entries in that column go to 0 in both
// buffers, so vertices without in-edges keep the all-zero state the gather skips.
// In the packed block a row holds whatever columns the last packed pass covered, so a seed
// that is left without in-edges and without seeds (a vertex the passes skip) gets its whole
// packed row cleared; the rows of other vertices are rewritten by every pass.
__global__ void k_batch_zero_slot(float* X0, float* X1, float* P0, float* P1, const uint64_t* __restrict__ col_ptr,
                                  const uint64_t* __restrict__ seedmask, const uint32_t* __restrict__ ids, uint32_t n,
                                  uint32_t slot) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = ids[i];
    X0[(uint64_t)v * BK + slot] = 0.0f;
    X1[(uint64_t)v * BK + slot] = 0.0f;
    if (seedmask[v] == 0ull && col_ptr[v] == col_ptr[v + 1])
      for (int k = 0; k < BATCH_PACKED_W; ++k) {
        P0[(uint64_t)v * BATCH_PACKED_W + k] = 0.0f;
        P1[(uint64_t)v * BATCH_PACKED_W + k] = 0.0f;
      }
  }
}

__global__ void k_batch_set_slot(uint32_t* status, double* init_dist, uint32_t* iters, double* last_delta,
                                 uint32_t slot, uint32_t st, double dist) {
  status[slot] = st;
  init_dist[slot] = dist;
  iters[slot] = 0;
  last_delta[slot] = 0.0;
}

// Columns of the current block -> dense [n][V] (internal order).
__global__ void k_batch_extract(const float* X0, const float* X1, const batch_ctl* __restrict__ ctl,
                                uint32_t V, const uint32_t* __restrict__ slots, uint32_t n,
                                float* __restrict__ out) {
  const float* X = ctl->cur ? X1 : X0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t i = 0; i < n; ++i) out[(uint64_t)i * V + v] = X[v * BK + slots[i]];
}

// Top-k support: per-slot histogram of the top 11 bits of the (non-negative)
// float pattern, then candidates at or above the threshold bin.
// Histogram of the 2048 high-bit buckets of the positive scores of each slot: every
// block counts its vertices in LDS, then adds its non-zero buckets to the global one.
__global__ __launch_bounds__(256) void k_batch_hist(const float* X0, const float* X1,
                                                    const batch_ctl* __restrict__ ctl, uint32_t V,
                                                    const uint32_t* __restrict__ slots, uint32_t n,
                                                    unsigned int* __restrict__ hist) {
  __shared__ unsigned int sh[2048];
  const float* X = ctl->cur ? X1 : X0;
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t j = threadIdx.x; j < 2048; j += blockDim.x) sh[j] = 0u;
    __syncthreads();
    const uint32_t slot = slots[i];
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
         v += (uint64_t)gridDim.x * blockDim.x) {
      const float f = X[v * BK + slot];
      if (f > 0.f) atomicAdd(&sh[__float_as_uint(f) >> 20], 1u);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < 2048; j += blockDim.x)
      if (sh[j]) atomicAdd(&hist[i * 2048u + j], sh[j]);
    __syncthreads();
  }
}
__global__ void k_batch_candidates(const float* X0, const float* X1, const batch_ctl* __restrict__ ctl,
                                   uint32_t V, const uint32_t* __restrict__ slots, uint32_t n,
                                   const uint32_t* __restrict__ thr_bin, uint32_t cap,
                                   unsigned int* __restrict__ count, uint32_t* __restrict__ cid,
                                   float* __restrict__ cval) {
  const float* X = ctl->cur ? X1 : X0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t i = 0; i < n; ++i) {
      const float f = X[v * BK + slots[i]];
      if (f > 0.f && (__float_as_uint(f) >> 20) >= thr_bin[i]) {
        const unsigned int p = atomicAdd(&count[i], 1u);
        if (p < cap) {
          cid[(uint64_t)i * cap + p] = (uint32_t)v;
          cval[(uint64_t)i * cap + p] = f;
        }
      }
    }
}

}  // namespace pe

using namespace pe;

// ================================================================================
// host side
// ================================================================================

struct pe_ppr_batch_s {
  pe_graph_s* g = nullptr;
  pe_ctx_s* ctx = nullptr;  // the graph's context; destroy uses it, not the graph (which may be gone)
  pe_ppr_batch_opts_t opts{};
  dbuf<float> X0, X1;
  dbuf<float> P0, P1;  // packed rows [V][BATCH_PACKED_W], same parity as X0 / X1
  dbuf<uint64_t> seedmask;
  dbuf<bt_task> tasks;
  dbuf<bt_long> longs;
  dbuf<float2> partial;
  dbuf<double> part_delta;
  dbuf<batch_ctl> ctl;
  dbuf<uint32_t> status, iters;
  dbuf<double> init_dist, last_delta;
  uint64_t ntasks = 0, nlong = 0, nchunks = 0;
  unsigned task_grid = 0, fix_grid = 0;
  std::vector<std::vector<uint32_t>> slot_seeds;  // internal ids per slot
  std::vector<uint32_t> host_status;
  stream_graph sg;  // the captured iterations, replayed by step()
  uint32_t graph_iters = 0;

  void build_tasks(uint32_t C) {
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t V = g->V;
    std::vector<uint64_t> cp((uint64_t)V + 1);
    PE_HIP(hipMemcpyAsync(cp.data(), g->col_ptr.p, ((uint64_t)V + 1) * 8, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    std::vector<bt_task> t;
    std::vector<bt_long> l;
    uint64_t chunk = 0;
    uint32_t v = 0;
    const uint32_t max_group = PE_BATCH_GROUP;
    while (v < V) {
      const uint64_t deg = cp[v + 1] - cp[v];
      if (deg > C) {
        // sqrt-sized chunks keep both the chunk work and the fix-up short
        uint64_t cs = std::max<uint64_t>(C, (uint64_t)std::ceil(std::sqrt((double)deg)));
        cs = (cs + 63) / 64 * 64;
        const uint32_t n = (uint32_t)((deg + cs - 1) / cs);
        PE_REQUIRE(chunk + n < 0xFFFFFFFFull, PE_ERR_RANGE, "too many chunks");
        bt_long lc{v, (uint32_t)chunk, n};
        l.push_back(lc);
        for (uint32_t k = 0; k < n; ++k) {
          bt_task tk{};
          tk.v0 = v;
          tk.v1 = v + 1;
          tk.e0 = cp[v] + (uint64_t)k * cs;
          tk.e1 = std::min<uint64_t>(cp[v + 1], tk.e0 + cs);
          tk.chunk = (uint32_t)(chunk + k);
          t.push_back(tk);
        }
        chunk += n;
        ++v;
        continue;
      }
      bt_task tk{};
      tk.v0 = v;
      tk.e0 = cp[v];
      tk.chunk = 0xFFFFFFFFu;
      uint64_t edges = 0;
      while (v < V) {
        const uint64_t d = cp[v + 1] - cp[v];
        if (d > C || (v > tk.v0 && (edges + d > C || v - tk.v0 >= max_group))) break;
        edges += d;
        ++v;
      }
      tk.v1 = v;
      tk.e1 = cp[v];
      t.push_back(tk);
    }
    ntasks = t.size();
    nlong = l.size();
    nchunks = chunk;
    tasks.alloc(ctx, std::max<uint64_t>(1, ntasks));
    longs.alloc(ctx, std::max<uint64_t>(1, nlong));
    partial.alloc(ctx, std::max<uint64_t>(1, nchunks) * BK);
    if (ntasks) PE_HIP(hipMemcpyAsync(tasks.p, t.data(), ntasks * sizeof(bt_task), hipMemcpyHostToDevice, s));
    if (nlong) PE_HIP(hipMemcpyAsync(longs.p, l.data(), nlong * sizeof(bt_long), hipMemcpyHostToDevice, s));
    PE_HIP(hipStreamSynchronize(s));
  }

  void launch_iteration() {
    hipStream_t s = g->ctx->stream;
    {
      ktime_scope kt(g->ctx, PE_KT_PPR_BLOCK);
      // the plan picks one of the four pass kernels; the others return at once
      k_batch_plan<<<1, BK, 0, s>>>(ctl.p, status.p);
      k_batch_tasks<4, true><<<task_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, g->od.p, seedmask.p, tasks.p,
                                                       ntasks, X0.p, X1.p, P0.p, P1.p, ctl.p, status.p, init_dist.p,
                                                       opts.damping, partial.p, part_delta.p);
      k_batch_tasks<4, false><<<task_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, g->od.p, seedmask.p, tasks.p,
                                                        ntasks, X0.p, X1.p, P0.p, P1.p, ctl.p, status.p, init_dist.p,
                                                        opts.damping, partial.p, part_delta.p);
      k_batch_tasks<2, false><<<task_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, g->od.p, seedmask.p, tasks.p,
                                                        ntasks, X0.p, X1.p, P0.p, P1.p, ctl.p, status.p, init_dist.p,
                                                        opts.damping, partial.p, part_delta.p);
      k_batch_tasks<1, false><<<task_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, g->od.p, seedmask.p, tasks.p,
                                                        ntasks, X0.p, X1.p, P0.p, P1.p, ctl.p, status.p, init_dist.p,
                                                        opts.damping, partial.p, part_delta.p);
    }
    k_batch_fixup<<<fix_grid, 256, 0, s>>>(g->od.p, seedmask.p, longs.p, nlong, X0.p, X1.p, P0.p, P1.p, ctl.p, status.p,
                                           init_dist.p, opts.damping, partial.p,
                                           part_delta.p + (uint64_t)task_grid * BK);
    k_batch_finish<<<1, FIN_GROUPS * BK, 0, s>>>(part_delta.p, task_grid + fix_grid, ctl.p, status.p, iters.p,
                                    last_delta.p, opts.tol, opts.max_iters);
    PE_HIP(hipGetLastError());
  }

  void step(uint32_t n) {
    if (n == 0) return;
    hipStream_t s = g->ctx->stream;
    if (g->ctx->timing) {
      for (uint32_t i = 0; i < n; ++i) launch_iteration();
      return;
    }
    if (!sg.exec || graph_iters != n) {
      graph_iters = 0;
      sg.begin(s);  // waits for launches of the previous executable first
      for (uint32_t i = 0; i < n; ++i) launch_iteration();
      sg.end();
      graph_iters = n;
    }
    sg.launch();
  }
};

namespace {

void check_slot(pe_ppr_batch_s* b, uint32_t slot) {
  PE_REQUIRE(slot < b->opts.k, PE_ERR_RANGE, "slot out of range");
}

// Slots given by the caller -> device list (all must be DONE)
dbuf<uint32_t> device_slots(pe_ppr_batch_s* b, const uint32_t* slots, uint32_t n) {
  dbuf<uint32_t> d(b->g->ctx, std::max<uint32_t>(1, n));
  PE_HIP(hipMemcpyAsync(d.p, slots, (size_t)n * 4, hipMemcpyHostToDevice, b->g->ctx->stream));
  return d;
}

void refresh_status(pe_ppr_batch_s* b) {
  hipStream_t s = b->g->ctx->stream;
  b->host_status.resize(BK);
  PE_HIP(hipMemcpyAsync(b->host_status.data(), b->status.p, BK * 4, hipMemcpyDeviceToHost, s));
  PE_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

int pe_ppr_batch_create(pe_graph_t* g, const pe_ppr_batch_opts_t* opts, pe_ppr_batch_t** out) {
  if (!g || !opts || !out) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(opts->k == 64, PE_ERR_INVALID, "k must be 64 (one column per lane of a wave)");
    PE_REQUIRE(opts->max_iters >= 1, PE_ERR_INVALID, "max_iters must be >= 1");
    PE_REQUIRE(opts->damping > 0 && opts->damping < 1, PE_ERR_INVALID, "damping must be in (0, 1)");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    auto b = std::make_unique<pe_ppr_batch_s>();
    b->g = g;
    b->ctx = g->ctx;
    b->opts = *opts;
    if (!b->opts.task_edges) b->opts.task_edges = 4096;
    const uint64_t cells = (uint64_t)std::max<uint32_t>(1, g->V) * BK;
    b->X0.alloc(ctx, cells);
    b->X1.alloc(ctx, cells);
    const uint64_t pcells = (uint64_t)std::max<uint32_t>(1, g->V) * BATCH_PACKED_W;
    b->P0.alloc(ctx, pcells);
    b->P1.alloc(ctx, pcells);
    PE_HIP(hipMemsetAsync(b->P0.p, 0, pcells * 4, s));
    PE_HIP(hipMemsetAsync(b->P1.p, 0, pcells * 4, s));
    b->seedmask.alloc(ctx, std::max<uint32_t>(1, g->V));
    PE_HIP(hipMemsetAsync(b->X0.p, 0, cells * 4, s));
    PE_HIP(hipMemsetAsync(b->X1.p, 0, cells * 4, s));
    PE_HIP(hipMemsetAsync(b->seedmask.p, 0, (uint64_t)std::max<uint32_t>(1, g->V) * 8, s));
    b->build_tasks(b->opts.task_edges);
    b->task_grid = persistent_grid(ctx, (b->ntasks + 3) / 4, 8);
    b->fix_grid = persistent_grid(ctx, std::max<uint64_t>(1, (b->nlong + 3) / 4), 2);
    b->part_delta.alloc(ctx, (uint64_t)(b->task_grid + b->fix_grid) * BK);
    b->ctl.alloc(ctx, 1);
    b->status.alloc(ctx, BK);
    b->iters.alloc(ctx, BK);
    b->init_dist.alloc(ctx, BK);
    b->last_delta.alloc(ctx, BK);
    batch_ctl c0{};
    c0.pk_n = 0xFFFFFFFFu;  // no packed pass yet
    auto limit = [](uint32_t v, uint32_t widest) {
      return v == 0 ? widest : v == PE_BATCH_PLAN_OFF ? 0u : std::min(v, widest);
    };
    c0.max_packed = limit(opts->plan_packed, (uint32_t)BATCH_PACKED_W);
    c0.max_edges4 = limit(opts->plan_edges4, 16u);
    c0.max_edges2 = limit(opts->plan_edges2, 32u);
    PE_HIP(hipMemcpyAsync(b->ctl.p, &c0, sizeof(batch_ctl), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemsetAsync(b->status.p, 0, BK * 4, s));
    PE_HIP(hipMemsetAsync(b->iters.p, 0, BK * 4, s));
    PE_HIP(hipMemsetAsync(b->init_dist.p, 0, BK * 8, s));
    PE_HIP(hipMemsetAsync(b->last_delta.p, 0, BK * 8, s));
    PE_HIP(hipStreamSynchronize(s));
    b->slot_seeds.resize(BK);
    b->host_status.assign(BK, COL_EMPTY);
    *out = b.release();
  });
}

int pe_ppr_batch_destroy(pe_ppr_batch_t* b) {
  if (!b) return PE_ERR_INVALID;
  (void)hipStreamSynchronize(b->ctx->stream);
  delete b;
  return PE_OK;
}

int pe_ppr_batch_bytes(const pe_ppr_batch_t* b, uint64_t* bytes) {
  if (!b || !bytes) return PE_ERR_INVALID;
  *bytes = b->X0.bytes() + b->X1.bytes() + b->P0.bytes() + b->P1.bytes() + b->seedmask.bytes() + b->tasks.bytes() + b->longs.bytes() +
           b->partial.bytes() + b->part_delta.bytes();
  return PE_OK;
}

int pe_ppr_batch_admit(pe_ppr_batch_t* b, uint32_t slot, const uint32_t* seeds, uint32_t nseeds) {
  if (!b || !seeds || nseeds == 0) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] {
    check_slot(b, slot);
    refresh_status(b);
    PE_REQUIRE(b->host_status[slot] == COL_EMPTY, PE_ERR_STATE, "slot is not free (release it first)");
    pe_graph_s* g = b->g;
    for (uint32_t i = 0; i < nseeds; ++i) PE_REQUIRE(seeds[i] < g->V, PE_ERR_RANGE, "seed vertex >= V");
    std::vector<uint32_t> ids(nseeds);
    to_internal_ids(g, seeds, nseeds, ids.data());
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    hipStream_t s = g->ctx->stream;
    dbuf<uint32_t> d(g->ctx, ids.size());
    PE_HIP(hipMemcpyAsync(d.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, s));
    k_batch_seed_bits<<<grid_for(ids.size(), 256, 64), 256, 0, s>>>(b->seedmask.p, d.p, (uint32_t)ids.size(), slot, 1);
    k_batch_set_slot<<<1, 1, 0, s>>>(b->status.p, b->init_dist.p, b->iters.p, b->last_delta.p, slot, COL_LOADING,
                                     1.0 / (double)ids.size());
    PE_HIP(hipGetLastError());
    PE_HIP(hipStreamSynchronize(s));
    b->slot_seeds[slot] = std::move(ids);
    b->host_status[slot] = COL_LOADING;
  });
}

int pe_ppr_batch_release(pe_ppr_batch_t* b, uint32_t slot) {
  if (!b) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] {
    check_slot(b, slot);
    hipStream_t s = b->g->ctx->stream;
    auto& ids = b->slot_seeds[slot];
    if (!ids.empty()) {
      dbuf<uint32_t> d(b->g->ctx, ids.size());
      PE_HIP(hipMemcpyAsync(d.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, s));
      k_batch_seed_bits<<<grid_for(ids.size(), 256, 64), 256, 0, s>>>(b->seedmask.p, d.p, (uint32_t)ids.size(), slot, 0);
      k_batch_zero_slot<<<grid_for(ids.size(), 256, 64), 256, 0, s>>>(b->X0.p, b->X1.p, b->P0.p, b->P1.p,
                                                                     b->g->col_ptr.p, b->seedmask.p, d.p,
                                                                     (uint32_t)ids.size(), slot);
      ids.clear();
    }
    k_batch_set_slot<<<1, 1, 0, s>>>(b->status.p, b->init_dist.p, b->iters.p, b->last_delta.p, slot, COL_EMPTY, 0.0);
    PE_HIP(hipGetLastError());
    PE_HIP(hipStreamSynchronize(s));
    b->host_status[slot] = COL_EMPTY;
  });
}

int pe_ppr_batch_step(pe_ppr_batch_t* b, uint32_t iterations) {
  if (!b) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] { b->step(iterations); });
}

int pe_ppr_batch_poll(pe_ppr_batch_t* b, uint32_t* status, uint32_t* iters, double* last_delta) {
  if (!b) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] {
    hipStream_t s = b->g->ctx->stream;
    if (status) PE_HIP(hipMemcpyAsync(status, b->status.p, BK * 4, hipMemcpyDeviceToHost, s));
    if (iters) PE_HIP(hipMemcpyAsync(iters, b->iters.p, BK * 4, hipMemcpyDeviceToHost, s));
    if (last_delta) PE_HIP(hipMemcpyAsync(last_delta, b->last_delta.p, BK * 8, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    if (status) b->host_status.assign(status, status + BK);
  });
}

int pe_ppr_batch_results(pe_ppr_batch_t* b, const uint32_t* slots, uint32_t n, double* ranks) {
  if (!b || (n && (!slots || !ranks))) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] {
    if (!n) return;
    refresh_status(b);
    for (uint32_t i = 0; i < n; ++i) {
      check_slot(b, slots[i]);
      PE_REQUIRE(b->host_status[slots[i]] == COL_DONE, PE_ERR_STATE, "slot has no finished result");
    }
    pe_graph_s* g = b->g;
    hipStream_t s = g->ctx->stream;
    const uint32_t V = g->V;
    auto ds = device_slots(b, slots, n);
    dbuf<float> dense(g->ctx, (uint64_t)n * std::max<uint32_t>(1, V));
    k_batch_extract<<<grid_for(V, 256, 8192), 256, 0, s>>>(b->X0.p, b->X1.p, b->ctl.p, V, ds.p, n, dense.p);
    PE_HIP(hipGetLastError());
    std::vector<float> tmp(V);
    for (uint32_t i = 0; i < n; ++i) {
      download_external(g, dense.p + (uint64_t)i * V, tmp.data(), 4);
      for (uint32_t v = 0; v < V; ++v) ranks[(uint64_t)i * V + v] = tmp[v];
    }
  });
}

int pe_ppr_batch_topk(pe_ppr_batch_t* b, const uint32_t* slots, uint32_t n, uint32_t k, uint32_t* ids,
                      float* vals) {
  if (!b || (n && (!slots || !ids || !vals)) || k == 0) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] {
    if (!n) return;
    refresh_status(b);
    for (uint32_t i = 0; i < n; ++i) {
      check_slot(b, slots[i]);
      PE_REQUIRE(b->host_status[slots[i]] == COL_DONE, PE_ERR_STATE, "slot has no finished result");
    }
    pe_graph_s* g = b->g;
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t V = g->V;
    auto ds = device_slots(b, slots, n);
    dbuf<unsigned int> hist(ctx, (uint64_t)n * 2048);
    PE_HIP(hipMemsetAsync(hist.p, 0, (uint64_t)n * 2048 * 4, s));
    k_batch_hist<<<persistent_grid(ctx, grid_for(V, 256), 4), 256, 0, s>>>(b->X0.p, b->X1.p, b->ctl.p, V, ds.p, n, hist.p);
    std::vector<unsigned int> h((uint64_t)n * 2048);
    PE_HIP(hipMemcpyAsync(h.data(), hist.p, h.size() * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> thr(n);
    uint32_t need_cap = k;
    for (uint32_t i = 0; i < n; ++i) {
      uint64_t acc = 0;
      int bin = 2047;
      for (; bin > 0; --bin) {
        acc += h[(uint64_t)i * 2048 + bin];
        if (acc >= k) break;
      }
      thr[i] = (uint32_t)std::max(bin, 1);  // bin 0 holds only denormals; zeros are excluded
      uint64_t cnt = 0;
      for (int b2 = 2047; b2 >= (int)thr[i]; --b2) cnt += h[(uint64_t)i * 2048 + b2];
      need_cap = std::max<uint32_t>(need_cap, (uint32_t)std::min<uint64_t>(cnt, 1u << 24));
    }
    dbuf<uint32_t> dthr(ctx, n);
    PE_HIP(hipMemcpyAsync(dthr.p, thr.data(), n * 4, hipMemcpyHostToDevice, s));
    dbuf<unsigned int> count(ctx, n);
    PE_HIP(hipMemsetAsync(count.p, 0, n * 4, s));
    dbuf<uint32_t> cid(ctx, (uint64_t)n * need_cap);
    dbuf<float> cval(ctx, (uint64_t)n * need_cap);
    k_batch_candidates<<<grid_for(V, 256, 8192), 256, 0, s>>>(b->X0.p, b->X1.p, b->ctl.p, V, ds.p, n, dthr.p,
                                                               need_cap, count.p, cid.p, cval.p);
    PE_HIP(hipGetLastError());
    std::vector<unsigned int> hc(n);
    std::vector<uint32_t> hid((uint64_t)n * need_cap);
    std::vector<float> hval((uint64_t)n * need_cap);
    PE_HIP(hipMemcpyAsync(hc.data(), count.p, n * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(hid.data(), cid.p, hid.size() * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(hval.data(), cval.p, hval.size() * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    // input ids for the tie-break (rank desc, id asc: trustrank.hpp:109-112)
    std::vector<uint32_t> perm;
    if (g->relabeled) {
      perm.resize(V);
      PE_HIP(hipMemcpyAsync(perm.data(), g->perm.p, (uint64_t)V * 4, hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
    }
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t m = std::min<uint32_t>(hc[i], need_cap);
      std::vector<std::pair<float, uint32_t>> c(m);
      for (uint32_t q = 0; q < m; ++q) {
        const uint32_t iv = hid[(uint64_t)i * need_cap + q];
        c[q] = {hval[(uint64_t)i * need_cap + q], g->relabeled ? perm[iv] : iv};
      }
      std::sort(c.begin(), c.end(), [](const std::pair<float, uint32_t>& a, const std::pair<float, uint32_t>& bb) {
        return a.first > bb.first || (a.first == bb.first && a.second < bb.second);
      });
      // fewer positive entries than k: the remaining places go to the smallest
      // ids with rank 0, as the reference's sort would order them
      std::vector<uint8_t> taken;
      uint32_t next_zero = 0;
      for (uint32_t q = 0; q < k; ++q) {
        if (q < c.size()) {
          ids[(uint64_t)i * k + q] = c[q].second;
          vals[(uint64_t)i * k + q] = c[q].first;
        } else {
          if (taken.empty()) {
            taken.assign(V, 0);
            for (auto& pr : c) taken[pr.second] = 1;
          }
          while (next_zero < V && taken[next_zero]) ++next_zero;
          ids[(uint64_t)i * k + q] = next_zero < V ? next_zero : 0xFFFFFFFFu;
          vals[(uint64_t)i * k + q] = 0.0f;
          ++next_zero;
        }
      }
    }
  });
}

}  // extern "C"
