// Binned (partition-centric) pull: the aggregate_message pass of the FP32 rank
// iterations (message_passing.hpp:88-144 driven by pagerank.hpp:104-140 and its
// weighted / personalized variants) with no random global gathers.
//
// The merge-path gather (pull.hip) reads x[u] once per edge at random addresses;
// on a 2^27-vertex graph those 4-byte gathers are bound by L1/L2 miss handling,
// not by HBM bandwidth.  Here every random access goes to LDS:
//   scatter: a block stages one chunk of x (2^src_bits sources) in LDS and writes
//            the contribution x[u] of every edge leaving the chunk into the run of
//            its destination bin (runs are contiguous, so reads and writes stream);
//   gather:  a block streams one bin's contributions (2^dst_bits destinations) and
//            adds them into LDS accumulators, then applies restart, damping and the
//            out-degree division to the whole bin (the apply of pull.hip, fused).
// (Beamer, Asanovic, Patterson, "Reducing PageRank communication via propagation
// blocking", IPDPS 2017; Lakhotia, Kannan, Prasanna, "Accelerating PageRank using
// partition-centric processing", USENIX ATC 2018.)
//
// Sums are exact and order-independent.  Each contribution (non-negative, finite)
// becomes the integer floor(x * 2^F) (the FP64 products w * x of weighted graphs are
// formed exactly first), and these integers are added with 64-bit LDS atomics into one
// word per destination.  For unweighted iterations the scatter stages each source as
// a 32-bit code of its term (pb_fixed.hpp), so the gather's per-entry work is a mask
// and two shifts; weighted iterations carry the FP32 values.  F is fixed per iteration from a bound M on the total of all
// contributions (pb_frac_bits: M * 2^F <= 2^61): for the rank iterations M is the
// rank mass, at most V for PageRank (ranks start at 1 and the restart adds
// (1 - d) per vertex while the damped flow keeps at most d of the mass) and at most 1
// for PPR, so no sum can reach 2^64.  Integer addition is associative: every sum is
// independent of order, launch shape and layout parameters, and results are
// bit-reproducible.  The truncation moves each term by less than 2^-F (2^-33 for
// PageRank on 2^27 vertices, 2^-60 for PPR).

// Layout (bin order): every CSC entry (v <- u) gets the key
//   bin = v >> dst_bits | chunk = u >> src_bits | u & smask | v & dmask
// and the entries are radix-sorted by it.  An update is one (bin, source) pair: the
// scatter writes the source's contribution once per update, not once per edge
// (partition-centric, as in Lakhotia et al.), and every entry carries its destination
// plus a flag bit marking the first entry of a new update; the update count before
// each 512-entry step plus a prefix count of the flags lets the gather find each
// entry's update.  Scatter and gather blocks take their units from a queue.  A run is the updates of one (bin, chunk)
// pair.  The scatter walks run records (runs cut to at most run_cap updates) grouped
// by chunk, in units of about one CU's share of the updates; the gather walks bins, split into units of at most 2^unit_log2 entries
// (units of a split bin meet in global 64-bit accumulators, applied by a last pass).
#include <algorithm>
#include <numeric>
#include <vector>

#include "pb_fixed.hpp"
#include "pe_internal.hpp"
#include "pull.hpp"

namespace pe {

static constexpr int PB_THREADS = 1024;  // layout build kernels
#ifndef PB_SCATTER_THREADS
#define PB_SCATTER_THREADS 1024
#endif
#ifndef PB_GATHER_THREADS
#define PB_GATHER_THREADS 1024
#endif
#ifndef PB_MAX_SRC_BITS
#define PB_MAX_SRC_BITS 15  // 128 KiB of FP32 sources in LDS
#endif
static constexpr int PB_MAX_DST_BITS = 14;  // 128 KiB of 64-bit accumulators in LDS
#ifndef PB_DST_BITS
#define PB_DST_BITS 14  // destinations per bin (default)
#endif
static constexpr uint32_t PB_SEG = 1u << 16;      // entries per segment of the run scan
#ifndef PB_SCATTER_RECORDS
#define PB_SCATTER_RECORDS 8  // run records per wave at a time in the scatter
#endif


// Units are taken from a queue (q[0]: next unit, q[1]: blocks finished): the unit sizes
// vary, and a block that has finished its unit takes the next one.  The last block to
// finish sets the queue back to zero for the next launch.  The unit index is read back
// as a wave-uniform value (every thread sees the same one), so the compiler treats the
// loop's exit test as uniform and keeps the loop whole around the block's barriers.
// A loop body must end with a barrier after any thread-0-only work: the compiler may
// otherwise merge that work with the next take across the back edge and let the other
// threads reach the next unit's barrier ahead of thread 0.
__device__ __forceinline__ uint32_t pb_take_unit(uint32_t* q, uint32_t* s_unit) {
  if (threadIdx.x == 0) *s_unit = atomicAdd(&q[0], 1u);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*s_unit);
}

__device__ __forceinline__ void pb_queue_done(uint32_t* q) {
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&q[1], 1u) == gridDim.x - 1) {
      q[0] = 0u;
      q[1] = 0u;
    }
  }
}

// Adds one converted term into a destination's accumulator word.
__device__ __forceinline__ void acc_add(unsigned long long* acc, uint64_t v) {
  if (v) atomicAdd(acc, (unsigned long long)v);
}

// The same for a coded term (fixed_code_f32): zero exactly when its significand field is,
// since the left shift loses no bits.
__device__ __forceinline__ void acc_add_code(unsigned long long* acc, uint32_t c) {
  const uint32_t m = c & 0xFFFFFFu;
  if (m) atomicAdd(acc, (unsigned long long)m << (c >> 24));
}

// ---- scatter ----------------------------------------------------------------------------
// A wave takes RB run records at a time (a record holds at most 256 updates, one pass of
// 4 x 64 lanes).  The chain per record group -- descriptors, then the sources of its
// updates, then LDS lookups and stores -- is software-pipelined across groups: the
// descriptors are loaded two groups ahead and the sources one group ahead, so a wave
// always has the next group's loads in flight while it stores the current one.
template <int RB>
struct pb_rec_group {
  uint64_t s[RB];
  uint32_t len[RB];
};

template <int RB>
__device__ __forceinline__ void pb_load_group(pb_rec_group<RB>& g, uint64_t r, uint64_t r1,
                                              const uint64_t* __restrict__ rec_start,
                                              const uint32_t* __restrict__ rec_len) {
#pragma unroll
  for (int q = 0; q < RB; ++q) {
    const bool ok = r + q < r1;
    g.s[q] = ok ? rec_start[r + q] : 0ull;
    g.len[q] = ok ? rec_len[r + q] : 0u;
  }
}

template <int RB>
__device__ __forceinline__ void pb_load_sources(const pb_rec_group<RB>& g, const uint16_t* __restrict__ src_upd,
                                                int lane, uint32_t (&sl)[RB][4]) {
#pragma unroll
  for (int q = 0; q < RB; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t k = j * 64 + lane;
      sl[q][j] = k < g.len[q] ? (uint32_t)src_upd[g.s[q] + k] : 0u;
    }
}

// CODED (unweighted iterations): the chunk is staged as 32-bit fixed-point codes
// (fixed_code_f32 at F = frac), one conversion per source, and the contributions are
// codes; otherwise (weighted: the gather forms w * x exactly) they are the FP32 values.
template <int RB, int THREADS, bool CODED>
__global__ __launch_bounds__(THREADS) void k_pb_scatter(
    const pb_unit1* __restrict__ units, uint32_t nunits, const uint64_t* __restrict__ rec_start,
    const uint32_t* __restrict__ rec_len, const uint16_t* __restrict__ src_upd, uint32_t V, int src_bits,
    int frac, const float* buf0, const float* buf1, const iter_ctl* __restrict__ ctl,
    uint32_t* __restrict__ values, uint32_t* __restrict__ queue) {
  extern __shared__ uint32_t pb_scatter_lds[];  // the chunk's 2^src_bits sources, then the unit slot
  uint32_t* xs = pb_scatter_lds;
  const float* x = buf0;
  if (ctl) {
    if (ctl->done) return;
    if (ctl->cur) x = buf1;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int WAVES = THREADS / 64;
  constexpr uint64_t RSTEP = (uint64_t)WAVES * RB;
  uint32_t* s_unit = reinterpret_cast<uint32_t*>(pb_scatter_lds + (1u << src_bits));
  for (;;) {
    const uint32_t ui = pb_take_unit(queue, s_unit);
    if (ui >= nunits) break;
    const pb_unit1 u = units[ui];
    const uint64_t base = (uint64_t)u.chunk << src_bits;
    const uint32_t n = (uint32_t)std::min<uint64_t>(1ull << src_bits, (uint64_t)V - base);
    uint64_t r = u.r0 + (uint64_t)wave * RB;
    // the first groups' loads go out before the chunk is staged
    pb_rec_group<RB> cur, nxt;
    pb_load_group<RB>(cur, r, u.r1, rec_start, rec_len);
    pb_load_group<RB>(nxt, r + RSTEP, u.r1, rec_start, rec_len);
    uint32_t sl[RB][4];
    pb_load_sources<RB>(cur, src_upd, lane, sl);
    if ((n & 3u) == 0) {  // base is a multiple of 16 sources, so x + base is 16-byte aligned
      const float4* x4 = reinterpret_cast<const float4*>(x + base);
      uint4* s4 = reinterpret_cast<uint4*>(xs);
      for (uint32_t i = tid; i < (n >> 2); i += THREADS) {
        const float4 v = x4[i];
        s4[i] = CODED ? make_uint4(fixed_code_f32(v.x, frac), fixed_code_f32(v.y, frac), fixed_code_f32(v.z, frac),
                                   fixed_code_f32(v.w, frac))
                      : make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z),
                                   __float_as_uint(v.w));
      }
    } else {
      for (uint32_t i = tid; i < n; i += THREADS)
        xs[i] = CODED ? fixed_code_f32(x[base + i], frac) : __float_as_uint(x[base + i]);
    }
    __syncthreads();
    for (; r < u.r1; r += RSTEP) {
      pb_rec_group<RB> nn;
      pb_load_group<RB>(nn, r + 2 * RSTEP, u.r1, rec_start, rec_len);
      uint32_t sl2[RB][4];
      pb_load_sources<RB>(nxt, src_upd, lane, sl2);
#pragma unroll
      for (int q = 0; q < RB; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t k = j * 64 + lane;
          if (k < cur.len[q]) values[cur.s[q] + k] = xs[sl[q][j]];
        }
      cur = nxt;
      nxt = nn;
#pragma unroll
      for (int q = 0; q < RB; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) sl[q][j] = sl2[q][j];
    }
    __syncthreads();
  }
  pb_queue_done(queue);
}

// ---- gather + apply -------------------------------------------------------------------------
// A wave step is PB_STEP = 512 aligned entries, 8 per lane (one 16-byte load of
// destination words); a wave works on GU steps at a time, all their loads in flight
// together.  An entry's update index is step_base[step] (updates before the step) plus a
// wave-wide prefix count of the update flags, so no per-group count is read; each entry
// then loads its value by update index (entries of one update read one address) and adds
// its term into its destination's LDS word.  Steps cover the unit's range rounded out to
// whole steps; entries outside the unit count for the prefix but add nothing.
static constexpr uint32_t PB_STEP = 512;
#ifndef PB_GATHER_STEPS
#define PB_GATHER_STEPS 4  // wave steps in flight per wave
#endif
#ifndef PB_GATHER_STEPS_W
#define PB_GATHER_STEPS_W 2  // the same for weighted graphs (two more float4 per lane and step)
#endif

// Exclusive prefix sum over the wave of per-lane counts below 16: the count's four bits
// are summed separately, each as the number of lanes below this one whose bit is set
// (a ballot and the two mbcnt halves) -- no cross-lane data movement through LDS.
__device__ __forceinline__ uint32_t wave_excl_scan16(uint32_t v) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const uint64_t m = __ballot((v >> b) & 1u);
    r += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << b;
  }
  return r;
}

template <bool WEIGHTED, int GU, int THREADS>
__global__ __launch_bounds__(THREADS) void k_pb_gather(
    const pb_unit2* __restrict__ units, uint32_t nunits, const uint32_t* __restrict__ values,
    const float* __restrict__ w, const uint16_t* __restrict__ dst_local, const uint32_t* __restrict__ step_base,
    uint32_t V, int dst_bits, int frac,
    float* buf0, float* buf1, const uint32_t* __restrict__ od, rank_args a, const iter_ctl* __restrict__ ctl,
    unsigned long long* __restrict__ acc, double* __restrict__ part_delta, double* __restrict__ part_rsum,
    uint32_t* __restrict__ queue) {
  // 2^dst_bits accumulators, then the reduction, then the unit slot
  extern __shared__ unsigned long long pb_gather_lds[];
  constexpr int PB_WAVES = THREADS / 64;
  unsigned long long* sacc = pb_gather_lds;
  double* s_red = reinterpret_cast<double*>(pb_gather_lds + (1u << dst_bits));
  uint32_t* s_unit = reinterpret_cast<uint32_t*>(s_red + 2 * PB_WAVES);
  bool final_iter = a.final_iter;
  const float* x = buf0;
  float* y = buf1;
  if (ctl) {
    if (ctl->done) return;
    final_iter = (ctl->t + 1 == ctl->iters);
    if (ctl->cur) {
      x = buf1;
      y = buf0;
    }
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t pd = 1u << dst_bits;
  double delta = 0.0, rsum = 0.0;
  for (;;) {
    const uint32_t ui = pb_take_unit(queue, s_unit);
    if (ui >= nunits) break;
    const pb_unit2 u = units[ui];
    for (uint32_t i = tid; i < pd; i += THREADS) sacc[i] = 0ull;
    __syncthreads();
    const uint64_t s0 = u.e0 / PB_STEP, s1 = (u.e1 + PB_STEP - 1) / PB_STEP;
    for (uint64_t sb = s0 + wave; sb < s1; sb += (uint64_t)PB_WAVES * GU) {
      uint4 dd[GU];
      uint32_t base[GU];
      float4 wa[GU], wb[GU];
      // Every step's loads are issued without a branch (a step past the unit loads the
      // unit's last step again, whose entries then fail the range test below and add
      // nothing), so the compiler counts the loads in flight and waits for each one only.
#pragma unroll
      for (int q = 0; q < GU; ++q) {
        const uint64_t st = std::min<uint64_t>(sb + (uint64_t)q * PB_WAVES, s1 - 1);
        dd[q] = reinterpret_cast<const uint4*>(dst_local + st * PB_STEP)[lane];
        base[q] = step_base[st];
        if (WEIGHTED) {
          const float4* wq = reinterpret_cast<const float4*>(w + st * PB_STEP) + 2 * lane;
          wa[q] = wq[0];
          wb[q] = wq[1];
        }
      }
      uint32_t xv[GU][8];  // contributions: fixed-point codes, or FP32 bits when WEIGHTED
#pragma unroll
      for (int q = 0; q < GU; ++q) {
        const uint32_t raw[8] = {dd[q].x & 0xFFFFu, dd[q].x >> 16, dd[q].y & 0xFFFFu, dd[q].y >> 16,
                                 dd[q].z & 0xFFFFu, dd[q].z >> 16, dd[q].w & 0xFFFFu, dd[q].w >> 16};
        uint32_t heads = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) heads += raw[k] >> 15;
        uint32_t idx = base[q] + wave_excl_scan16(heads) - 1u;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          idx += raw[k] >> 15;
          xv[q][k] = values[idx];
        }
      }
#pragma unroll
      for (int q = 0; q < GU; ++q) {
        const uint64_t st = sb + (uint64_t)q * PB_WAVES;  // past the unit: every entry fails the range test
        const uint64_t e0 = st * PB_STEP + 8u * (uint32_t)lane;
        float wv[8] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
        if (WEIGHTED) {
          wv[0] = wa[q].x; wv[1] = wa[q].y; wv[2] = wa[q].z; wv[3] = wa[q].w;
          wv[4] = wb[q].x; wv[5] = wb[q].y; wv[6] = wb[q].z; wv[7] = wb[q].w;
        }
        const uint32_t d[8] = {dd[q].x & 0x7FFFu, (dd[q].x >> 16) & 0x7FFFu, dd[q].y & 0x7FFFu,
                               (dd[q].y >> 16) & 0x7FFFu, dd[q].z & 0x7FFFu, (dd[q].z >> 16) & 0x7FFFu,
                               dd[q].w & 0x7FFFu, (dd[q].w >> 16) & 0x7FFFu};
        if (e0 >= u.e0 && e0 + 8 <= u.e1) {
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            if (WEIGHTED)
              acc_add(&sacc[d[k]], fixed64_f64((double)wv[k] * (double)__uint_as_float(xv[q][k]), frac));
            else
              acc_add_code(&sacc[d[k]], xv[q][k]);
          }
        } else {  // the unit's first or last group
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            if (e0 + k < u.e0 || e0 + k >= u.e1) continue;
            if (WEIGHTED)
              acc_add(&sacc[d[k]], fixed64_f64((double)wv[k] * (double)__uint_as_float(xv[q][k]), frac));
            else
              acc_add_code(&sacc[d[k]], xv[q][k]);
          }
        }
      }
    }
    __syncthreads();
    const uint64_t v0 = (uint64_t)u.bin << dst_bits;
    if (u.slot == 0xFFFFFFFFu) {
      for (uint32_t i = tid; i < pd; i += THREADS) {
        const uint64_t v = v0 + i;
        if (v >= V) break;
        delta += apply_row<float>(a, final_iter, (uint32_t)v, from_fixed64(sacc[i], frac), x, y, od, rsum);
      }
    } else {
      unsigned long long* ga = acc + ((uint64_t)u.slot << dst_bits);
      for (uint32_t i = tid; i < pd; i += THREADS)
        if (sacc[i]) atomicAdd(&ga[i], sacc[i]);
    }
    __syncthreads();
  }
  pb_queue_done(queue);
  block_reduce2<THREADS>(delta, rsum, s_red);
  if (tid == 0) {
    part_delta[blockIdx.x] = delta;
    part_rsum[blockIdx.x] = rsum;
  }
}

// Bins summed by several units: apply from the global accumulators, then clear them.
__global__ __launch_bounds__(256) void k_pb_apply_split(
    const uint32_t* __restrict__ split_bin, uint32_t nsplit, int dst_bits, int frac, uint32_t V,
    unsigned long long* __restrict__ acc, float* buf0, float* buf1, const uint32_t* __restrict__ od,
    rank_args a, const iter_ctl* __restrict__ ctl, double* __restrict__ part_delta,
    double* __restrict__ part_rsum) {
  __shared__ double s_red[8];
  bool final_iter = a.final_iter;
  const float* x = buf0;
  float* y = buf1;
  if (ctl) {
    if (ctl->done) return;
    final_iter = (ctl->t + 1 == ctl->iters);
    if (ctl->cur) {
      x = buf1;
      y = buf0;
    }
  }
  const uint32_t pd = 1u << dst_bits;
  const uint64_t n = (uint64_t)nsplit << dst_bits;
  double delta = 0.0, rsum = 0.0;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t slot = j >> dst_bits, i = j & (pd - 1);
    unsigned long long* ga = acc + (slot << dst_bits);
    const uint64_t sum = ga[i];
    ga[i] = 0ull;
    const uint64_t v = ((uint64_t)split_bin[slot] << dst_bits) + i;
    if (v < V) delta += apply_row<float>(a, final_iter, (uint32_t)v, from_fixed64(sum, frac), x, y, od, rsum);
  }
  block_reduce2<256>(delta, rsum, s_red);
  if (threadIdx.x == 0) {
    part_delta[blockIdx.x] = delta;
    part_rsum[blockIdx.x] = rsum;
  }
}

// Total contribution of a state x along the edges: sum over sources of od(u) * x(u)
// (weighted graphs: x(u), their weights sum to 1 per source).  A block's partial is
// -1 when some source value is negative or not finite.
__global__ __launch_bounds__(256) void k_pb_mass(const float* __restrict__ x, const uint32_t* __restrict__ od,
                                                 uint32_t V, int weighted, double* __restrict__ part) {
  __shared__ double s_red[8];
  double acc = 0.0, bad = 0.0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = od[v];
    if (!d) continue;
    const float xv = x[v];
    if (!(xv >= 0.0f) || isinf(xv)) bad = 1.0;
    else acc += weighted ? (double)xv : (double)d * (double)xv;
  }
  block_reduce2<256>(acc, bad, s_red);
  if (threadIdx.x == 0) part[blockIdx.x] = bad > 0.0 ? -1.0 : acc;
}

double binned_mass_bound(pe_graph_s* g, const float* x, bool weighted) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const unsigned G = grid_for(g->V, 256, 1024);
  dbuf<double> part(ctx, G);
  k_pb_mass<<<G, 256, 0, s>>>(x, g->od.p, g->V, weighted ? 1 : 0, part.p);
  PE_HIP(hipGetLastError());
  std::vector<double> h(G);
  PE_HIP(hipMemcpyAsync(h.data(), part.p, G * sizeof(double), hipMemcpyDeviceToHost, s));
  PE_HIP(hipStreamSynchronize(s));
  double total = 0.0;
  for (double p : h) {
    if (p < 0.0) return 0.0;
    total += p;
  }
  // FP32 weights are rounded per edge (their sum per source exceeds 1 by < 2^-23)
  return std::max(total * (1.0 + 0x1p-20), 0x1p-60);
}

// ---- layout build -----------------------------------------------------------------------------
// One thread per column; a hub column keeps its thread longer, once per build.
__global__ void k_pb_keys(const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ row_idx, uint32_t V,
                          int sb, int db, int cbits, uint64_t* __restrict__ keys) {
  const uint64_t smask = (1ull << sb) - 1ull, dmask = (1ull << db) - 1ull;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t hi_v = ((v >> db) << cbits), lo_v = v & dmask;
    const uint64_t e1 = col_ptr[v + 1];
    for (uint64_t e = col_ptr[v]; e < e1; ++e) {
      const uint64_t u = row_idx[e];
      keys[e] = ((hi_v | (u >> sb)) << (db + sb)) | ((u & smask) << db) | lo_v;
    }
  }
}

// First entry of a new update (bin, chunk, source) and of a new run (bin, chunk).
__device__ __forceinline__ bool pb_new(const uint64_t* keys, uint64_t e, int shift) {
  return e == 0 || (keys[e] >> shift) != (keys[e - 1] >> shift);
}

// Updates and run heads counted per segment of PB_SEG entries.
__global__ __launch_bounds__(1024) void k_pb_count(const uint64_t* __restrict__ keys, uint64_t nnz, int db, int shift,
                                                   uint32_t* __restrict__ seg_upd, uint32_t* __restrict__ seg_run) {
  __shared__ uint32_t s_cnt[2];
  if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t e0 = (uint64_t)blockIdx.x * PB_SEG, e1 = std::min<uint64_t>(nnz, e0 + PB_SEG);
  uint32_t cu = 0, cr = 0;
  for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    cu += pb_new(keys, e, db) ? 1u : 0u;
    cr += pb_new(keys, e, shift) ? 1u : 0u;
  }
  atomicAdd(&s_cnt[0], cu);
  atomicAdd(&s_cnt[1], cr);
  __syncthreads();
  if (threadIdx.x == 0) {
    seg_upd[blockIdx.x] = s_cnt[0];
    seg_run[blockIdx.x] = s_cnt[1];
  }
}

// In entry order: destination words with the update flag, update counts of aligned
// groups, the source of every update, and (update, entry, run id) of every run head.
// upd_off / run_off are the exclusive scans of seg_upd / seg_run.
__global__ __launch_bounds__(1024) void k_pb_write(const uint64_t* __restrict__ keys, uint64_t nnz, int sb, int db,
                                                   const uint64_t* __restrict__ upd_off,
                                                   const uint64_t* __restrict__ run_off,
                                                   uint16_t* __restrict__ dst_local, uint32_t* __restrict__ gbase,
                                                   uint16_t* __restrict__ src_upd, uint64_t* __restrict__ run_upd,
                                                   uint64_t* __restrict__ run_pos, uint64_t* __restrict__ run_id) {
  __shared__ uint32_t s_u[16], s_r[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int shift = sb + db;
  const uint64_t smask = (1ull << sb) - 1ull, dmask = (1ull << db) - 1ull;
  const uint64_t e0 = (uint64_t)blockIdx.x * PB_SEG, e1 = std::min<uint64_t>(nnz, e0 + PB_SEG);
  uint64_t ou = upd_off[blockIdx.x], orr = run_off[blockIdx.x];
  for (uint64_t b = e0; b < e1; b += 1024) {
    const uint64_t e = b + threadIdx.x;
    const bool in = e < e1;
    const bool fu = in && pb_new(keys, e, db);
    const bool fr = in && pb_new(keys, e, shift);
    const uint64_t mu = __ballot(fu), mr = __ballot(fr);
    const uint64_t below = (1ull << lane) - 1ull;
    if (lane == 0) {
      s_u[wave] = (uint32_t)__popcll(mu);
      s_r[wave] = (uint32_t)__popcll(mr);
    }
    __syncthreads();
    uint32_t bu = 0, br = 0, tu = 0, tr = 0;
    for (int i = 0; i < 16; ++i) {
      if (i < wave) {
        bu += s_u[i];
        br += s_r[i];
      }
      tu += s_u[i];
      tr += s_r[i];
    }
    if (in) {
      const uint64_t k = keys[e];
      const uint64_t u_ex = ou + bu + (uint64_t)__popcll(mu & below);  // updates before e
      dst_local[e] = (uint16_t)((k & dmask) | (fu ? 0x8000ull : 0ull));
      if ((e % PB_STEP) == 0) gbase[e / PB_STEP] = (uint32_t)u_ex;
      if (fu) src_upd[u_ex] = (uint16_t)((k >> db) & smask);
      if (fr) {
        const uint64_t r = orr + br + (uint64_t)__popcll(mr & below);
        run_upd[r] = u_ex;
        run_pos[r] = e;
        run_id[r] = k >> shift;
      }
    }
    ou += tu;
    orr += tr;
    __syncthreads();
  }
}

__global__ void k_pb_zero(unsigned long long* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

static int pb_bits_for(uint64_t n) {
  int b = 0;
  while ((1ull << b) < n) ++b;
  return b;
}

template <typename T>
static void upload(pe_ctx_s* ctx, dbuf<T>& d, const std::vector<T>& h) {
  d.alloc(ctx, std::max<size_t>(1, h.size()));
  if (!h.empty()) PE_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}

void binned_build(pe_graph_s* g, const pe_binned_opts_t* opts) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  // Gather units: by default the largest power of two up to half a CU's share of the
  // entries, so bins are split only where that evens out the work (RMAT-27 on 256 CUs:
  // 2^22 entries, 2.53-2.57 ms against 2.66 ms at 2^20 and 3.8 ms at 2^24; small graphs
  // get more units than bins).
  uint32_t unit_log2 = 10;
  {
    const uint64_t share = g->nnz / (uint64_t)std::max(1, ctx->num_cus) / 2;
    while (unit_log2 < 30 && (1ull << (unit_log2 + 1)) <= share) ++unit_log2;
  }
  pe_binned_opts_t o{(uint32_t)PB_MAX_SRC_BITS, (uint32_t)PB_DST_BITS, 256u, unit_log2};
  if (opts) {
    if (opts->src_bits) o.src_bits = opts->src_bits;
    if (opts->dst_bits) o.dst_bits = opts->dst_bits;
    if (opts->run_cap) o.run_cap = opts->run_cap;
    if (opts->unit_log2) o.unit_log2 = opts->unit_log2;
  }
  PE_REQUIRE(o.src_bits >= 4 && o.src_bits <= (uint32_t)PB_MAX_SRC_BITS, PE_ERR_INVALID, "src_bits must be in [4, 15]");
  PE_REQUIRE(o.dst_bits >= 4 && o.dst_bits <= (uint32_t)PB_MAX_DST_BITS, PE_ERR_INVALID, "dst_bits must be in [4, 14]");
  PE_REQUIRE(o.run_cap >= 64 && o.run_cap <= 256, PE_ERR_INVALID,
             "run_cap must be in [64, 256] (a scatter pass covers 256 updates of a record)");
  PE_REQUIRE(o.unit_log2 >= 10 && o.unit_log2 <= 30, PE_ERR_INVALID, "unit_log2 must be in [10, 30]");
  if (g->pb && g->pb->src_bits == o.src_bits && g->pb->dst_bits == o.dst_bits && g->pb->run_cap == o.run_cap &&
      g->pb->unit_log2 == o.unit_log2)
    return;
  g->pb.reset();
  auto L = std::make_unique<pb_layout>();
  const int sb = (int)o.src_bits, db = (int)o.dst_bits;
  L->src_bits = o.src_bits;
  L->dst_bits = o.dst_bits;
  L->run_cap = o.run_cap;
  L->unit_log2 = o.unit_log2;
  const uint32_t V = g->V;
  const uint64_t nnz = g->nnz;
  L->nchunks = (uint32_t)((V + (1ull << sb) - 1) >> sb);
  L->nbins = (uint32_t)((V + (1ull << db) - 1) >> db);
  const int cbits = pb_bits_for(L->nchunks), bbits = pb_bits_for(L->nbins);
  const int shift = sb + db;  // run id = key >> shift = bin << cbits | chunk
  const int end_bit = bbits + cbits + shift;
  PE_REQUIRE(end_bit <= 64, PE_ERR_RANGE, "binned key does not fit 64 bits");
  L->weighted = g->has_weights;

  std::vector<uint64_t> h_upd, h_pos, h_id;
  {
    dbuf<uint64_t> keys(ctx, std::max<uint64_t>(1, nnz));
    if (nnz) {
      k_pb_keys<<<grid_for(V, 256, 1u << 20), 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, V, sb, db, cbits, keys.p);
      PE_HIP(hipGetLastError());
    }
    if (L->weighted) {
      const uint64_t wpad = (nnz + PB_STEP - 1) / PB_STEP * PB_STEP + PB_STEP;
      L->w.alloc(ctx, wpad);
      k_pb_zero<<<grid_for(wpad / 2, 256, 65536), 256, 0, s>>>(reinterpret_cast<unsigned long long*>(L->w.p), wpad / 2);
      if (nnz) PE_HIP(hipMemcpyAsync(L->w.p, g->wn.p, nnz * sizeof(float), hipMemcpyDeviceToDevice, s));
      sort_keys(ctx, keys, &L->w, nnz, end_bit);
    } else {
      sort_keys(ctx, keys, nullptr, nnz, end_bit);
    }
    // padded to whole gather steps with zero words (no update flag, masked by range)
    const uint64_t padded = (nnz + PB_STEP - 1) / PB_STEP * PB_STEP + PB_STEP;
    L->dst_local.alloc(ctx, padded);
    k_pb_zero<<<grid_for(padded / 4, 256, 65536), 256, 0, s>>>(reinterpret_cast<unsigned long long*>(L->dst_local.p),
                                                               padded / 4);
    L->gbase.alloc(ctx, padded / PB_STEP + 1);
    if (nnz) {
      const uint64_t nseg = (nnz + PB_SEG - 1) / PB_SEG;
      PE_REQUIRE(nseg < 0x7FFFFFFFull, PE_ERR_RANGE, "too many entries for the update scan");
      dbuf<uint32_t> seg_upd(ctx, nseg), seg_run(ctx, nseg);
      k_pb_count<<<(unsigned)nseg, 1024, 0, s>>>(keys.p, nnz, db, shift, seg_upd.p, seg_run.p);
      PE_HIP(hipGetLastError());
      std::vector<uint32_t> hu(nseg), hr(nseg);
      PE_HIP(hipMemcpyAsync(hu.data(), seg_upd.p, nseg * 4, hipMemcpyDeviceToHost, s));
      PE_HIP(hipMemcpyAsync(hr.data(), seg_run.p, nseg * 4, hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      std::vector<uint64_t> ou(nseg), orr(nseg);
      uint64_t nupd = 0, nruns = 0;
      for (uint64_t i = 0; i < nseg; ++i) {
        ou[i] = nupd;
        orr[i] = nruns;
        nupd += hu[i];
        nruns += hr[i];
      }
      PE_REQUIRE(nupd < 0xFFFFFFFFull, PE_ERR_RANGE,
                 "binned layout: more than 2^32 (source, bin) updates; use the merge-path layout");
      L->nupdates = nupd;
      L->src_upd.alloc(ctx, nupd);
      dbuf<uint64_t> upd_off(ctx, nseg), run_off(ctx, nseg), run_upd(ctx, nruns), run_pos(ctx, nruns),
          run_id(ctx, nruns);
      PE_HIP(hipMemcpyAsync(upd_off.p, ou.data(), nseg * 8, hipMemcpyHostToDevice, s));
      PE_HIP(hipMemcpyAsync(run_off.p, orr.data(), nseg * 8, hipMemcpyHostToDevice, s));
      k_pb_write<<<(unsigned)nseg, 1024, 0, s>>>(keys.p, nnz, sb, db, upd_off.p, run_off.p, L->dst_local.p,
                                                 L->gbase.p, L->src_upd.p, run_upd.p, run_pos.p, run_id.p);
      PE_HIP(hipGetLastError());
      h_upd.resize(nruns);
      h_pos.resize(nruns);
      h_id.resize(nruns);
      PE_HIP(hipMemcpyAsync(h_upd.data(), run_upd.p, nruns * 8, hipMemcpyDeviceToHost, s));
      PE_HIP(hipMemcpyAsync(h_pos.data(), run_pos.p, nruns * 8, hipMemcpyDeviceToHost, s));
      PE_HIP(hipMemcpyAsync(h_id.data(), run_id.p, nruns * 8, hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
    }
  }

  // host: chunk-major run records over updates, scatter units, bin ranges, gather units
  const uint64_t nruns = h_pos.size(), nupd = L->nupdates;
  const uint64_t cmask = (1ull << cbits) - 1ull;
  std::vector<uint64_t> rec_per_chunk(L->nchunks + 1, 0);
  for (uint64_t r = 0; r < nruns; ++r) {
    const uint64_t len = (r + 1 < nruns ? h_upd[r + 1] : nupd) - h_upd[r];
    rec_per_chunk[(h_id[r] & cmask) + 1] += (len + o.run_cap - 1) / o.run_cap;
  }
  for (uint32_t c = 0; c < L->nchunks; ++c) rec_per_chunk[c + 1] += rec_per_chunk[c];
  const uint64_t nrec = rec_per_chunk[L->nchunks];
  L->nrecords = nrec;
  std::vector<uint64_t> rs(nrec);
  std::vector<uint32_t> rl(nrec);
  {
    std::vector<uint64_t> fill(rec_per_chunk.begin(), rec_per_chunk.end() - 1);
    for (uint64_t r = 0; r < nruns; ++r) {  // bin-major, so records stay in bin order inside a chunk
      const uint64_t end = r + 1 < nruns ? h_upd[r + 1] : nupd;
      uint64_t& f = fill[h_id[r] & cmask];
      for (uint64_t p = h_upd[r]; p < end; p += o.run_cap) {
        rs[f] = p;
        rl[f] = (uint32_t)std::min<uint64_t>(o.run_cap, end - p);
        ++f;
      }
    }
  }
  // Scatter units: a chunk's records, cut at about one CU's share of all updates.  The
  // largest go first from the queue, so the hub chunks' units start together, one per CU
  // or so; smaller caps re-stage hub chunks more often, larger ones leave CUs idle
  // (RMAT-27: 1.52 ms at updates / CUs against 1.74 ms at 2^20 and 2.27 ms at 2^23).
  std::vector<pb_unit1> u1;
  std::vector<uint64_t> u1_size;
  const uint64_t ncu = (uint64_t)std::max(1, ctx->num_cus);
  const uint64_t cap1 = std::max<uint64_t>(1ull << 16, (nupd + ncu - 1) / ncu);
  for (uint32_t c = 0; c < L->nchunks; ++c) {
    uint64_t r = rec_per_chunk[c], acc = 0, r0 = r;
    for (; r < rec_per_chunk[c + 1]; ++r) {
      if (acc > 0 && acc + rl[r] > cap1) {
        u1.push_back(pb_unit1{r0, r, c, 0u});
        u1_size.push_back(acc);
        r0 = r;
        acc = 0;
      }
      acc += rl[r];
    }
    if (acc > 0) {
      u1.push_back(pb_unit1{r0, r, c, 0u});
      u1_size.push_back(acc);
    }
  }
  std::vector<uint64_t> bin_ptr(L->nbins + 1, nnz);
  {
    uint32_t b = 0;
    for (uint64_t r = 0; r < nruns; ++r) {
      const uint32_t rb = (uint32_t)(h_id[r] >> cbits);
      while (b <= rb) bin_ptr[b++] = h_pos[r];
    }
    while (b <= L->nbins) bin_ptr[b++] = nnz;
  }
  const uint64_t unit2 = 1ull << o.unit_log2;
  std::vector<pb_unit2> u2;
  std::vector<uint32_t> split;
  for (uint32_t b = 0; b < L->nbins; ++b) {
    const uint64_t e0 = bin_ptr[b], e1 = bin_ptr[b + 1];
    if (e1 - e0 <= unit2) {
      u2.push_back(pb_unit2{e0, e1, b, 0xFFFFFFFFu});
    } else {
      const uint32_t slot = (uint32_t)split.size();
      split.push_back(b);
      for (uint64_t e = e0; e < e1; e += unit2) u2.push_back(pb_unit2{e, std::min(e1, e + unit2), b, slot});
    }
  }
  // largest units first: persistent blocks take units round-robin
  {
    std::vector<uint32_t> idx(u1.size());
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return u1_size[a] > u1_size[b]; });
    std::vector<pb_unit1> t(u1.size());
    for (size_t i = 0; i < idx.size(); ++i) t[i] = u1[idx[i]];
    u1.swap(t);
    std::stable_sort(u2.begin(), u2.end(),
                     [](const pb_unit2& a, const pb_unit2& b) { return (a.e1 - a.e0) > (b.e1 - b.e0); });
  }
  // (the unit queues hand out one ticket per block past the last unit: counts stay far below 2^32)
  PE_REQUIRE(u1.size() < (1ull << 31) && u2.size() < (1ull << 31), PE_ERR_RANGE, "too many binned units");
  L->n1 = (uint32_t)u1.size();
  L->n2 = (uint32_t)u2.size();
  L->nsplit = (uint32_t)split.size();
  upload(ctx, L->rec_start, rs);
  upload(ctx, L->rec_len, rl);
  upload(ctx, L->units1, u1);
  upload(ctx, L->units2, u2);
  upload(ctx, L->split_bin, split);
  L->values.alloc(ctx, std::max<uint64_t>(1, nupd));
  const uint64_t nacc = std::max<uint64_t>(1, (uint64_t)L->nsplit << db);
  L->acc.alloc(ctx, nacc);
  k_pb_zero<<<grid_for(nacc, 256, 65536), 256, 0, s>>>(L->acc.p, nacc);
  PE_HIP(hipGetLastError());
  L->work.alloc(ctx, 4);
  k_pb_zero<<<1, 256, 0, s>>>(reinterpret_cast<unsigned long long*>(L->work.p), 2);
  PE_HIP(hipGetLastError());
  PE_HIP(hipStreamSynchronize(s));
  g->pb = std::move(L);
}

// Blocks of `threads` that fit one CU with `lds` bytes each (160 KiB of LDS, 2048 threads).
static unsigned pb_blocks_per_cu(size_t lds, int threads) {
  const unsigned by_lds = (unsigned)std::max<size_t>(1, (160u << 10) / std::max<size_t>(1, lds));
  const unsigned by_threads = (unsigned)std::max(1, 2048 / threads);
  return std::max(1u, std::min(by_lds, by_threads));
}

static size_t pb_gather_lds_bytes(uint32_t dst_bits) {
  return ((size_t)8 << dst_bits) + (size_t)16 * (PB_GATHER_THREADS / 64) + 16;
}

unsigned binned_grid_gather(pe_graph_s* g) {
  return persistent_grid(g->ctx, std::max<uint32_t>(1, g->pb->n2),
                         pb_blocks_per_cu(pb_gather_lds_bytes(g->pb->dst_bits), PB_GATHER_THREADS));
}

unsigned binned_grid_split(pe_graph_s* g) {
  return grid_for((uint64_t)g->pb->nsplit << g->pb->dst_bits, 256, (unsigned)std::max(1, g->ctx->num_cus) * 4);
}

void binned_iteration(pe_graph_s* g, float* b0, float* b1, bool weighted, const rank_args& a, iter_ctl* ctl,
                      double* part_delta, double* part_rsum) {
  PE_REQUIRE(g->pb, PE_ERR_STATE, "binned layout not built");
  const pb_layout& L = *g->pb;
  PE_REQUIRE(!weighted || L.weighted, PE_ERR_STATE, "weighted iteration on a layout built without weights");
  PE_REQUIRE(a.mass > 0.0, PE_ERR_INTERNAL, "binned iteration without a bound on the contributions");
  const int frac = pb_frac_bits(a.mass);
  hipStream_t s = g->ctx->stream;
  const size_t lds1 = ((size_t)4 << L.src_bits) + 16;
  const unsigned g1 = persistent_grid(g->ctx, std::max<uint32_t>(1, L.n1), pb_blocks_per_cu(lds1, PB_SCATTER_THREADS));
  const size_t lds2 = pb_gather_lds_bytes(L.dst_bits);
  const unsigned g2 = binned_grid_gather(g), gs = binned_grid_split(g);
  {
    ktime_scope kt(g->ctx, PE_KT_PULL_BIN_SCATTER);
    uint32_t* vals = reinterpret_cast<uint32_t*>(L.values.p);
    if (weighted)
      k_pb_scatter<PB_SCATTER_RECORDS, PB_SCATTER_THREADS, false><<<g1, PB_SCATTER_THREADS, lds1, s>>>(
          L.units1.p, L.n1, L.rec_start.p, L.rec_len.p, L.src_upd.p, g->V, (int)L.src_bits, frac, b0, b1, ctl, vals,
          L.work.p);
    else
      k_pb_scatter<PB_SCATTER_RECORDS, PB_SCATTER_THREADS, true><<<g1, PB_SCATTER_THREADS, lds1, s>>>(
          L.units1.p, L.n1, L.rec_start.p, L.rec_len.p, L.src_upd.p, g->V, (int)L.src_bits, frac, b0, b1, ctl, vals,
          L.work.p);
    PE_HIP(hipGetLastError());
  }
  {
    ktime_scope kt(g->ctx, PE_KT_PULL_BIN_GATHER);
    if (weighted)
      k_pb_gather<true, PB_GATHER_STEPS_W, PB_GATHER_THREADS><<<g2, PB_GATHER_THREADS, lds2, s>>>(L.units2.p, L.n2, reinterpret_cast<const uint32_t*>(L.values.p), L.w.p, L.dst_local.p, L.gbase.p, g->V,
                                                  (int)L.dst_bits, frac, b0, b1, g->od.p, a, ctl, L.acc.p, part_delta,
                                                  part_rsum, L.work.p + 2);
    else
      k_pb_gather<false, PB_GATHER_STEPS, PB_GATHER_THREADS><<<g2, PB_GATHER_THREADS, lds2, s>>>(L.units2.p, L.n2, reinterpret_cast<const uint32_t*>(L.values.p), nullptr, L.dst_local.p, L.gbase.p, g->V,
                                                   (int)L.dst_bits, frac, b0, b1, g->od.p, a, ctl, L.acc.p, part_delta,
                                                   part_rsum, L.work.p + 2);
    PE_HIP(hipGetLastError());
    k_pb_apply_split<<<gs, 256, 0, s>>>(L.split_bin.p, L.nsplit, (int)L.dst_bits, frac, g->V, L.acc.p, b0, b1, g->od.p,
                                        a, ctl, part_delta + g2, part_rsum + g2);
    PE_HIP(hipGetLastError());
  }
}

}  // namespace pe

extern "C" int pe_graph_set_pull_layout(pe_graph_t* g, int layout, const pe_binned_opts_t* opts) {
  if (!g) return PE_ERR_INVALID;
  return pe::guarded(g->ctx, [&] {
    if (layout == PE_PULL_MERGE_PATH) {
      g->pull_layout = PE_PULL_MERGE_PATH;
      g->pb.reset();
      return;
    }
    PE_REQUIRE(layout == PE_PULL_BINNED, PE_ERR_INVALID, "layout must be PE_PULL_MERGE_PATH or PE_PULL_BINNED");
    pe::binned_build(g, opts);
    g->pull_layout = PE_PULL_BINNED;
  });
}
