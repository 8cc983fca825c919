// HyperANF: plato::algo::hyperanf (plato/algo/hyperanf/hyperanf.hpp:80-197) on the device.
//
// Every vertex holds a HyperLogLog counter of 2^P one-byte registers
// (plato/util/hyperloglog.hpp:32-142), started with its own id (MurmurHash3_x86_32 of
// the 4-byte id, seed 0x5f61767a: register = hash >> (32 - P), rank = 1 + leading
// zeros of hash << P, at most 33 - P).  An iteration gives every vertex the
// register-wise max of its counter and its in-neighbours' counters of the previous
// iteration: the reference's push (active vertices merge into their out-neighbours'
// deltas) and pull (every vertex merges its in-neighbours) deliver exactly that, since
// a counter that did not change was already merged downstream.  Register max is
// order-independent, so the counters are bit-exact; anf is the FP64 sum of the
// per-vertex estimates, and the loop stops after `iterations` or when anf grows by less
// than DBL_EPSILON (that anf is dropped).  The average distance is GetAvgDistance over
// the anfs (hyperanf.hpp:58-67).
//
// Layout: counters [V][2^P] bytes, double-buffered (Jacobi).  Pull: one wave per
// vertex, registers handled as 32-bit words of four; 64- and 128-register counters take
// 4 or 2 in-edges per wave step (a row is 16 or 32 lanes), larger ones one edge per step
// with each lane owning 2^P / 256 words; several rows are loaded before they are merged.
// Vertices with more than HANF_HUB_DEG in-edges (the hubs) only copy their own row in
// that pass; a second pass cuts their in-edges into chunks spread over all waves, each
// merging its chunk's maximum into the hub's row, so no wave walks a hub alone.
#include <cfloat>
#include <cmath>
#include <vector>

#include "pe_internal.hpp"

namespace pe {

static constexpr uint64_t HANF_HUB_DEG = 4096;    // in-edges above which a vertex is a hub
static constexpr uint64_t HANF_HUB_CHUNK = 2048;  // in-edges of a hub per wave task
static constexpr int HANF_HUB_U = 8;             // rows each lane of the hub pass keeps in flight
static constexpr uint64_t HANF_LOW_DEG = 32;    // counters of up to 128 registers: fewer in-edges
                                                 // take a lane group per vertex

__device__ __forceinline__ uint32_t hanf_rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// MurmurHash3_x86_32 of one little-endian 32-bit word (len 4), the reference's hash of
// a vertex id (plato/util/hash.hpp:33-96, A. Appleby's published algorithm)
__device__ __forceinline__ uint32_t hanf_murmur_u32(uint32_t key, uint32_t seed) {
  uint32_t k = key * 0xcc9e2d51u;
  k = hanf_rotl(k, 15);
  k *= 0x1b873593u;
  uint32_t h = seed ^ k;
  h = hanf_rotl(h, 13);
  h = h * 5 + 0xe6546b64u;
  h ^= 4u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__global__ void k_hanf_init(uint8_t* __restrict__ regs, uint32_t V, int P, const uint32_t* __restrict__ perm) {
  const uint64_t m = 1ull << P;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t id = perm ? perm[v] : (uint32_t)v;  // counters hash input ids
    const uint32_t h = hanf_murmur_u32(id, 0x5f61767au);
    const uint32_t w = h << P;
    const uint32_t q = 32u - (uint32_t)P;
    const uint32_t rank = w ? (uint32_t)__clz(w) + 1u : q + 1u;
    regs[v * m + (h >> q)] = (uint8_t)(rank < q + 1u ? rank : q + 1u);
  }
}

__global__ void k_hanf_zero(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

__device__ __forceinline__ uint32_t byte_max4(uint32_t a, uint32_t b) {
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < 32; i += 8) r |= max((a >> i) & 0xFFu, (b >> i) & 0xFFu) << i;
  return r;
}

// Pull for counters of 64 or 128 registers (16 or 32 words): a wave takes EPS in-edges
// per step (64 / EPS lanes per row, one word each), U steps in flight, and folds the
// EPS partial rows together with shuffles at the end.  0 is the identity of max.
template <int EPS, int U>
__global__ __launch_bounds__(256) void k_hanf_pull_small(const uint64_t* __restrict__ col_ptr,
                                                         const uint32_t* __restrict__ row_idx, uint32_t V,
                                                         uint64_t min_deg, const uint32_t* __restrict__ cur,
                                                         uint32_t* __restrict__ nxt) {
  constexpr int WPR = 64 / EPS;  // words per counter
  const int lane = threadIdx.x & 63, slot = lane / WPR, w = lane % WPR;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  for (uint64_t v = wave_id; v < V; v += nwaves) {
    const uint64_t b = col_ptr[v], e = col_ptr[v + 1];
    if (e - b < min_deg) continue;  // k_hanf_pull_lowdeg
    uint32_t acc = slot == 0 ? cur[v * WPR + w] : 0u;
    if (e - b > HANF_HUB_DEG) {  // a hub: its own row now, its in-edges in k_hanf_pull_hubs
      if (slot == 0) nxt[v * WPR + w] = acc;
      continue;
    }
    for (uint64_t k0 = b; k0 < e; k0 += (uint64_t)EPS * U) {
      uint32_t val[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t k = k0 + (uint64_t)u * EPS + slot;
        val[u] = k < e ? cur[(uint64_t)row_idx[k] * WPR + w] : 0u;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) acc = byte_max4(acc, val[u]);
    }
    for (int off = WPR; off < 64; off <<= 1) acc = byte_max4(acc, (uint32_t)__shfl_xor(acc, off, 64));
    if (slot == 0) nxt[v * WPR + w] = acc;
  }
}

// Pull for the vertices with fewer than max_deg in-edges and counters of WPR <= 32
// words: one group of WPR lanes per vertex, 64 / WPR vertices per wave step, U in-edges
// of each in flight, so the vertices' short chains of dependent loads overlap.
template <int WPR, int U>
__global__ __launch_bounds__(256) void k_hanf_pull_lowdeg(const uint64_t* __restrict__ col_ptr,
                                                          const uint32_t* __restrict__ row_idx, uint32_t V,
                                                          uint64_t max_deg, const uint32_t* __restrict__ cur,
                                                          uint32_t* __restrict__ nxt) {
  constexpr int G = 64 / WPR;  // vertices per wave step
  const int lane = threadIdx.x & 63, grp = lane / WPR, w = lane % WPR;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  for (uint64_t v0 = wave_id * G; v0 < V; v0 += nwaves * G) {
    const uint64_t v = v0 + grp;
    uint64_t b = 0, e = 0;
    bool mine = false;
    if (v < V) {
      b = col_ptr[v];
      e = col_ptr[v + 1];
      mine = e - b < max_deg;
    }
    if (!mine) e = b;
    uint32_t acc = mine ? cur[v * WPR + w] : 0u;
    // the wave steps until its longest list is done
    uint64_t len = e - b;
    for (int off = WPR; off < 64; off <<= 1) len = std::max<uint64_t>(len, (uint64_t)__shfl_xor((long long)len, off, 64));
    for (uint64_t k0 = 0; k0 < len; k0 += U) {
      uint32_t val[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t k = b + k0 + u;
        val[u] = k < e ? cur[(uint64_t)row_idx[k] * WPR + w] : 0u;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) acc = byte_max4(acc, val[u]);
    }
    if (mine) nxt[v * WPR + w] = acc;
  }
}

// Pull with W 32-bit words (4 W registers) per lane and chunk, P >= 8: a counter row
// of `row` words is handled in chunks of 64 W words (larger counters walk the in-edges
// once per chunk), U in-edges in flight.
template <int W, int U>
__global__ __launch_bounds__(256) void k_hanf_pull_words(const uint64_t* __restrict__ col_ptr,
                                                         const uint32_t* __restrict__ row_idx, uint32_t V,
                                                         uint64_t row, const uint32_t* __restrict__ cur,
                                                         uint32_t* __restrict__ nxt) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  for (uint64_t v = wave_id; v < V; v += nwaves) {
    const uint64_t b = col_ptr[v], e = col_ptr[v + 1];
    if (e - b > HANF_HUB_DEG) {  // a hub: its own row now, its in-edges in k_hanf_pull_hubs
      for (uint64_t i = lane; i < row; i += 64) nxt[v * row + i] = cur[v * row + i];
      continue;
    }
    for (uint64_t c0 = 0; c0 < row; c0 += 64 * W) {
      uint32_t acc[W];
#pragma unroll
      for (int j = 0; j < W; ++j) acc[j] = cur[v * row + c0 + (uint64_t)j * 64 + lane];
      for (uint64_t k0 = b; k0 < e; k0 += U) {
        uint32_t val[U][W];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const bool ok = k0 + u < e;
          const uint64_t urow = ok ? (uint64_t)row_idx[k0 + u] * row + c0 : 0;
#pragma unroll
          for (int j = 0; j < W; ++j) val[u][j] = ok ? cur[urow + (uint64_t)j * 64 + lane] : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < W; ++j) acc[j] = byte_max4(acc[j], val[u][j]);
      }
#pragma unroll
      for (int j = 0; j < W; ++j) nxt[v * row + c0 + (uint64_t)j * 64 + lane] = acc[j];
    }
  }
}

// Register-wise max of x into *p (a compare-and-swap loop: every retry sees a word that
// only grew, so it ends).
__device__ __forceinline__ void byte_max_merge(uint32_t* p, uint32_t x) {
  uint32_t old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    const uint32_t nv = byte_max4(old, x);
    if (nv == old) return;
    const uint32_t prev = atomicCAS(p, old, nv);
    if (prev == old) return;
    old = prev;
  }
}

// Hub in-edges, one chunk per wave at a time: the chunk's register-wise max (rows of
// wpr words; up to 64 words: 64 / wpr in-edges per step, larger rows 64 words per step
// with the in-edges walked once per 64-word block) merged into the hub's next row.
__global__ __launch_bounds__(256) void k_hanf_pull_hubs(const uint64_t* __restrict__ col_ptr,
                                                        const uint32_t* __restrict__ row_idx,
                                                        const uint32_t* __restrict__ hubs,
                                                        const unsigned long long* __restrict__ nhub,
                                                        const uint64_t* __restrict__ start, uint64_t wpr,
                                                        const uint32_t* __restrict__ cur, uint32_t* __restrict__ nxt) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  const uint64_t n = *nhub;
  if (n == 0) return;
  const uint64_t total = start[n];
  for (uint64_t c = wave_id; c < total; c += nwaves) {
    uint64_t k0, k1;
    const uint32_t v = hub_chunk(col_ptr, hubs, start, n, HANF_HUB_CHUNK, c, &k0, &k1);
    if (wpr <= 64) {
      const uint64_t eps = 64 / wpr, slot = (uint64_t)lane / wpr, w = (uint64_t)lane % wpr;
      uint32_t acc = 0;
      for (uint64_t k = k0 + slot; k < k1; k += eps * HANF_HUB_U) {
        uint32_t val[HANF_HUB_U];
#pragma unroll
        for (int u = 0; u < HANF_HUB_U; ++u) {
          const uint64_t kk = k + (uint64_t)u * eps;
          val[u] = kk < k1 ? cur[(uint64_t)row_idx[kk] * wpr + w] : 0u;
        }
#pragma unroll
        for (int u = 0; u < HANF_HUB_U; ++u) acc = byte_max4(acc, val[u]);
      }
      for (uint64_t off = wpr; off < 64; off <<= 1) acc = byte_max4(acc, (uint32_t)__shfl_xor(acc, (int)off, 64));
      if (slot == 0) byte_max_merge(&nxt[(uint64_t)v * wpr + w], acc);
    } else {
      for (uint64_t cw = lane; cw < wpr; cw += 64) {
        uint32_t acc = 0;
        for (uint64_t k = k0; k < k1; ++k) acc = byte_max4(acc, cur[(uint64_t)row_idx[k] * wpr + cw]);
        byte_max_merge(&nxt[(uint64_t)v * wpr + cw], acc);
      }
    }
  }
}

// 2^-r exactly, from the exponent bits (r <= 33)
__device__ __forceinline__ double hanf_pow2_neg(uint32_t r) {
  return __longlong_as_double((long long)(1023u - r) << 52);
}

// The estimate of one counter from its register sum s (exact: at most 2^16 terms of 2^-r)
// and its zero registers (hyperloglog.hpp:77-116, with its small- and large-range
// corrections); alpha is alpha_m * m^2.
__device__ __forceinline__ double hanf_estimate(double alpha, double m, double s, uint32_t zeros) {
  double est = alpha / s;
  if (est <= 5.0 / 2.0 * m) {
    if (zeros) est = m * log(m / zeros);
  } else if (est > (1.0 / 30.0) * 4294967296.0) {
    est = -4294967296.0 * log(1.0 - (est / 4294967296.0));
  }
  return est;
}

__device__ __forceinline__ double hanf_alpha(int P) {
  const double m = (double)(1u << P);
  double alpha;
  switch (P) {
    case 4: alpha = 0.673; break;
    case 5: alpha = 0.697; break;
    case 6: alpha = 0.709; break;
    default: alpha = 0.7213 / (1.0 + 1.079 / m); break;
  }
  return alpha * m * m;
}

// Estimates of counters of at most 1024 registers, summed per block in a fixed order:
// m / 16 lanes per counter (16 registers each, one 16-byte load), 1024 / m counters per
// wave step, so a wave keeps many counters' loads in flight instead of one.
__global__ __launch_bounds__(256) void k_hanf_estimate_rows(const uint8_t* __restrict__ regs, uint32_t V, int P,
                                                            double* __restrict__ part) {
  __shared__ double s_w[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t m = 1u << P, L = m / 16, sub = (uint32_t)lane % L, slot = (uint32_t)lane / L, per = 64 / L;
  const double alpha = hanf_alpha(P);
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  double total = 0.0;  // lane's running sum over its counters, in order
  for (uint64_t v0 = wave_id * per; v0 < V; v0 += nwaves * per) {
    const uint64_t v = v0 + slot;
    double s = 0.0;
    uint32_t zeros = 0;
    if (v < V) {
      const uint4 q = *reinterpret_cast<const uint4*>(regs + v * m + (uint64_t)sub * 16);
      const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int b = 0; b < 32; b += 8) {
          const uint32_t r = (w4[i] >> b) & 0xFFu;
          s += hanf_pow2_neg(r);
          zeros += r == 0 ? 1u : 0u;
        }
    }
    for (uint32_t off = 1; off < L; off <<= 1) {
      s += __shfl_xor(s, (int)off, 64);
      zeros += (uint32_t)__shfl_xor((int)zeros, (int)off, 64);
    }
    if (v < V && sub == 0) total += hanf_estimate(alpha, (double)m, s, zeros);
  }
  for (int off = 32; off > 0; off >>= 1) total += __shfl_down(total, off, 64);
  if (lane == 0) s_w[wave] = total;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Per-vertex estimate (hyperloglog.hpp:77-116) of larger counters, one wave per counter,
// summed per block in a fixed order.
// The register sum of 2^-reg is exact in FP64 (at most 2^16 terms of 2^-k, k <= 27).
__global__ __launch_bounds__(256) void k_hanf_estimate(const uint8_t* __restrict__ regs, uint32_t V, int P,
                                                       double* __restrict__ part) {
  __shared__ double s_w[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t m = 1ull << P;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  const double alpha = hanf_alpha(P);
  double total = 0.0;  // lane 0's running sum over this wave's vertices, in order
  for (uint64_t v = wave_id; v < V; v += nwaves) {
    double s = 0.0;
    uint32_t zeros = 0;
    for (uint64_t i = lane; i < m; i += 64) {
      const uint32_t r = regs[v * m + i];
      s += hanf_pow2_neg(r);
      zeros += r == 0 ? 1u : 0u;
    }
    for (int off = 32; off > 0; off >>= 1) {
      s += __shfl_xor(s, off, 64);
      zeros += __shfl_xor(zeros, off, 64);
    }
    total += hanf_estimate(alpha, (double)m, s, zeros);
  }
  if (lane == 0) s_w[wave] = total;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// n-step degrees: the estimate of each target's counter as the reference stores it,
// static_cast<vid_t>(estimate) - 1 (nstepdegrees.hpp:412-421); 0 for other vertices.
// One wave per vertex; targets (input ids, a bitmap) are looked up through perm.
__global__ __launch_bounds__(256) void k_nstep_degree(const uint8_t* __restrict__ regs, uint32_t V, int P,
                                                      const uint64_t* __restrict__ targets,
                                                      const uint32_t* __restrict__ perm, uint32_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t m = 1ull << P;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  const double alpha = hanf_alpha(P);
  for (uint64_t v = wave_id; v < V; v += nwaves) {
    const uint32_t id = perm ? perm[v] : (uint32_t)v;
    if (targets && !((targets[id >> 6] >> (id & 63)) & 1ull)) {
      if (lane == 0) out[v] = 0u;
      continue;
    }
    double s = 0.0;
    uint32_t zeros = 0;
    for (uint64_t i = lane; i < m; i += 64) {
      const uint32_t r = regs[v * m + i];
      s += hanf_pow2_neg(r);
      zeros += r == 0 ? 1u : 0u;
    }
    for (int off = 32; off > 0; off >>= 1) {
      s += __shfl_xor(s, off, 64);
      zeros += __shfl_xor(zeros, off, 64);
    }
    if (lane == 0) out[v] = (uint32_t)hanf_estimate(alpha, (double)m, s, zeros) - 1u;
  }
}

// Any word of a differs from b: *flag = 1.
__global__ void k_rows_differ(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t n,
                              unsigned int* __restrict__ flag) {
  bool d = false;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    d |= a[i] != b[i];
  if (__ballot(d) && (threadIdx.x & 63) == 0) *flag = 1u;
}

__global__ void k_flag_zero(unsigned int* f) { *f = 0u; }

// counters in input-id order
__global__ void k_hanf_rows_external(const uint8_t* __restrict__ in, const uint32_t* __restrict__ rank_of,
                                     uint32_t V, uint64_t m, uint8_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (uint64_t)V * m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = i / m, j = i % m;
    out[i] = in[(uint64_t)rank_of[v] * m + j];
  }
}

}  // namespace pe

using namespace pe;

// One HyperANF pull over the adjacency (ptr, idx): nxt[v] = register-wise max of cur[v]
// and cur[u] over the adjacency of v.  hubs: the adjacency's hub table.
static void hanf_pull(pe_ctx_s* ctx, const uint64_t* ptr, const uint32_t* idx, uint32_t V, int P, hub_table& hubs,
                      const uint8_t* cur, uint8_t* nxt, unsigned grid) {
  hipStream_t s = ctx->stream;
  const uint64_t m = 1ull << P;
  switch (P) {
    case 6:
      k_hanf_pull_lowdeg<16, 4><<<grid, 256, 0, s>>>(ptr, idx, V, HANF_LOW_DEG, (const uint32_t*)cur, (uint32_t*)nxt);
      k_hanf_pull_small<4, 8><<<grid, 256, 0, s>>>(ptr, idx, V, HANF_LOW_DEG, (const uint32_t*)cur, (uint32_t*)nxt);
      break;
    case 7:
      k_hanf_pull_lowdeg<32, 4><<<grid, 256, 0, s>>>(ptr, idx, V, HANF_LOW_DEG, (const uint32_t*)cur, (uint32_t*)nxt);
      k_hanf_pull_small<2, 4><<<grid, 256, 0, s>>>(ptr, idx, V, HANF_LOW_DEG, (const uint32_t*)cur, (uint32_t*)nxt);
      break;
#define HANF_W(PP, W, U)                                                                                 \
  case PP:                                                                                               \
    k_hanf_pull_words<W, U><<<grid, 256, 0, s>>>(ptr, idx, V, m / 4, (const uint32_t*)cur, (uint32_t*)nxt); \
    break;
    HANF_W(8, 1, 8) HANF_W(9, 2, 4) HANF_W(10, 4, 4) HANF_W(11, 8, 2) HANF_W(12, 16, 2) HANF_W(13, 16, 2)
    HANF_W(14, 16, 2) HANF_W(15, 16, 2) HANF_W(16, 16, 2)
#undef HANF_W
  }
  k_hanf_pull_hubs<<<grid, 256, 0, s>>>(ptr, idx, hubs.hubs.p, hubs.count.p, hubs.start.p, m / 4,
                                        (const uint32_t*)cur, (uint32_t*)nxt);
  PE_HIP(hipGetLastError());
}

extern "C" int pe_hyperanf(pe_graph_t* g, const pe_hyperanf_opts_t* opts, double* anfs, uint32_t cap,
                           uint32_t* nanfs, double* avg_distance, uint8_t* registers, double* seconds) {
  if (!g || !opts || !anfs || cap == 0) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    const int P = (int)opts->bits;
    PE_REQUIRE(P >= 6 && P <= 16, PE_ERR_INVALID, "bits must be in [6, 16] (hyperanf_test supported range)");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t V = g->V;
    const uint64_t m = 1ull << P, bytes = std::max<uint64_t>(4, (uint64_t)V * m);
    dbuf<uint8_t> a(ctx, bytes), b(ctx, bytes);
    const unsigned est_grid = persistent_grid(ctx, grid_for(V, 4), 4);
    dbuf<double> part(ctx, est_grid);
    std::vector<double> hpart(est_grid);
    const unsigned pull_grid = persistent_grid(ctx, grid_for(V, 4), 8);
    const uint32_t* perm = g->relabeled ? g->perm.p : nullptr;
    hub_table hubs;
    auto anf_of = [&](const uint8_t* regs) {
      if (P <= 10)
        k_hanf_estimate_rows<<<est_grid, 256, 0, s>>>(regs, V, P, part.p);
      else
        k_hanf_estimate<<<est_grid, 256, 0, s>>>(regs, V, P, part.p);
      PE_HIP(hipGetLastError());
      PE_HIP(hipMemcpyAsync(hpart.data(), part.p, est_grid * sizeof(double), hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      double t = 0.0;
      for (double x : hpart) t += x;
      return t;
    };
    auto pull = [&](const uint8_t* cur, uint8_t* nxt) {
      hanf_pull(ctx, g->col_ptr.p, g->row_idx.p, V, P, hubs, cur, nxt, pull_grid);
    };
    PE_HIP(hipEventRecord(ctx->t0, s));
    hub_table_build(ctx, g->col_ptr.p, V, g->nnz, HANF_HUB_DEG, HANF_HUB_CHUNK, hubs);
    k_hanf_zero<<<grid_for(bytes / 4, 256, 65536), 256, 0, s>>>((uint32_t*)a.p, bytes / 4);
    k_hanf_init<<<grid_for(V, 256, 8192), 256, 0, s>>>(a.p, V, P, perm);
    PE_HIP(hipGetLastError());
    std::vector<double> out;
    double last = anf_of(a.p);
    out.push_back(last);
    uint8_t* cur = a.p;
    uint8_t* nxt = b.p;
    for (uint32_t it = 0; it < opts->iterations; ++it) {
      pull(cur, nxt);
      const double anf = anf_of(nxt);
      std::swap(cur, nxt);
      if (anf - last < DBL_EPSILON) break;  // hyperanf.hpp:189-191
      out.push_back(anf);
      last = anf;
    }
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    for (size_t i = 0; i < out.size() && i < cap; ++i) anfs[i] = out[i];
    if (nanfs) *nanfs = (uint32_t)out.size();
    if (avg_distance) {  // GetAvgDistance (hyperanf.hpp:58-67)
      double avg = 0.0;
      for (size_t i = 1; i < out.size(); ++i) avg += i * (out[i] - out[i - 1]);
      *avg_distance = avg / (out.back() - out.front());
    }
    if (registers) {
      if (g->relabeled) {
        k_hanf_rows_external<<<grid_for((uint64_t)V * m, 256, 65536), 256, 0, s>>>(cur, g->rank_of.p, V, m, nxt);
        PE_HIP(hipGetLastError());
        PE_HIP(hipMemcpyAsync(registers, nxt, (uint64_t)V * m, hipMemcpyDeviceToHost, s));
      } else {
        PE_HIP(hipMemcpyAsync(registers, cur, (uint64_t)V * m, hipMemcpyDeviceToHost, s));
      }
      PE_HIP(hipStreamSynchronize(s));
    }
    if (seconds) *seconds = ms * 1e-3;
  });
}

extern "C" int pe_nstep_degrees(pe_graph_t* g, const uint64_t* target_words, uint32_t step, uint32_t bits,
                                int type, uint32_t* in_degrees, uint32_t* out_degrees, double* seconds) {
  if (!g) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    const int P = (int)bits;
    PE_REQUIRE(P >= 6 && P <= 16, PE_ERR_INVALID, "bits must be in [6, 16] (nstepdegrees supported range)");
    PE_REQUIRE(type >= PE_NSTEP_IN && type <= PE_NSTEP_BOTH, PE_ERR_INVALID, "type must be PE_NSTEP_IN, OUT or BOTH");
    PE_REQUIRE(!(type & PE_NSTEP_IN) || in_degrees, PE_ERR_INVALID, "in_degrees is required for the in pass");
    PE_REQUIRE(!(type & PE_NSTEP_OUT) || out_degrees, PE_ERR_INVALID, "out_degrees is required for the out pass");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t V = g->V;
    const uint64_t m = 1ull << P, bytes = std::max<uint64_t>(4, (uint64_t)V * m);
    const uint32_t* perm = g->relabeled ? g->perm.p : nullptr;
    dbuf<uint64_t> tw;
    if (target_words) {
      const uint64_t nw = ((uint64_t)V >> 6) + 1;
      tw.alloc(ctx, nw);
      PE_HIP(hipMemcpyAsync(tw.p, target_words, nw * 8, hipMemcpyHostToDevice, s));
    }
    dbuf<uint8_t> a(ctx, bytes), b(ctx, bytes);
    dbuf<uint32_t> dev_deg(ctx, std::max<uint32_t>(1, V));
    dbuf<unsigned int> flag(ctx, 1);
    const unsigned grid = persistent_grid(ctx, grid_for(V, 4), 8);
    PE_HIP(hipEventRecord(ctx->t0, s));
    // one pass per direction: the in pass gathers over in-edges (the reference's
    // propagate on the forward engine), the out pass over out-edges (on the reversed one)
    auto pass = [&](const uint64_t* ptr, const uint32_t* idx, uint32_t* host_out) {
      hub_table hubs;
      hub_table_build(ctx, ptr, V, g->nnz, HANF_HUB_DEG, HANF_HUB_CHUNK, hubs);
      k_hanf_zero<<<grid_for(bytes / 4, 256, 65536), 256, 0, s>>>((uint32_t*)a.p, bytes / 4);
      k_hanf_init<<<grid_for(V, 256, 8192), 256, 0, s>>>(a.p, V, P, perm);
      PE_HIP(hipGetLastError());
      uint8_t* cur = a.p;
      uint8_t* nxt = b.p;
      for (uint32_t it = 0; it < step; ++it) {  // nstepdegrees.hpp:343-409
        hanf_pull(ctx, ptr, idx, V, P, hubs, cur, nxt, grid);
        k_flag_zero<<<1, 1, 0, s>>>(flag.p);
        k_rows_differ<<<grid_for(bytes / 4, 256, 8192), 256, 0, s>>>((const uint32_t*)cur, (const uint32_t*)nxt,
                                                                     bytes / 4, flag.p);
        PE_HIP(hipGetLastError());
        unsigned int changed = 0;
        PE_HIP(hipMemcpyAsync(&changed, flag.p, 4, hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        std::swap(cur, nxt);
        if (!changed) break;  // no book grew: further steps change nothing
      }
      k_nstep_degree<<<grid, 256, 0, s>>>(cur, V, P, tw.p, perm, dev_deg.p);
      PE_HIP(hipGetLastError());
      download_external(g, dev_deg.p, host_out, 4);
    };
    if (type & PE_NSTEP_IN) pass(g->col_ptr.p, g->row_idx.p, in_degrees);
    if (type & PE_NSTEP_OUT) {
      if (!g->directed && (type & PE_NSTEP_IN)) {
        std::memcpy(out_degrees, in_degrees, (size_t)V * 4);  // undirected: out = in (nstepdegrees.hpp:456-463)
      } else {
        PE_REQUIRE(g->has_csr || g->csr_is_csc, PE_ERR_STATE, "the out pass needs out-edges (PE_GRAPH_BUILD_CSR)");
        pass(g->csr_ptr(), g->csr_idx(), out_degrees);
      }
    }
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    if (seconds) *seconds = ms * 1e-3;
  });
}
