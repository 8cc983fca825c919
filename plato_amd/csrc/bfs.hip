// Direction-switching BFS on the dual-mode engine.
//
// Replaces plato::algo::breadth_first_search (plato/algo/bfs/bfs.hpp:55-155)
// with the engine's push/pull choice (plato/engine/dualmode.hpp:261-279):
// push iff (double)A / (double)E' < threshold, A = sum of od over the frontier.
//
//  push (top-down): an unvisited out-neighbour of a frontier vertex is claimed with
//       an integer atomicMin on its level (the claim that the reference does with
//       cas on parent, bfs.hpp:104-116).  Frontiers with few active edges are walked
//       one wave per vertex and append their claims to the next queue (one
//       wave-aggregated atomic per step); larger ones are cut into equal ranges of
//       edges, one per wave, and set their claims in the next bitmap only.
//  pull (bottom-up): one lane per vertex, 64 consecutive vertices per wave, so
//       a wave owns one 64-bit word of the next-frontier bitmap and writes it
//       with a ballot, no atomics; an unvisited vertex scans its in-edges for
//       the first neighbour in the current frontier bitmap and stops there
//       (bfs.hpp:119-128).  Vertices with more than CAP in-edges that found
//       nothing in their first CAP continue in a second kernel, one wave each.
// Every step leaves the next frontier as a bitmap and counts its size and active
// edges on the device, so the next decision needs no scan; a queue is built from the
// bitmap by compaction only when a top-down step needs one and none was appended.
#include <algorithm>
#include <memory>

#include "pe_internal.hpp"


namespace pe {

static constexpr uint32_t UNVIS = 0xFFFFFFFFu;
#ifndef PULL_CAP
#define PULL_CAP 32  // in-edges a pull lane scans before its vertex goes to the wave-per-vertex pass
#endif

struct bfs_counters {
  unsigned long long next_count;
  unsigned long long next_active;
  unsigned long long heavy_count;
  unsigned long long examined;  // pull: in-edges a sequential scan reads, each list up to its first hit
};

__device__ __forceinline__ unsigned wave_prefix(uint64_t mask) {
  // lanes below this one that are set in mask
  return __popcll(mask & ((1ull << (threadIdx.x & 63)) - 1ull));
}

// A search from root: levels UNVIS but the root's 0, the visited bitmap empty, the
// frontier {root} as a bitmap of fwords words and as a queue, its size and active edges
// in *cnt.
__global__ void k_bfs_init(uint32_t* __restrict__ level, uint64_t* __restrict__ vis, uint32_t V, uint64_t nwords,
                           uint32_t root, uint64_t* __restrict__ fbits, uint64_t fwords, uint32_t* __restrict__ queue,
                           bfs_counters* __restrict__ cnt, const uint32_t* __restrict__ od) {
  const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = tid; v < V; v += stride) level[v] = (v == root) ? 0u : UNVIS;
  const uint64_t nw = nwords > fwords ? nwords : fwords;
  for (uint64_t i = tid; i < nw; i += stride) {
    if (i < nwords) vis[i] = 0ull;
    if (i < fwords) fbits[i] = (i == (root >> 6)) ? 1ull << (root & 63) : 0ull;
  }
  if (tid == 0) {
    queue[0] = root;
    cnt->next_count = 1;
    cnt->next_active = od[root];
    cnt->heavy_count = 0;
    cnt->examined = 0;
  }
}

// Before a top-down step: the current frontier joins the visited bitmap, the next is cleared.
__global__ void k_bfs_push_prep(uint64_t* __restrict__ vis, const uint64_t* __restrict__ cur, uint64_t nwords,
                                uint64_t* __restrict__ next, uint64_t next_words) {
  const uint64_t n = nwords > next_words ? nwords : next_words;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < nwords) {
      const uint64_t c = cur[i];
      if (c) vis[i] |= c;
    }
    if (i < next_words) next[i] = 0ull;
  }
}

// Claims w at level L unless it was visited before this step.  The visited bitmap (one
// bit per vertex, read-only during the step, small enough to stay in L2) rules out
// nearly every edge of a late top-down step without touching the level array; a
// target not in it is claimed with an atomicMin on its level, as the reference claims
// with a cas on the parent (bfs.hpp:104-116).
__device__ __forceinline__ bool bfs_claim(uint32_t w, const uint64_t* __restrict__ vis, uint32_t* __restrict__ level,
                                          uint32_t L) {
  if ((vis[w >> 6] >> (w & 63)) & 1ull) return false;
  return level[w] == UNVIS && atomicMin(&level[w], L) == UNVIS;
}

__global__ void k_bfs_reset(bfs_counters* __restrict__ cnt) {
  cnt->next_count = 0;
  cnt->next_active = 0;
  cnt->heavy_count = 0;
  cnt->examined = 0;
}

// wave-aggregated append of the lanes in `mask`; returns this lane's slot
__device__ __forceinline__ unsigned long long wave_append(uint64_t mask,
                                                          unsigned long long* counter) {
  unsigned long long base = 0;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, 64);
  return base + wave_prefix(mask);
}

// Exclusive scan of the queued vertices' out-degrees into pref[0..n] (pref[n] = their
// sum), in blocks of DEG_BLOCK entries: per-block sums, a one-block scan of those sums,
// then each block recomputes its degrees and writes its offsets.
static constexpr uint32_t DEG_BLOCK = 1024;  // entries per block: 256 threads x 4

__device__ __forceinline__ uint64_t queued_degree(const uint32_t* __restrict__ queue, uint64_t n,
                                                  const uint64_t* __restrict__ ptr, uint64_t q) {
  if (q >= n) return 0ull;
  const uint32_t v = queue[q];
  return ptr[v + 1] - ptr[v];
}

// block-wide exclusive scan of one value per thread (256 threads); *total gets the sum
__device__ __forceinline__ uint64_t block_excl_scan256(uint64_t x, uint64_t* total) {
  __shared__ uint64_t s_wave[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t incl = x;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint64_t before = 0, sum = 0;
  for (int k = 0; k < 4; ++k) {
    before += k < wave ? s_wave[k] : 0ull;
    sum += s_wave[k];
  }
  __syncthreads();
  *total = sum;
  return before + incl - x;
}

__global__ __launch_bounds__(256) void k_deg_block_sums(const uint32_t* __restrict__ queue, uint64_t n,
                                                        const uint64_t* __restrict__ ptr, uint64_t* __restrict__ bsum) {
  const uint64_t base = blockIdx.x * (uint64_t)DEG_BLOCK;
  uint64_t d = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) d += queued_degree(queue, n, ptr, base + (uint64_t)k * 256 + threadIdx.x);
  uint64_t total = 0;
  (void)block_excl_scan256(d, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// one block: bsum[0..nb) -> its exclusive scan, in place
__global__ __launch_bounds__(256) void k_scan_block_sums(uint64_t* __restrict__ bsum, uint64_t nb) {
  uint64_t carry = 0;
  for (uint64_t c0 = 0; c0 < nb; c0 += 256) {
    const uint64_t i = c0 + threadIdx.x;
    const uint64_t x = i < nb ? bsum[i] : 0ull;
    uint64_t total = 0;
    const uint64_t ex = block_excl_scan256(x, &total);
    if (i < nb) bsum[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(256) void k_deg_scan_write(const uint32_t* __restrict__ queue, uint64_t n,
                                                        const uint64_t* __restrict__ ptr,
                                                        const uint64_t* __restrict__ boff, uint64_t* __restrict__ pref) {
  const uint64_t q0 = blockIdx.x * (uint64_t)DEG_BLOCK + (uint64_t)threadIdx.x * 4;
  uint64_t d[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) d[k] = queued_degree(queue, n, ptr, q0 + k);
  uint64_t total = 0;
  uint64_t off = boff[blockIdx.x] + block_excl_scan256(d[0] + d[1] + d[2] + d[3], &total);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (q0 + k < n) pref[q0 + k] = off;
    off += d[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) pref[n] = boff[blockIdx.x] + total;
}

// Edge-balanced top-down step.  The frontier's out-edges, numbered 0..T-1 through the
// exclusive scan pref[0..n] of the queued vertices' degrees (pref[n] = T), are cut into
// one contiguous range per wave, so a hub's edges spread over many waves and a run of
// low-degree vertices fills every lane.  A wave finds the vertex of its first edge with
// a 64-way search, then keeps a window of 64 queued vertices in registers (lane t: the
// first edge number of vertex wbase+t and its CSR offset) and resolves each lane's
// vertex with a 6-step shuffle search; the window moves only when the step's edges run
// past it.  Claims are set in the next bitmap and counted once per wave (no queue).
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
  return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(256) void k_bfs_push_eb(
    const uint32_t* __restrict__ queue, uint64_t n, const uint64_t* __restrict__ pref,
    const uint64_t* __restrict__ ptr, const uint32_t* __restrict__ idx, const uint64_t* __restrict__ vis,
    uint32_t* __restrict__ level, uint32_t L, uint64_t* __restrict__ next_bits, bfs_counters* __restrict__ cnt,
    const uint32_t* __restrict__ od) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id();
  const uint64_t nwaves = grid_waves();
  const uint64_t T = pref[n];
  const uint64_t per = ((T + nwaves - 1) / nwaves + 63) & ~63ull;
  const uint64_t e0 = wave_id * per;
  const uint64_t e1 = std::min<uint64_t>(T, e0 + per);
  unsigned long long active = 0, found = 0;
  if (e0 < e1) {
    // the vertex of edge e0: the last q in [0, n) with pref[q] <= e0 (pref[0] = 0)
    uint64_t lo = 0, hi = n - 1;
    while (lo < hi) {
      const uint64_t step = (hi - lo + 64) >> 6;
      const uint64_t p = lo + (uint64_t)(lane + 1) * step;
      const bool ok = p <= hi && pref[p] <= e0;
      const uint64_t nlo = lo + (uint64_t)__popcll(__ballot(ok)) * step;
      hi = std::min<uint64_t>(hi, nlo + step - 1);
      lo = nlo;
    }
    uint64_t wbase = lo;
    uint64_t wpref = 0, wsrc = 0, wlast = 0;  // pref[wbase+lane] (T past n), ptr[queue[.]], pref[wbase+64]
    auto load_window = [&] {
      const uint64_t q = wbase + lane;
      wpref = q < n ? pref[q] : T;
      wsrc = q < n ? ptr[queue[q]] : 0ull;
      wlast = pref[std::min<uint64_t>(wbase + 64, n)];
    };
    load_window();
    for (uint64_t j0 = e0; j0 < e1; j0 += 64) {
      if (j0 + 63 >= wlast && wlast < T) {
        // move the window to the vertex of edge j0 (in steps of up to 63 slots while j0
        // lies past it; zero-degree vertices can make a window hold fewer than 64 edges)
        do {
          const int t0 = __popcll(__ballot(wpref <= j0)) - 1;
          wbase += (uint64_t)t0;
          load_window();
        } while (j0 >= wlast);
      }
      const uint64_t j = j0 + lane;
      // window slot of edge j: the last t with wpref[t] <= j (all lanes take part)
      int t = 0;
#pragma unroll
      for (int sft = 32; sft > 0; sft >>= 1) {
        const uint64_t pv = shfl64(wpref, t + sft);
        if (pv <= j) t += sft;
      }
      const uint64_t tp = shfl64(wpref, t), ts = shfl64(wsrc, t);
      bool claimed = false;
      uint32_t w = 0;
      if (j < e1) {
        uint64_t i;
        if (j < wlast) {
          i = ts + (j - tp);
        } else {  // past the window (zero-degree vertices in it): walk on
          uint64_t q = wbase + 63;
          while (pref[q + 1] <= j) ++q;
          i = ptr[queue[q]] + (j - pref[q]);
        }
        w = idx[i];
        claimed = bfs_claim(w, vis, level, L);
      }
      if (claimed) {
        atomicOr((unsigned long long*)&next_bits[w >> 6], 1ull << (w & 63));
        active += od[w];
        found += 1;
      }
    }
  }
  block_add2(active, found, &cnt->next_active, &cnt->next_count);
}

__global__ __launch_bounds__(256) void k_bfs_push(
    const uint32_t* __restrict__ queue, uint64_t n, const uint64_t* __restrict__ ptr,
    const uint32_t* __restrict__ idx, const uint64_t* __restrict__ vis, uint32_t* __restrict__ level, uint32_t L,
    uint32_t* __restrict__ next_queue, uint64_t* __restrict__ next_bits,
    bfs_counters* __restrict__ cnt, const uint32_t* __restrict__ od) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id();
  const uint64_t nwaves = grid_waves();
  unsigned long long active = 0;
  for (uint64_t q = wave_id; q < n; q += nwaves) {
    const uint32_t v = queue[q];
    const uint64_t b = ptr[v], e = ptr[v + 1];
    for (uint64_t base = b; base < e; base += 64) {
      const uint64_t i = base + lane;
      bool claimed = false;
      uint32_t w = 0;
      if (i < e) {
        w = idx[i];
        claimed = bfs_claim(w, vis, level, L);
      }
      const uint64_t m = __ballot(claimed);
      if (m) {
        const unsigned long long pos = wave_append(m, &cnt->next_count);
        if (claimed) {
          next_queue[pos] = w;
          atomicOr((unsigned long long*)&next_bits[w >> 6], 1ull << (w & 63));
          active += od[w];
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) active += __shfl_down(active, off, 64);
  if (lane == 0 && active) atomicAdd(&cnt->next_active, active);
}

// first_in[v] = lowest in-neighbour of v (the first of its ascending in-edge list:
// with degree relabeling, its highest-degree one), UNVIS for none
__global__ void k_first_in(const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ row_idx, uint32_t V,
                           uint32_t* __restrict__ first_in) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = col_ptr[v];
    first_in[v] = b < col_ptr[v + 1] ? row_idx[b] : UNVIS;
  }
}

// Bottom-up step.  The first PULL_HOT_WORDS words of the frontier bitmap (ids below
// 64 * PULL_HOT_WORDS) are staged in LDS once per block: with degree relabeling those
// are the hubs, which most in-edge lists start with, so most membership tests hit LDS.
static constexpr int PULL_SCAN_UNROLL = 4;  // in-edges a pull lane loads per round trip
#ifndef PULL_BLOCKS_PER_CU
#define PULL_BLOCKS_PER_CU 2  // two 1024-thread blocks, the most a CU holds
#endif
static constexpr uint32_t PULL_THREADS = 1024;
#ifndef PULL_HOT_WORDS
#define PULL_HOT_WORDS 8192  // 64 KiB: ids below 2^19 (two blocks per CU fit in LDS)
#endif

__global__ __launch_bounds__(PULL_THREADS) void k_bfs_pull(
    const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ row_idx,
    const uint32_t* __restrict__ first_in, uint32_t V, uint64_t nwords, uint64_t* __restrict__ vis,
    uint32_t* __restrict__ level,
    const uint64_t* __restrict__ cur_bits, uint32_t L, uint64_t* __restrict__ next_bits,
    uint64_t* __restrict__ heavy_bits, bfs_counters* __restrict__ cnt, const uint32_t* __restrict__ od) {
  __shared__ uint64_t hot[PULL_HOT_WORDS];
  const uint64_t nh = nwords < (uint64_t)PULL_HOT_WORDS ? nwords : (uint64_t)PULL_HOT_WORDS;
  for (uint64_t i = threadIdx.x; i < nh; i += blockDim.x) hot[i] = cur_bits[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  unsigned long long active = 0, found = 0;
  uint32_t examined = 0;  // a lane's share stays far below 2^32 (at most PULL_CAP per vertex)
  for (uint64_t word = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; word < nwords;
       word += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t v = word * 64 + lane;
    // visited before this step: the visited bitmap and the current frontier; the frontier
    // joins the bitmap here (this wave alone owns the word in this kernel)
    const uint64_t cw = word < nh ? hot[word] : cur_bits[word];
    const uint64_t seen = vis[word] | cw;
    bool claimed = false, is_heavy = false;
    if (v < V && !((seen >> lane) & 1ull)) {
      // the first in-edge from the dense first_in array (one coalesced word per lane);
      // the list is read only when that neighbour is not in the frontier
      const uint32_t f = first_in[v];
      if (f != UNVIS) {
        const uint64_t fw = (f >> 6) < nh ? hot[f >> 6] : cur_bits[f >> 6];
        examined += 1;  // the list's first entry
        if ((fw >> (f & 63)) & 1ull) {
          claimed = true;
        } else {
          const uint64_t b = col_ptr[v], e = col_ptr[v + 1];
          const uint64_t lim = (e - b > PULL_CAP) ? b + PULL_CAP : e;
          // PULL_SCAN_UNROLL in-edges per round trip: their ids are loaded together, then tested
          for (uint64_t i = b + 1; i < lim && !claimed; i += PULL_SCAN_UNROLL) {
            uint32_t us[PULL_SCAN_UNROLL];
#pragma unroll
            for (int k = 0; k < PULL_SCAN_UNROLL; ++k) us[k] = i + k < lim ? row_idx[i + k] : UNVIS;
            uint32_t hits = 0;  // bit k: entry i + k is in the frontier
#pragma unroll
            for (int k = 0; k < PULL_SCAN_UNROLL; ++k) {
              const uint32_t u = us[k];
              if (u != UNVIS) {
                const uint64_t wd = (u >> 6) < nh ? hot[u >> 6] : cur_bits[u >> 6];
                hits |= (uint32_t)((wd >> (u & 63)) & 1ull) << k;
              }
            }
            if (hits) {
              claimed = true;
              examined += (uint32_t)(i - b) + (uint32_t)__builtin_ctz(hits);  // entries b+1 .. the first hit
            }
          }
          if (!claimed) examined += (uint32_t)(lim - b - 1);  // every entry up to the cap
          is_heavy = !claimed && e > lim;
        }
      }
    }
    // the next frontier and the vertices left for the wave-per-vertex pass as bitmap
    // words (one store per wave, no atomics); counts are summed once per wave at the end
    const uint64_t m = __ballot(claimed);
    const uint64_t hm = __ballot(is_heavy);
    if (lane == 0) {
      next_bits[word] = m;
      heavy_bits[word] = hm;
      if (cw) vis[word] = seen;
      found += (unsigned long long)__popcll(m);
    }
    if (claimed) {
      level[v] = L;
      active += od[v];
    }
  }
  block_add3(active, found, (unsigned long long)examined, &cnt->next_active, &cnt->next_count, &cnt->examined);
}

// second pull pass: one wave per heavy vertex, 64 in-edges per step, stop at the first hit
__global__ __launch_bounds__(256) void k_bfs_pull_heavy(
    const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ row_idx,
    uint32_t* __restrict__ level, const uint64_t* __restrict__ cur_bits, uint32_t L,
    uint64_t* __restrict__ next_bits, const uint32_t* __restrict__ heavy, bfs_counters* __restrict__ cnt,
    const uint32_t* __restrict__ od) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id();
  const uint64_t nwaves = grid_waves();
  const unsigned long long n = cnt->heavy_count;
  unsigned long long active = 0, examined = 0;
  for (uint64_t h = wave_id; h < n; h += nwaves) {
    const uint32_t v = heavy[h];
    const uint64_t b = col_ptr[v] + PULL_CAP, e = col_ptr[v + 1];
    bool found = false;
    uint64_t scanned = e - b;  // entries past the cap a sequential scan reads
    for (uint64_t base = b; base < e && !found; base += 64) {
      const uint64_t i = base + lane;
      bool hit = false;
      if (i < e) {
        const uint32_t u = row_idx[i];
        hit = (cur_bits[u >> 6] >> (u & 63)) & 1ull;
      }
      const uint64_t hits = __ballot(hit);
      found = hits != 0ull;
      if (found) scanned = base - b + (uint64_t)__builtin_ctzll(hits) + 1;
    }
    if (lane == 0) examined += scanned;
    if (found && lane == 0) {
      level[v] = L;
      atomicOr((unsigned long long*)&next_bits[v >> 6], 1ull << (v & 63));
      atomicAdd(&cnt->next_count, 1ull);
      active += od[v];
    }
  }
  if (lane == 0 && active) atomicAdd(&cnt->next_active, active);
  if (lane == 0 && examined) atomicAdd(&cnt->examined, examined);
}

static constexpr unsigned COMPACT_BLOCKS_PER_CU = 4;

// Frontier bitmap -> queue.  Each block owns one contiguous range of words: it counts
// the range's set bits, reserves that many queue slots with one atomic, then writes the
// ids tile by tile (256 words, one per thread) at offsets from a block-wide scan.  One
// atomic per block: a counter that every wave bumps serialises in the memory-side atomic
// unit.  Ids come out ascending inside each block's slice.
__global__ __launch_bounds__(256) void k_frontier_compact(const uint64_t* __restrict__ words, uint64_t nwords,
                                                          uint32_t V, uint32_t* __restrict__ queue,
                                                          unsigned long long* __restrict__ count) {
  __shared__ unsigned long long s_part[4];
  __shared__ unsigned s_wave[4];
  __shared__ unsigned long long s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t vwords = std::min<uint64_t>(nwords, (uint64_t)(V >> 6) + ((V & 63) ? 1 : 0));
  const uint64_t per = (vwords + gridDim.x - 1) / gridDim.x;
  const uint64_t w0 = std::min<uint64_t>(vwords, blockIdx.x * per), w1 = std::min<uint64_t>(vwords, w0 + per);
  auto word_at = [&](uint64_t wi) -> uint64_t {
    if (wi >= w1) return 0ull;
    uint64_t w = words[wi];
    if (wi == (uint64_t)(V >> 6) && (V & 63)) w &= (1ull << (V & 63)) - 1ull;  // bits >= V are not vertices
    return w;
  };
  unsigned long long c = 0;
  for (uint64_t wi = w0 + threadIdx.x; wi < w1; wi += 256) c += (unsigned long long)__popcll(word_at(wi));
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if (lane == 0) s_part[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long total = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    s_base = total ? atomicAdd(count, total) : 0ull;
  }
  __syncthreads();
  unsigned long long pos = s_base;
  for (uint64_t t0 = w0; t0 < w1; t0 += 256) {
    const uint64_t wi = t0 + threadIdx.x;
    uint64_t w = word_at(wi);
    const unsigned cnt = (unsigned)__popcll(w);
    unsigned incl = cnt;
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    unsigned before = 0, tile = 0;
    for (int k = 0; k < 4; ++k) {
      const unsigned x = s_wave[k];
      before += k < wave ? x : 0u;
      tile += x;
    }
    __syncthreads();  // s_wave is rewritten by the next tile
    uint64_t p = pos + before + incl - cnt;
    while (w) {
      queue[p++] = (uint32_t)(wi * 64 + (uint64_t)(__ffsll((unsigned long long)w) - 1));
      w &= w - 1;
    }
    pos += tile;
  }
}

// Compacts the set bits of a bitmap (ids below V) into queue[*count ...]; *count
// must be zero or hold the entries already queued.
void frontier_compact(pe_ctx_s* ctx, const uint64_t* words, uint64_t nwords, uint32_t V, uint32_t* queue,
                      unsigned long long* count) {
  const uint64_t vwords = std::min<uint64_t>(nwords, (uint64_t)(V >> 6) + ((V & 63) ? 1 : 0));
  const unsigned grid = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((uint64_t)std::max(1, ctx->num_cus) * COMPACT_BLOCKS_PER_CU, (vwords + 255) / 256));
  k_frontier_compact<<<grid, 256, 0, ctx->stream>>>(words, nwords, V, queue, count);
  PE_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_reached_edges(const uint32_t* __restrict__ level,
                                                       const uint32_t* __restrict__ od, uint32_t V,
                                                       unsigned long long* __restrict__ out) {
  unsigned long long acc = 0, cntv = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    if (level[v] != UNVIS) { acc += od[v]; ++cntv; }
  block_add2(acc, cntv, out, out + 1);
}

// Device state of one BFS engine, reused across roots.
// top-down levels with at most this many active edges append to a queue (one wave per
// vertex); larger ones run edge-balanced and leave the next frontier as a bitmap
#ifndef PUSH_APPEND_EDGES
#define PUSH_APPEND_EDGES 4096
#endif
static constexpr uint64_t EB_WAVE_EDGES = 512;  // fewest edges an edge-balanced wave takes

struct bfs_state {
  pe_graph_s* g;
  uint64_t nwords;
  dbuf<uint32_t> level, q0, q1, heavy;
  dbuf<uint64_t> hb;  // vertices left for the wave-per-vertex pull pass, as a bitmap
  dbuf<uint64_t> b0, b1;
  // vertices visited before the current frontier (with it: all visited so far);
  // the frontier joins it at the start of each step
  dbuf<uint64_t> vis;
  dbuf<bfs_counters> cnt;  // two sets: consecutive steps alternate
  dbuf<unsigned long long> red;
  dbuf<uint64_t> pref;  // scan of the queued vertices' out-degrees (edge-balanced push)
  dbuf<uint64_t> bsum;  // its per-block sums

  explicit bfs_state(pe_graph_s* graph) : g(nullptr) {
    bind(graph);
    pe_ctx_s* ctx = g->ctx;
    const uint32_t V = std::max<uint32_t>(1, g->V);
    nwords = ((uint64_t)g->V >> 6) + 1;
    level.alloc(ctx, V);
    q0.alloc(ctx, V);
    q1.alloc(ctx, V);
    heavy.alloc(ctx, V);
    hb.alloc(ctx, nwords);
    b0.alloc(ctx, nwords);
    b1.alloc(ctx, nwords);
    vis.alloc(ctx, nwords);
    cnt.alloc(ctx, 2);
    red.alloc(ctx, 2);
    pref.alloc(ctx, (uint64_t)V + 1);
    bsum.alloc(ctx, ((uint64_t)V + DEG_BLOCK - 1) / DEG_BLOCK);
  }

  // Walk the edges of `graph` from now on (a graph over the same vertex ids).
  void bind(pe_graph_s* graph) {
    g = graph;
    if (!g->first_in.p) {  // kept with the graph for later searches
      g->first_in.alloc(g->ctx, std::max<uint32_t>(1, g->V));
      k_first_in<<<grid_for(g->V, 256, 8192), 256, 0, g->ctx->stream>>>(g->col_ptr.p, g->row_idx.p, g->V,
                                                                         g->first_in.p);
      PE_HIP(hipGetLastError());
    }
  }

  // One top-down level from the queue qc (n vertices, A active edges; A = 0 if unknown)
  // at level L.  Returns whether the next frontier was queued into qn as well as set in bn.
  bool push_level(const uint32_t* qc, uint64_t n, uint64_t A, uint32_t L, uint32_t* qn, uint64_t* bn,
                  bfs_counters* c) {
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const unsigned push_grid = persistent_grid(ctx, 1u << 20, 8);
    if (A != 0 && A <= PUSH_APPEND_EDGES) {
      const unsigned grid = (unsigned)std::min<uint64_t>(push_grid, (n + 3) / 4);
      k_bfs_push<<<grid, 256, 0, s>>>(qc, n, g->csr_ptr(), g->csr_idx(), vis.p, level.p, L, qn, bn, c,
                                      g->od.p);
      PE_HIP(hipGetLastError());
      return true;
    }
    PE_REQUIRE(n > 0, PE_ERR_INTERNAL, "top-down step over an empty frontier");
    const uint64_t nb = (n + DEG_BLOCK - 1) / DEG_BLOCK;
    k_deg_block_sums<<<(unsigned)nb, 256, 0, s>>>(qc, n, g->csr_ptr(), bsum.p);
    k_scan_block_sums<<<1, 256, 0, s>>>(bsum.p, nb);
    k_deg_scan_write<<<(unsigned)nb, 256, 0, s>>>(qc, n, g->csr_ptr(), bsum.p, pref.p);
    const uint64_t waves = A ? std::max<uint64_t>(1, (A + EB_WAVE_EDGES - 1) / EB_WAVE_EDGES) : (uint64_t)push_grid * 4;
    const unsigned grid = (unsigned)std::min<uint64_t>(push_grid, (waves + 3) / 4);
    k_bfs_push_eb<<<grid, 256, 0, s>>>(qc, n, pref.p, g->csr_ptr(), g->csr_idx(), vis.p, level.p, L, bn, c,
                                       g->od.p);
    PE_HIP(hipGetLastError());
    return false;
  }

  // Algorithmic bytes of the search (counted per level): push 8|F| + 4 sum_F deg + 4|new|
  // (offsets, edge ids, level writes), pull 8|U| + 4 examined + V/8 (offsets of the unvisited
  // vertices, the in-edges a sequential scan reads up to its first hit, the next bitmap).
  uint64_t alg_bytes = 0;

  int run(uint32_t root, double threshold, uint8_t* modes, uint64_t* frontier, uint64_t* actv,
          uint32_t max_epochs, double* seconds) {
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint64_t Ep = g->directed ? g->edges : g->edges * 2;  // dualmode.hpp:261-262
    const unsigned push_grid = persistent_grid(ctx, 1u << 20, 8);
    const unsigned pull_grid = persistent_grid(ctx, grid_for(nwords * 64, PULL_THREADS), PULL_BLOCKS_PER_CU);
    const unsigned heavy_grid = persistent_grid(ctx, 1u << 20, 4);
    PE_HIP(hipEventRecord(ctx->t0, s));
    k_bfs_init<<<grid_for(g->V, 256, 8192), 256, 0, s>>>(level.p, vis.p, g->V, nwords, root, b0.p, nwords, q0.p,
                                                         cnt.p, g->od.p);
    // counters: step e adds into set (e + 1) & 1 (the root's are in set 0); the kernel that
    // mails a step's counters to the host zeroes the next step's set
    auto set = [&](int e) { return cnt.p + ((e + 1) & 1); };
    constexpr uint32_t CW = sizeof(bfs_counters) / 8;
    PE_HIP(hipGetLastError());
    const bfs_counters* h = reinterpret_cast<const bfs_counters*>(
        mail_post(ctx, reinterpret_cast<const unsigned long long*>(set(-1)), CW,
                  reinterpret_cast<unsigned long long*>(set(0)), CW));
    uint64_t n = h->next_count, A = h->next_active;
    uint32_t* qc = q0.p;
    uint32_t* qn = q1.p;
    uint64_t* bc = b0.p;
    uint64_t* bn = b1.p;
    int epoch = 0;
    bool queued = true;  // qc holds the current frontier as a queue
    alg_bytes = 0;
    uint64_t visited = n;  // visited so far, the current frontier included
    while (n != 0) {
      const bool push = ((double)A / (double)Ep) < threshold;  // dualmode.hpp:279
      if ((uint32_t)epoch < max_epochs) {
        if (modes) modes[epoch] = push ? 1 : 0;
        if (frontier) frontier[epoch] = n;
        if (actv) actv[epoch] = A;
      }
      const uint32_t L = (uint32_t)epoch + 1;
      bfs_counters* c = set(epoch);
      if (push) {
        PE_REQUIRE(g->has_csr, PE_ERR_STATE, "push mode needs out-edges (PE_GRAPH_BUILD_CSR)");
        k_bfs_push_prep<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(vis.p, bc, nwords, bn, nwords);
        ktime_scope kt(ctx, PE_KT_BFS_PUSH);
        // a step that left the frontier as a bitmap only: queue it (heavy_count as scratch)
        if (!queued) frontier_compact(ctx, bc, nwords, g->V, qc, &c->heavy_count);
        queued = push_level(qc, n, A, L, qn, bn, c);
      } else {
        ktime_scope kt(ctx, PE_KT_BFS_PULL);
        k_bfs_pull<<<pull_grid, PULL_THREADS, 0, s>>>(g->col_ptr.p, g->row_idx.p, g->first_in.p, g->V, nwords,
                                                      vis.p, level.p, bc, L, bn, hb.p, c, g->od.p);
        frontier_compact(ctx, hb.p, nwords, g->V, heavy.p, &c->heavy_count);
        k_bfs_pull_heavy<<<heavy_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, level.p, bc, L, bn, heavy.p, c,
                                                    g->od.p);
        queued = false;
      }
      PE_HIP(hipGetLastError());
      h = reinterpret_cast<const bfs_counters*>(mail_post(ctx, reinterpret_cast<const unsigned long long*>(c), CW,
                                                          reinterpret_cast<unsigned long long*>(set(epoch + 1)), CW));
      if (push) alg_bytes += 8 * n + 4 * A + 4 * h->next_count;
      else alg_bytes += 8 * ((uint64_t)g->V - visited) + 4 * h->examined + (uint64_t)g->V / 8;
      n = h->next_count;
      A = h->next_active;
      visited += n;
      std::swap(qc, qn);
      std::swap(bc, bn);
      ++epoch;
      PE_REQUIRE(epoch <= (int)g->V + 1, PE_ERR_INTERNAL, "BFS did not terminate");
    }
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    if (seconds) *seconds = ms * 1e-3;
    return epoch;
  }
};

}  // namespace pe

using namespace pe;

extern "C" {

}  // extern "C"

struct pe_bfs_s {
  pe_graph_s* g;
  pe_ctx_s* ctx = nullptr;  // the graph's context; destroy uses it, not the graph (which may be gone)
  std::unique_ptr<bfs_state> st;
  uint32_t epoch = 0;
};

extern "C" {

int pe_bfs_create(pe_graph_t* g, pe_bfs_t** out) {
  if (!g || !out) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(g->ctx, [&] {
    auto b = std::make_unique<pe_bfs_s>();
    b->g = g;
    b->ctx = g->ctx;
    b->st = std::make_unique<bfs_state>(g);
    *out = b.release();
  });
}

int pe_bfs_destroy(pe_bfs_t* b) {
  if (!b) return PE_ERR_INVALID;
  (void)hipStreamSynchronize(b->ctx->stream);
  delete b;
  return PE_OK;
}

int pe_bfs_reset(pe_bfs_t* b, uint32_t root, pe_bitmap_t* frontier) {
  if (!b || !frontier) return PE_ERR_INVALID;
  pe_graph_s* g = b->g;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(root < g->V, PE_ERR_RANGE, "root >= V");
    PE_REQUIRE(frontier->bits >= g->V, PE_ERR_INVALID, "bitmap smaller than the vertex set");
    uint32_t iroot = root;
    to_internal_ids(g, &root, 1, &iroot);
    hipStream_t s = g->ctx->stream;
    k_bfs_init<<<grid_for(g->V, 256, 8192), 256, 0, s>>>(b->st->level.p, b->st->vis.p, g->V, b->st->nwords, iroot,
                                                         frontier->p, frontier->words, b->st->q0.p, b->st->cnt.p,
                                                         g->od.p);
    PE_HIP(hipGetLastError());
    PE_HIP(hipStreamSynchronize(s));
    b->epoch = 0;
  });
}

int pe_bfs_step(pe_bfs_t* b, int push, pe_bitmap_t* current, pe_bitmap_t* next, uint64_t* activated,
                uint64_t* next_active_edges) {
  if (!b || !current || !next) return PE_ERR_INVALID;
  pe_graph_s* g = b->g;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(current->bits >= g->V && next->bits >= g->V, PE_ERR_INVALID, "bitmap smaller than the vertex set");
    PE_REQUIRE(current != next, PE_ERR_INVALID, "current and next must differ");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    bfs_state& st = *b->st;
    const uint32_t L = b->epoch + 1;
    const uint64_t nwords = ((uint64_t)g->V >> 6) + 1;
    k_bfs_reset<<<1, 1, 0, s>>>(st.cnt.p);
    if (push) {
      PE_REQUIRE(g->has_csr, PE_ERR_STATE, "push mode needs out-edges (PE_GRAPH_BUILD_CSR)");
      // compact the frontier bitmap into a queue (count in heavy_count, reused as scratch)
      frontier_compact(ctx, current->p, nwords, g->V, st.q0.p, &st.cnt.p->heavy_count);
      k_bfs_push_prep<<<grid_for(std::max(nwords, next->words), 256, 4096), 256, 0, s>>>(st.vis.p, current->p, nwords,
                                                                                          next->p, next->words);
      PE_HIP(hipGetLastError());
      bfs_counters h{};
      PE_HIP(hipMemcpyAsync(&h, st.cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      ktime_scope kt(ctx, PE_KT_BFS_PUSH);
      st.push_level(st.q0.p, h.heavy_count, 0, L, st.q1.p, next->p, st.cnt.p);
    } else {
      ktime_scope kt(ctx, PE_KT_BFS_PULL);
      k_bfs_pull<<<persistent_grid(ctx, grid_for(nwords * 64, PULL_THREADS), PULL_BLOCKS_PER_CU), PULL_THREADS, 0,
                   s>>>(g->col_ptr.p, g->row_idx.p, g->first_in.p, g->V, nwords, st.vis.p, st.level.p, current->p,
                        L, next->p, st.hb.p, st.cnt.p, g->od.p);
      frontier_compact(ctx, st.hb.p, nwords, g->V, st.heavy.p, &st.cnt.p->heavy_count);
      k_bfs_pull_heavy<<<persistent_grid(ctx, 1u << 20, 4), 256, 0, s>>>(
          g->col_ptr.p, g->row_idx.p, st.level.p, current->p, L, next->p, st.heavy.p, st.cnt.p, g->od.p);
    }
    PE_HIP(hipGetLastError());
    bfs_counters h{};
    PE_HIP(hipMemcpyAsync(&h, st.cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    if (activated) *activated = h.next_count;
    if (next_active_edges) *next_active_edges = h.next_active;
    b->epoch += 1;
  });
}

int pe_bfs_get_levels(pe_bfs_t* b, int32_t* levels) {
  if (!b || !levels) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] { download_external(b->g, b->st->level.p, levels, 4); });
}

int pe_bfs_bind(pe_bfs_t* b, pe_graph_t* g) {
  if (!b || !g) return PE_ERR_INVALID;
  return guarded(b->g->ctx, [&] {
    PE_REQUIRE(g->ctx == b->g->ctx, PE_ERR_INVALID, "graph of another context");
    PE_REQUIRE(g->id_space == b->g->id_space && g->V == b->g->V, PE_ERR_INVALID,
               "graph over other vertex ids (bind g or its reverse made with PE_GRAPH_SAME_IDS)");
    b->g = g;
    b->st->bind(g);
    PE_HIP(hipStreamSynchronize(g->ctx->stream));
  });
}

int pe_bfs(pe_graph_t* g, uint32_t root, double threshold, int32_t* levels, uint8_t* modes,
           uint64_t* frontier, uint64_t* active_edges, uint32_t max_epochs, pe_bfs_stats_t* stats) {
  if (!g) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(root < g->V, PE_ERR_RANGE, "root >= V");
    uint32_t iroot = root;
    to_internal_ids(g, &root, 1, &iroot);
    bfs_state st(g);
    double secs = 0.0;
    const int epochs = st.run(iroot, threshold, modes, frontier, active_edges, max_epochs, &secs);
    hipStream_t s = g->ctx->stream;
    if (levels) download_external(g, st.level.p, levels, 4);
    PE_HIP(hipMemsetAsync(st.red.p, 0, 16, s));
    k_reached_edges<<<grid_for(g->V, 256, 1024), 256, 0, s>>>(st.level.p, g->od.p, g->V, st.red.p);
    PE_HIP(hipGetLastError());
    unsigned long long r[2] = {0, 0};
    PE_HIP(hipMemcpyAsync(r, st.red.p, 16, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    if (stats) {
      stats->epochs = (uint32_t)epochs;
      stats->visited = (uint32_t)(r[1] & 0xFFFFFFFFull);
      stats->reached_edges = r[0];
      stats->seconds = secs;
      stats->alg_bytes = st.alg_bytes;
    }
  });
}

}  // extern "C"
