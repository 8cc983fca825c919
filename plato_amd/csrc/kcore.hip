// k-core decomposition: plato::algo::kcore_algo_t::compute_shell_index
// (plato/algo/kcore/kcore.hpp:80-268) on the device.
//
// The reference iterates the h-index operator of Montresor et al. (IEEE TPDS 2012)
// synchronously: coreness starts at the adjacency length (0xFFFFFFFF for vertices
// without one), every vertex is active in epoch 0, and in each epoch every active
// vertex v with est = min(len(v), estimate[v]) takes the largest i in [1, est] such
// that at least i neighbours u have min(est, estimate[u]) >= i, when that i is below
// its coreness (v is then modified and its neighbours become active for the next
// epoch); a vertex with no neighbours gets 0; when no i qualifies the value is kept.
// Estimates are the previous epoch's values (Jacobi), so the result and the per-epoch
// counts do not depend on the order of the vertices: they are compared bit for bit
// with the oracle.
//
// Per epoch: the active bitmap is compacted into a queue; one lane per vertex of at
// most KC_SMALL neighbours counts them in a byte histogram of its own in LDS; longer
// adjacencies whose answer cannot exceed KC_HIST (min(length, estimate) <= KC_HIST) go to
// a second queue handled one wave per vertex with a KC_HIST-bin LDS histogram, and the
// others (adjacencies above KC_HUGE, or estimates above KC_HIST) to a third, one
// 1024-thread block per vertex with a KC_HHIST-bin histogram (when the answer reaches
// its last bin, a binary search over [KC_HHIST, est] with block-wide counts finishes
// the job); adjacencies above KC_GIANT are cut into chunks whose block histograms add
// up in global memory, so no block walks a giant alone.
// Modified vertices set their neighbours' bits in the next bitmap (test, then atomic
// OR, as the reference's get_bit / set_bit).  Then estimate = coreness for the queue.
#include <algorithm>
#include <vector>

#include "pe_internal.hpp"

namespace pe {

static constexpr uint32_t KC_UNSET = 0xFFFFFFFFu;
static constexpr uint32_t KC_SMALL = 64;    // neighbours handled by one lane
static constexpr uint32_t KC_HIST = 1024;   // histogram bins 0..KC_HIST of the wave-per-vertex pass
static constexpr uint32_t KC_PER = 17;      // bins per lane in its scan (64 * 17 > KC_HIST)
static constexpr int KC_THREADS = 256;
static constexpr uint32_t KC_HUGE = 4096;    // longer adjacencies take one block each
static constexpr int KC_HTHREADS = 1024;
static constexpr uint32_t KC_HHIST = 16384;  // histogram bins 0..KC_HHIST of the block pass
static constexpr uint32_t KC_HPER = 17;      // bins per thread in its scan (1024 * 17 > KC_HHIST)
static constexpr uint64_t KC_GIANT = 65536;  // longer adjacencies are cut into chunks over many blocks
static constexpr uint64_t KC_GCHUNK = 16384;

struct kc_counters {
  unsigned long long queued;    // active vertices this epoch
  unsigned long long nbig;      // of them, with KC_SMALL < neighbours <= KC_HUGE and min(neighbours, estimate) <= KC_HIST
  unsigned long long nhuge;     // of them, the other ones above KC_SMALL neighbours
  unsigned long long modified;  // vertices whose coreness dropped
};

__global__ void k_kc_init(const uint64_t* __restrict__ ptr, uint32_t V, uint32_t* __restrict__ core,
                          uint32_t* __restrict__ est) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t len = ptr[v + 1] - ptr[v];
    const uint32_t c = len ? (uint32_t)std::min<uint64_t>(len, 0xFFFFFFFEull) : KC_UNSET;
    core[v] = c;
    est[v] = c;
  }
}

__global__ void k_kc_fill(uint64_t* __restrict__ w, uint64_t nwords, uint32_t V) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nwords; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = i * 64;
    uint64_t m = 0ull;
    if (lo + 64 <= V) m = ~0ull;
    else if (lo < V) m = (1ull << (V - lo)) - 1ull;
    w[i] = m;
  }
}

__global__ void k_kc_zero(uint64_t* __restrict__ w, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    w[i] = 0ull;
}

__global__ void k_kc_reset(kc_counters* c) {
  c->queued = 0;
  c->nbig = 0;
  c->nhuge = 0;
  c->modified = 0;
}

// the reference's `if (!get_bit(u)) set_bit(u)` on the next active bitmap
__device__ __forceinline__ void kc_activate(uint64_t* next_bits, uint32_t u) {
  const uint64_t m = 1ull << (u & 63);
  if (!(next_bits[u >> 6] & m)) atomicOr((unsigned long long*)&next_bits[u >> 6], (unsigned long long)m);
}

// One lane per queued vertex of at most KC_SMALL neighbours (byte counts suffice).
// Grid-stride over the queue; longer adjacencies are marked in the big / huge bitmaps
// (one atomic OR per marked vertex, on its own word) and queued by compaction after the
// pass, and the modified count is summed once per wave: no per-wave atomics on a shared
// counter.
__global__ __launch_bounds__(KC_THREADS) void k_kc_small(
    const uint32_t* __restrict__ queue, uint64_t n, const uint64_t* __restrict__ ptr,
    const uint32_t* __restrict__ idx, const uint32_t* __restrict__ est, uint32_t* __restrict__ core,
    uint64_t* __restrict__ next_bits, uint64_t* __restrict__ big_bits, uint64_t* __restrict__ huge_bits,
    kc_counters* __restrict__ cnt) {
  __shared__ uint8_t hist[KC_THREADS][KC_SMALL + 1];
  uint8_t* h = hist[threadIdx.x];
  unsigned long long nmod = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = queue[i];
    const uint64_t b = ptr[v], e = ptr[v + 1];
    const uint64_t len = e - b;
    if (len == 0) {
      core[v] = 0;  // kcore.hpp:214-218
    } else if (len > KC_SMALL) {
      // the block pass also takes the vertices whose answer may exceed the wave pass's
      // KC_HIST bins (which would need its binary search over the adjacency)
      const bool block = len > KC_HUGE || std::min<uint64_t>(len, est[v]) > KC_HIST;
      atomicOr((unsigned long long*)&(block ? huge_bits : big_bits)[v >> 6], 1ull << (v & 63));
    } else {
      const uint32_t cap = std::min<uint32_t>((uint32_t)len, est[v]);
      for (uint32_t j = 0; j <= cap; ++j) h[j] = 0;
      for (uint64_t k = b; k < e; ++k) {
        const uint32_t eu = est[idx[k]];
        ++h[eu < cap ? eu : cap];
      }
      uint32_t sum = 0, best = 0;
      for (uint32_t j = cap; j > 0; --j) {
        sum += h[j];
        if (sum >= j) {
          best = j;
          break;
        }
      }
      if (best && best < core[v]) {
        core[v] = best;
        ++nmod;
        for (uint64_t k = b; k < e; ++k) kc_activate(next_bits, idx[k]);
      }
    }
  }
  block_add2(nmod, 0, &cnt->modified, &cnt->modified);
}

// One wave per queued vertex with more than KC_SMALL neighbours.
__global__ __launch_bounds__(KC_THREADS) void k_kc_big(
    const uint32_t* __restrict__ big, const kc_counters* __restrict__ cnt_in, const uint64_t* __restrict__ ptr,
    const uint32_t* __restrict__ idx, const uint32_t* __restrict__ est, uint32_t* __restrict__ core,
    uint64_t* __restrict__ next_bits, kc_counters* __restrict__ cnt) {
  __shared__ uint32_t hist_all[KC_THREADS / 64][64 * KC_PER];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t* hist = hist_all[wave];
  const uint64_t nbig = cnt_in->nbig;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  uint32_t nmod = 0;
  for (uint64_t q = wave_id; q < nbig; q += nwaves) {
    const uint32_t v = big[q];
    const uint64_t b = ptr[v], e = ptr[v + 1];
    const uint32_t cap = (uint32_t)std::min<uint64_t>(e - b, (uint64_t)est[v]);
    const uint32_t top = std::min(cap, KC_HIST);
    for (uint32_t j = lane; j <= top; j += 64) hist[j] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint64_t k = b + lane; k < e; k += 64) {
      const uint32_t eu = est[idx[k]];
      atomicAdd(&hist[eu < top ? eu : top], 1u);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // largest i in [1, top] with (count of bins >= i) >= i: lane l owns bins
    // [PER l, PER l + PER), and the bins of the lanes above it are suffix-summed
    const uint32_t j0 = (uint32_t)lane * KC_PER;
    uint32_t own = 0;
    for (uint32_t j = j0; j < j0 + KC_PER && j <= top; ++j) own += hist[j];
    uint32_t incl = own;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_down(incl, off, 64);
      if (lane + off < 64) incl += t;
    }
    uint32_t best = 0, sum = incl - own;
    for (int j = (int)std::min(j0 + KC_PER - 1, top); j >= (int)j0 && j >= 1; --j) {
      sum += hist[j];
      if (sum >= (uint32_t)j) {
        best = (uint32_t)j;
        break;
      }
    }
    for (int off = 32; off > 0; off >>= 1) best = std::max(best, (uint32_t)__shfl_xor(best, off, 64));
    if (best && best < core[v]) {
      if (lane == 0) core[v] = best;
      for (uint64_t k = b + lane; k < e; k += 64) kc_activate(next_bits, idx[k]);
      ++nmod;
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0 && nmod) atomicAdd(&cnt->modified, (unsigned long long)nmod);
}

// Block-wide h-index from a histogram in LDS (hist[j] = neighbours whose estimate, capped
// at top, is j, for j <= top = min(cap, KC_HHIST)): the largest j in [1, top] whose suffix
// sum reaches j; when that is KC_HHIST and cap is larger, a binary search on
// [KC_HHIST, cap] with block-wide counts over the adjacency [b, e) finishes it.  0 when
// no j qualifies.  Every thread of the block calls it and gets the answer.
__device__ uint32_t kc_block_hindex(const uint32_t* hist, uint32_t top, uint32_t cap, uint64_t b, uint64_t e,
                                    const uint32_t* __restrict__ idx, const uint32_t* __restrict__ est) {
  __shared__ uint32_t s_wave[KC_HTHREADS / 64];
  __shared__ uint32_t s_best;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int WAVES = KC_HTHREADS / 64;
  // thread t owns bins [17 t, 17 t + 17); suffix sums over the threads above it
  const uint32_t j0 = (uint32_t)tid * KC_HPER;
  uint32_t own = 0;
  for (uint32_t j = j0; j < j0 + KC_HPER && j <= top; ++j) own += hist[j];
  uint32_t incl = own;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_down(incl, off, 64);
    if (lane + off < 64) incl += t;
  }
  if (lane == 0) s_wave[wave] = incl;  // the wave's total
  if (tid == 0) s_best = 0;
  __syncthreads();
  uint32_t sum = incl - own;
  for (int w = wave + 1; w < WAVES; ++w) sum += s_wave[w];
  uint32_t best = 0;
  for (int j = (int)std::min(j0 + KC_HPER - 1, top); j >= (int)j0 && j >= 1; --j) {
    sum += hist[j];
    if (sum >= (uint32_t)j) {
      best = (uint32_t)j;
      break;
    }
  }
  for (int off = 32; off > 0; off >>= 1) best = std::max(best, (uint32_t)__shfl_xor(best, off, 64));
  if (lane == 0 && best) atomicMax(&s_best, best);
  __syncthreads();
  best = s_best;
  if (best == KC_HHIST && cap > KC_HHIST) {
    // binary search on [KC_HHIST, cap] with block-wide counts
    uint32_t lo = KC_HHIST, hi = cap;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      uint32_t c = 0;
      for (uint64_t k = b + tid; k < e; k += KC_HTHREADS) c += est[idx[k]] >= mid ? 1u : 0u;
      for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
      __syncthreads();
      if (tid == 0) s_best = 0;
      __syncthreads();
      if (lane == 0 && c) atomicAdd(&s_best, c);
      __syncthreads();
      if (s_best >= mid) lo = mid;
      else hi = mid - 1;
    }
    best = lo;
  }
  __syncthreads();  // s_best and s_wave are free again
  return best;
}

// A block's vertex v with h-index `best`: a drop below its coreness is recorded and its
// neighbours are activated.  Returns 1 for a modified vertex.
__device__ uint32_t kc_block_finish(uint32_t v, uint32_t best, uint64_t b, uint64_t e,
                                    const uint32_t* __restrict__ idx, uint32_t* __restrict__ core,
                                    uint64_t* __restrict__ next_bits) {
  const bool mod = best && best < core[v];
  __syncthreads();  // every thread has read core[v]
  if (mod) {
    if (threadIdx.x == 0) core[v] = best;
    for (uint64_t k = b + threadIdx.x; k < e; k += KC_HTHREADS) kc_activate(next_bits, idx[k]);
  }
  __syncthreads();
  return mod ? 1u : 0u;
}

// One block per queued vertex with more than KC_HUGE neighbours (the hubs, whose
// adjacency would otherwise keep one wave busy for the whole epoch), up to KC_GIANT;
// longer ones are left to the chunked passes below.
__global__ __launch_bounds__(KC_HTHREADS) void k_kc_huge(
    const uint32_t* __restrict__ huge, const kc_counters* __restrict__ cnt_in, const uint64_t* __restrict__ ptr,
    const uint32_t* __restrict__ idx, const uint32_t* __restrict__ est, uint32_t* __restrict__ core,
    uint64_t* __restrict__ next_bits, kc_counters* __restrict__ cnt) {
  __shared__ uint32_t hist[KC_HTHREADS * KC_HPER];
  const int tid = threadIdx.x;
  const uint64_t nhuge = cnt_in->nhuge;
  uint32_t nmod = 0;
  for (uint64_t q = blockIdx.x; q < nhuge; q += gridDim.x) {
    const uint32_t v = huge[q];
    const uint64_t b = ptr[v], e = ptr[v + 1];
    if (e - b > KC_GIANT) continue;  // uniform: k_kc_giant_*
    const uint32_t cap = (uint32_t)std::min<uint64_t>(e - b, (uint64_t)est[v]);
    const uint32_t top = std::min(cap, KC_HHIST);
    for (uint32_t j = tid; j <= top; j += KC_HTHREADS) hist[j] = 0;
    __syncthreads();
    for (uint64_t k = b + tid; k < e; k += KC_HTHREADS) {
      const uint32_t eu = est[idx[k]];
      atomicAdd(&hist[eu < top ? eu : top], 1u);
    }
    __syncthreads();
    const uint32_t best = kc_block_hindex(hist, top, cap, b, e, idx, est);
    nmod += kc_block_finish(v, best, b, e, idx, core, next_bits);
  }
  if (tid == 0 && nmod) atomicAdd(&cnt->modified, (unsigned long long)nmod);
}

// Giants (more than KC_GIANT neighbours): their adjacencies in chunks of KC_GCHUNK, one
// block per chunk, each adding its LDS histogram of the neighbours' capped estimates
// into the giant's histogram in global memory (ghist, KC_HHIST + 1 bins per giant);
// then one block per giant finds the h-index from it.  Inactive giants are skipped.
__global__ void k_kc_giant_zero(uint32_t* __restrict__ ghist, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    ghist[i] = 0u;
}

__global__ __launch_bounds__(KC_HTHREADS) void k_kc_giant_hist(
    const uint64_t* __restrict__ ptr, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ giants,
    const unsigned long long* __restrict__ ngiant, const uint64_t* __restrict__ start,
    const uint64_t* __restrict__ active, const uint32_t* __restrict__ est, uint32_t* __restrict__ ghist) {
  __shared__ uint32_t hist[KC_HHIST + 1];
  const int tid = threadIdx.x;
  const uint64_t n = *ngiant;
  if (n == 0) return;
  const uint64_t total = start[n];
  for (uint64_t c = blockIdx.x; c < total; c += gridDim.x) {
    uint64_t k0, k1, gi;
    const uint32_t v = hub_chunk(ptr, giants, start, n, KC_GCHUNK, c, &k0, &k1, &gi);
    if (!((active[v >> 6] >> (v & 63)) & 1ull)) continue;  // uniform
    const uint64_t len = ptr[v + 1] - ptr[v];
    const uint32_t cap = (uint32_t)std::min<uint64_t>(len, (uint64_t)est[v]);
    const uint32_t top = std::min(cap, KC_HHIST);
    for (uint32_t j = tid; j <= top; j += KC_HTHREADS) hist[j] = 0;
    __syncthreads();
    for (uint64_t k = k0 + tid; k < k1; k += KC_HTHREADS) {
      const uint32_t eu = est[idx[k]];
      atomicAdd(&hist[eu < top ? eu : top], 1u);
    }
    __syncthreads();
    uint32_t* gh = ghist + gi * (uint64_t)(KC_HHIST + 1);
    for (uint32_t j = tid; j <= top; j += KC_HTHREADS)
      if (hist[j]) atomicAdd(&gh[j], hist[j]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(KC_HTHREADS) void k_kc_giant_final(
    const uint64_t* __restrict__ ptr, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ giants,
    const unsigned long long* __restrict__ ngiant, const uint64_t* __restrict__ active,
    const uint32_t* __restrict__ est, const uint32_t* __restrict__ ghist, uint32_t* __restrict__ core,
    uint64_t* __restrict__ next_bits, kc_counters* __restrict__ cnt) {
  __shared__ uint32_t hist[KC_HTHREADS * KC_HPER];
  const int tid = threadIdx.x;
  const uint64_t n = *ngiant;
  uint32_t nmod = 0;
  for (uint64_t gi = blockIdx.x; gi < n; gi += gridDim.x) {
    const uint32_t v = giants[gi];
    if (!((active[v >> 6] >> (v & 63)) & 1ull)) continue;  // uniform
    const uint64_t b = ptr[v], e = ptr[v + 1];
    const uint32_t cap = (uint32_t)std::min<uint64_t>(e - b, (uint64_t)est[v]);
    const uint32_t top = std::min(cap, KC_HHIST);
    const uint32_t* gh = ghist + gi * (uint64_t)(KC_HHIST + 1);
    for (uint32_t j = tid; j <= top; j += KC_HTHREADS) hist[j] = gh[j];
    __syncthreads();
    const uint32_t best = kc_block_hindex(hist, top, cap, b, e, idx, est);
    nmod += kc_block_finish(v, best, b, e, idx, core, next_bits);
  }
  if (tid == 0 && nmod) atomicAdd(&cnt->modified, (unsigned long long)nmod);
}

// estimate = coreness for the epoch's active vertices (kcore.hpp:248-258)
__global__ void k_kc_sync(const uint32_t* __restrict__ queue, uint64_t n, const uint32_t* __restrict__ core,
                          uint32_t* __restrict__ est) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = queue[i];
    est[v] = core[v];
  }
}

}  // namespace pe

using namespace pe;

extern "C" int pe_kcore(pe_graph_t* g, int out_edges, uint32_t* coreness, uint64_t* epoch_modified,
                        uint64_t* epoch_next, uint32_t cap, pe_kcore_stats_t* stats) {
  if (!g || !coreness) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(!out_edges || g->has_csr, PE_ERR_STATE, "out-edge k-core needs PE_GRAPH_BUILD_CSR");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t V = g->V;
    const uint64_t* ptr = out_edges ? g->csr_ptr() : g->col_ptr.p;
    const uint32_t* idx = out_edges ? g->csr_idx() : g->row_idx.p;
    const uint64_t nwords = ((uint64_t)V >> 6) + 1;
    const uint32_t Vs = std::max<uint32_t>(1, V);
    dbuf<uint32_t> core(ctx, Vs), est(ctx, Vs), queue(ctx, Vs), big(ctx, Vs), huge(ctx, Vs);
    dbuf<uint64_t> cur(ctx, nwords), nxt(ctx, nwords), big_bits(ctx, nwords), huge_bits(ctx, nwords);
    dbuf<kc_counters> cnt(ctx, 1);
    kc_counters* h = static_cast<kc_counters*>(ctx->host_pinned);
    PE_HIP(hipEventRecord(ctx->t0, s));
    k_kc_init<<<grid_for(V, 256, 8192), 256, 0, s>>>(ptr, V, core.p, est.p);
    k_kc_fill<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(cur.p, nwords, V);  // active_current.fill()
    PE_HIP(hipGetLastError());
    const unsigned big_grid = persistent_grid(ctx, 1u << 20, 4);
    const unsigned small_grid = persistent_grid(ctx, 1u << 20, 8);
    const unsigned huge_grid = persistent_grid(ctx, 1u << 20, 1);
    // giants and their histograms (one per giant, KC_HHIST + 1 bins)
    hub_table giants;
    hub_table_build(ctx, ptr, V, g->nnz, KC_GIANT, KC_GCHUNK, giants);
    unsigned long long ngiant = 0;
    PE_HIP(hipMemcpyAsync(&ngiant, giants.count.p, sizeof(ngiant), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    const uint64_t ghist_words = ngiant * (uint64_t)(KC_HHIST + 1);
    dbuf<uint32_t> ghist(ctx, std::max<uint64_t>(1, ghist_words));
    uint32_t epoch = 0;
    uint64_t total_mod = 0;
    for (;;) {
      k_kc_reset<<<1, 1, 0, s>>>(cnt.p);
      frontier_compact(ctx, cur.p, nwords, V, queue.p, &cnt.p->queued);
      PE_HIP(hipMemcpyAsync(h, cnt.p, sizeof(kc_counters), hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      const uint64_t n = h->queued;
      if (epoch > 0) {
        if (epoch - 1 < cap && epoch_next) epoch_next[epoch - 1] = n;  // need_modified of the last epoch
        if (n == 0) break;                                              // kcore.hpp:264
      }
      k_kc_zero<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(nxt.p, nwords);
      if (n) {
        k_kc_zero<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(big_bits.p, nwords);
        k_kc_zero<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(huge_bits.p, nwords);
        k_kc_small<<<std::min<unsigned>(small_grid, grid_for(n, KC_THREADS)), KC_THREADS, 0, s>>>(
            queue.p, n, ptr, idx, est.p, core.p, nxt.p, big_bits.p, huge_bits.p, cnt.p);
        frontier_compact(ctx, big_bits.p, nwords, V, big.p, &cnt.p->nbig);
        frontier_compact(ctx, huge_bits.p, nwords, V, huge.p, &cnt.p->nhuge);
        k_kc_big<<<big_grid, KC_THREADS, 0, s>>>(big.p, cnt.p, ptr, idx, est.p, core.p, nxt.p, cnt.p);
        k_kc_huge<<<huge_grid, KC_HTHREADS, 0, s>>>(huge.p, cnt.p, ptr, idx, est.p, core.p, nxt.p, cnt.p);
        if (ngiant) {
          k_kc_giant_zero<<<grid_for(ghist_words, 256, 4096), 256, 0, s>>>(ghist.p, ghist_words);
          k_kc_giant_hist<<<huge_grid, KC_HTHREADS, 0, s>>>(ptr, idx, giants.hubs.p, giants.count.p, giants.start.p,
                                                            cur.p, est.p, ghist.p);
          k_kc_giant_final<<<(unsigned)std::min<uint64_t>(ngiant, huge_grid), KC_HTHREADS, 0, s>>>(
              ptr, idx, giants.hubs.p, giants.count.p, cur.p, est.p, ghist.p, core.p, nxt.p, cnt.p);
        }
        k_kc_sync<<<grid_for(n, 256, 8192), 256, 0, s>>>(queue.p, n, core.p, est.p);
      }
      PE_HIP(hipGetLastError());
      PE_HIP(hipMemcpyAsync(h, cnt.p, sizeof(kc_counters), hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      if (epoch < cap && epoch_modified) epoch_modified[epoch] = h->modified;
      total_mod += h->modified;
      std::swap(cur.p, nxt.p);
      ++epoch;
      // every epoch but the last lowers some coreness, which starts at most at its length
      PE_REQUIRE((uint64_t)epoch <= g->nnz + 2, PE_ERR_INTERNAL, "k-core did not terminate");
      if (n == 0) {  // nothing active in epoch 0: an empty graph
        if (epoch - 1 < cap && epoch_next) epoch_next[epoch - 1] = 0;
        break;
      }
    }
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    download_external(g, core.p, coreness, 4);
    if (stats) {
      stats->epochs = epoch;
      stats->modified = total_mod;
      stats->seconds = ms * 1e-3;
    }
  });
}
