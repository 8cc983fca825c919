// Connected components: plato::algo::connected_component_t::compute
// (plato/algo/cgm/connected_component.hpp:170-234) on the device engine.
//
// Labels start at the vertex ids and only ever take a smaller label: pushed along the
// out-edges of the vertices whose label changed in the last epoch (the push slot's
// write_min and set_bit, :193-205), or pulled by every vertex as the minimum of its
// own id and its in-neighbours' labels (:206-224).  The engine's rule picks push or
// pull per epoch from the active set's out-degree sum (dualmode.hpp:264-279), and the
// loop ends after an epoch that lowers no label.  The fixed point is unique: label(v)
// is the smallest input id from which v is reachable (on an undirected graph, the
// smallest id of its component), whatever the order in which labels spread; only the
// per-epoch activity depends on it, as in the reference, whose threads race the same way.
//
// Labels are kept in input ids, so a degree-relabeled graph gives the same labels.
#include <algorithm>
#include <vector>

#include "pe_internal.hpp"

namespace pe {

static constexpr uint32_t CC_PULL_CAP = 64;  // in-edges a pull lane takes; longer go one wave each
static constexpr uint64_t CC_HUB_DEG = 4096;    // longer still: chunks of CC_HUB_CHUNK over all waves
static constexpr uint64_t CC_HUB_CHUNK = 2048;

struct cc_counters {
  unsigned long long activated;  // sum of the slots' returns
  unsigned long long queued;     // push frontier size
  unsigned long long heavy;      // pull vertices with more than CC_PULL_CAP in-edges
};

__global__ void k_cc_init(uint32_t V, const uint32_t* __restrict__ perm, uint32_t* __restrict__ label) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
    label[v] = perm ? perm[v] : (uint32_t)v;
}

__global__ void k_cc_fill(uint64_t* __restrict__ w, uint64_t nwords, uint32_t V) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nwords; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = i * 64;
    w[i] = lo + 64 <= V ? ~0ull : (lo < V ? (1ull << (V - lo)) - 1ull : 0ull);
  }
}

__global__ void k_cc_zero(uint64_t* __restrict__ w, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    w[i] = 0ull;
}

__global__ void k_cc_reset(cc_counters* c) {
  c->activated = 0;
  c->queued = 0;
  c->heavy = 0;
}

// one atomic per block; every thread of the block calls it
__device__ __forceinline__ void cc_add(unsigned long long act, unsigned long long* dst) {
  block_add2(act, 0, dst, dst);
}

// Push: one wave per frontier vertex u over its out-edges; a neighbour whose label is
// larger takes u's label (write_min) and becomes active.
__global__ __launch_bounds__(256) void k_cc_push(const uint32_t* __restrict__ queue, const cc_counters* cnt_in,
                                                 const uint64_t* __restrict__ ptr, const uint32_t* __restrict__ idx,
                                                 uint32_t* __restrict__ label, uint64_t* __restrict__ next_bits,
                                                 cc_counters* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const uint64_t n = cnt_in->queued;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  unsigned long long act = 0;
  for (uint64_t q = wave_id; q < n; q += nwaves) {
    const uint32_t u = queue[q];
    const uint32_t lu = label[u];  // the message's label, read when it is sent
    const uint64_t b = ptr[u], e = ptr[u + 1];
    for (uint64_t k = b + lane; k < e; k += 64) {
      const uint32_t dst = idx[k];
      if (lu < label[dst]) {
        atomicMin(&label[dst], lu);
        const uint64_t m = 1ull << (dst & 63);
        if (!(next_bits[dst >> 6] & m)) atomicOr((unsigned long long*)&next_bits[dst >> 6], (unsigned long long)m);
        ++act;
      }
    }
  }
  cc_add(act, &cnt->activated);
}

// Pull: one lane per vertex, 64 consecutive vertices per wave (the wave owns one word of
// the next bitmap); vertices with more than CC_PULL_CAP in-edges are marked in a second
// bitmap, queued by compaction and finished by k_cc_pull_heavy.
// Only v's own lane or wave writes label[v] in a pull epoch.
__global__ __launch_bounds__(256) void k_cc_pull(const uint64_t* __restrict__ col_ptr,
                                                 const uint32_t* __restrict__ row_idx, uint32_t V, uint64_t nwords,
                                                 const uint32_t* __restrict__ perm, uint32_t* __restrict__ label,
                                                 uint64_t* __restrict__ next_bits, uint64_t* __restrict__ heavy_bits,
                                                 cc_counters* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  unsigned long long act = 0;
  for (uint64_t word = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; word < nwords;
       word += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t v = word * 64 + lane;
    bool upd = false, is_heavy = false;
    if (v < V) {
      const uint64_t b = col_ptr[v], e = col_ptr[v + 1];
      if (e - b > CC_PULL_CAP) {
        is_heavy = true;
      } else {
        uint32_t cur = perm ? perm[v] : (uint32_t)v;  // the mirror's own id (:209)
        for (uint64_t k = b; k < e; ++k) cur = std::min(cur, label[row_idx[k]]);
        if (cur < label[v]) {
          label[v] = cur;
          upd = true;
        }
      }
    }
    // bitmap words, one store per wave: no per-word atomics on a shared counter
    const uint64_t m = __ballot(upd), hm = __ballot(is_heavy);
    if (lane == 0) {
      next_bits[word] = m;
      heavy_bits[word] = hm;
    }
    act += upd ? 1 : 0;
  }
  cc_add(act, &cnt->activated);
}

__global__ __launch_bounds__(256) void k_cc_pull_heavy(const uint64_t* __restrict__ col_ptr,
                                                       const uint32_t* __restrict__ row_idx,
                                                       const uint32_t* __restrict__ perm,
                                                       uint32_t* __restrict__ label, uint64_t* __restrict__ next_bits,
                                                       const uint32_t* __restrict__ heavy, cc_counters* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const uint64_t n = cnt->heavy;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  unsigned long long act = 0;
  for (uint64_t h = wave_id; h < n; h += nwaves) {
    const uint32_t v = heavy[h];
    const uint64_t b = col_ptr[v], e = col_ptr[v + 1];
    if (e - b > CC_HUB_DEG) continue;  // k_cc_pull_hubs
    uint32_t cur = perm ? perm[v] : v;
    for (uint64_t k = b + lane; k < e; k += 64) cur = std::min(cur, label[row_idx[k]]);
    for (int off = 32; off > 0; off >>= 1) cur = std::min(cur, (uint32_t)__shfl_xor(cur, off, 64));
    if (lane == 0 && cur < label[v]) {
      label[v] = cur;
      atomicOr((unsigned long long*)&next_bits[v >> 6], 1ull << (v & 63));
      ++act;
    }
  }
  if (lane == 0 && act) atomicAdd(&cnt->activated, act);
}

// Pull for the hubs (more than CC_HUB_DEG in-edges), one chunk of in-edges per wave at
// a time: the chunk's minimum label lowers the hub's with atomicMin (several chunks of one
// hub may), and the hub counts once, when its bit in the next bitmap is first set.  A
// hub's own id is never below its label, so the chunks leave it out.
__global__ __launch_bounds__(256) void k_cc_pull_hubs(const uint64_t* __restrict__ col_ptr,
                                                      const uint32_t* __restrict__ row_idx,
                                                      const uint32_t* __restrict__ hubs,
                                                      const unsigned long long* __restrict__ nhub,
                                                      const uint64_t* __restrict__ start, uint32_t* __restrict__ label,
                                                      uint64_t* __restrict__ next_bits, cc_counters* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  const uint64_t n = *nhub;
  unsigned long long act = 0;
  if (n) {
    const uint64_t total = start[n];
    for (uint64_t c = wave_id; c < total; c += nwaves) {
      uint64_t k0, k1;
      const uint32_t v = hub_chunk(col_ptr, hubs, start, n, CC_HUB_CHUNK, c, &k0, &k1);
      uint32_t cur = 0xFFFFFFFFu;
      for (uint64_t k = k0 + lane; k < k1; k += 64) cur = std::min(cur, label[row_idx[k]]);
      for (int off = 32; off > 0; off >>= 1) cur = std::min(cur, (uint32_t)__shfl_xor(cur, off, 64));
      if (lane == 0 && cur < label[v] && cur < atomicMin(&label[v], cur)) {
        const unsigned long long bit = 1ull << (v & 63);
        if (!(atomicOr((unsigned long long*)&next_bits[v >> 6], bit) & bit)) ++act;
      }
    }
  }
  cc_add(act, &cnt->activated);
}

}  // namespace pe

using namespace pe;

extern "C" int pe_connected_components(pe_graph_t* g, uint32_t* labels, uint8_t* modes, uint64_t* actives,
                                       uint32_t cap, pe_cc_stats_t* stats) {
  if (!g || !labels) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t V = g->V;
    const uint64_t nwords = ((uint64_t)V >> 6) + 1;
    const uint32_t Vs = std::max<uint32_t>(1, V);
    dbuf<uint32_t> label(ctx, Vs), queue(ctx, Vs), heavy(ctx, Vs);
    dbuf<uint64_t> cur(ctx, nwords), nxt(ctx, nwords), hbits(ctx, nwords);
    dbuf<cc_counters> cnt(ctx, 1);
    cc_counters* h = static_cast<cc_counters*>(ctx->host_pinned);
    const uint32_t* perm = g->relabeled ? g->perm.p : nullptr;
    const uint64_t Ep = g->directed ? g->edges : g->edges * 2;  // dualmode.hpp:261-262
    PE_HIP(hipEventRecord(ctx->t0, s));
    k_cc_init<<<grid_for(V, 256, 8192), 256, 0, s>>>(V, perm, label.p);
    k_cc_fill<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(cur.p, nwords, V);  // active_current.fill()
    PE_HIP(hipGetLastError());
    const unsigned push_grid = persistent_grid(ctx, 1u << 20, 8);
    const unsigned pull_grid = persistent_grid(ctx, grid_for(nwords * 64, 256), 8);
    hub_table hubs;
    hub_table_build(ctx, g->col_ptr.p, V, g->nnz, CC_HUB_DEG, CC_HUB_CHUNK, hubs);
    uint64_t A = active_edges_dev(g, cur.p, nwords);
    uint32_t epoch = 0;
    for (uint64_t act = 1; act != 0; ++epoch) {
      const bool push = g->has_csr && ((double)A / (double)std::max<uint64_t>(Ep, 1)) < 0.05;  // dualmode.hpp:279
      k_cc_reset<<<1, 1, 0, s>>>(cnt.p);
      k_cc_zero<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(nxt.p, nwords);
      if (push) {
        frontier_compact(ctx, cur.p, nwords, V, queue.p, &cnt.p->queued);
        k_cc_push<<<push_grid, 256, 0, s>>>(queue.p, cnt.p, g->csr_ptr(), g->csr_idx(), label.p, nxt.p, cnt.p);
      } else {
        k_cc_pull<<<pull_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, V, nwords, perm, label.p, nxt.p, hbits.p,
                                            cnt.p);
        frontier_compact(ctx, hbits.p, nwords, V, heavy.p, &cnt.p->heavy);
        k_cc_pull_heavy<<<push_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, perm, label.p, nxt.p, heavy.p, cnt.p);
        k_cc_pull_hubs<<<push_grid, 256, 0, s>>>(g->col_ptr.p, g->row_idx.p, hubs.hubs.p, hubs.count.p, hubs.start.p,
                                                 label.p, nxt.p, cnt.p);
      }
      PE_HIP(hipGetLastError());
      PE_HIP(hipMemcpyAsync(h, cnt.p, sizeof(cc_counters), hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      act = h->activated;
      if (epoch < cap) {
        if (modes) modes[epoch] = push ? 1 : 0;
        if (actives) actives[epoch] = act;
      }
      std::swap(cur.p, nxt.p);
      if (act) A = active_edges_dev(g, cur.p, nwords);
      // every epoch but the last lowers a label, and labels are ids below V
      PE_REQUIRE((uint64_t)epoch <= (uint64_t)V * V + 1, PE_ERR_INTERNAL, "connected components did not terminate");
    }
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    download_external(g, label.p, labels, 4);
    if (stats) {
      stats->epochs = epoch;
      stats->seconds = ms * 1e-3;
      std::vector<uint8_t> seen(V, 0);
      uint32_t c = 0;
      for (uint32_t v = 0; v < V; ++v)
        if (!seen[labels[v]]) {
          seen[labels[v]] = 1;
          ++c;
        }
      stats->components = c;
    }
  });
}
