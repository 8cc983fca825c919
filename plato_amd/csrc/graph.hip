// Device graph construction: counter-based RMAT generator, CSC/CSR build by
// radix sort of packed (major, minor) keys, out-degrees, normalised weights and
// the merge-path tiling used by the pull gather.
//
// Layout (replaces dcsc_t, plato/graph/structure/dcsc.hpp:58-185, and bcsr_t,
// plato/graph/structure/bcsr.hpp:67-208):
//   col_ptr uint64[V+1]  over every id (no double compression; empty columns
//                        cost 8 bytes, and a column is found without a search)
//   row_idx uint32[nnz]  in-neighbours, ascending inside each column
//   od      uint32[V]    out-degree counted from the in-edges (structure.hpp:292-338)
//   wn      float[nnz]   weight / (sum of the source's out-weights)
//   row_ptr/col_idx      out-edges for push mode (shared with the CSC when symmetrized)
#include <algorithm>
#include <cstring>
#include <memory>

#include "pe_internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

namespace pe {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  // SplitMix64 finalizer (Steele, Lea, Flood: "Fast splittable pseudorandom
  // number generators", OOPSLA 2014).
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct rmat_params {
  uint32_t scale;
  uint64_t sk;      // mix64(seed)
  uint32_t t0, t1, t2;
};

// Edge e of the RMAT graph: bit (scale-1-l) of (src, dst) is decided by a
// 32-bit draw; draws come two per 64-bit hash of (seed, e, level pair).
__device__ __forceinline__ void rmat_edge(const rmat_params& p, uint64_t e, uint32_t& s,
                                          uint32_t& d) {
  s = 0;
  d = 0;
  for (uint32_t l = 0; l < p.scale; l += 2) {
    const uint64_t word = mix64(p.sk ^ (e * 32u + (l >> 1)));
    const uint32_t u0 = (uint32_t)word;
    uint32_t bit = 1u << (p.scale - 1u - l);
    if (u0 >= p.t0) {
      if (u0 < p.t1) d |= bit;
      else if (u0 < p.t2) s |= bit;
      else { s |= bit; d |= bit; }
    }
    if (l + 1 < p.scale) {
      const uint32_t u1 = (uint32_t)(word >> 32);
      bit >>= 1;
      if (u1 >= p.t0) {
        if (u1 < p.t1) d |= bit;
        else if (u1 < p.t2) s |= bit;
        else { s |= bit; d |= bit; }
      }
    }
  }
}

__device__ __forceinline__ float rmat_weight(uint64_t sk, uint64_t e) {
  const uint64_t r = mix64(sk ^ (e * 32u + 31u));
  return (float)((r >> 40) + 1u) * (1.0f / 16777216.0f);
}

__global__ void k_rmat_edges(rmat_params p, uint64_t e0, uint64_t n, uint32_t* __restrict__ src,
                             uint32_t* __restrict__ dst, float* __restrict__ w) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t s, d;
    rmat_edge(p, e0 + i, s, d);
    src[i] = s;
    dst[i] = d;
    if (w) w[i] = rmat_weight(p.sk, e0 + i);
  }
}

__global__ void k_edge_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                            const float* __restrict__ w_in, uint64_t E, int b, int csr, int sym,
                            uint64_t* __restrict__ keys, float* __restrict__ w) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = src[i], d = dst[i];
    const uint64_t maj = csr ? s : d, mnr = csr ? d : s;
    keys[i] = (maj << b) | mnr;
    if (sym) keys[E + i] = (mnr << b) | maj;
    if (w) {
      w[i] = w_in[i];
      if (sym) w[E + i] = w_in[i];
    }
  }
}

// Split sorted keys into ptr[V+1] and idx[nnz].
__global__ void k_split_keys(const uint64_t* __restrict__ keys, uint64_t nnz, int b, uint32_t V,
                             uint64_t* __restrict__ ptr, uint32_t* __restrict__ idx) {
  const uint64_t mask = (1ull << b) - 1ull;
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[e];
    const uint32_t maj = (uint32_t)(k >> b);
    idx[e] = (uint32_t)(k & mask);
    const int64_t prev = e ? (int64_t)(keys[e - 1] >> b) : -1;
    for (int64_t v = prev + 1; v <= (int64_t)maj; ++v) ptr[v] = e;
  }
}

// ptr[v] = nnz for every v above the last major key (grid-stride, not one thread)
__global__ void k_split_tail(const uint64_t* __restrict__ keys, uint64_t nnz, int b, uint32_t V,
                             uint64_t* __restrict__ ptr) {
  const uint64_t last = keys[nnz - 1] >> b;
  for (uint64_t v = last + 1 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= V;
       v += (uint64_t)gridDim.x * blockDim.x)
    ptr[v] = nnz;
}

static void split_keys(hipStream_t s, const uint64_t* keys, uint64_t nnz, int b, uint32_t V,
                       uint64_t* ptr, uint32_t* idx) {
  if (nnz) {
    k_split_keys<<<grid_for(nnz, 256, 1u << 20), 256, 0, s>>>(keys, nnz, b, V, ptr, idx);
    k_split_tail<<<grid_for(V, 256, 8192), 256, 0, s>>>(keys, nnz, b, V, ptr);
  } else {
    PE_HIP(hipMemsetAsync(ptr, 0, ((uint64_t)V + 1) * 8, s));
  }
  PE_HIP(hipGetLastError());
}

__global__ void k_out_degrees(const uint32_t* __restrict__ row_idx, uint64_t nnz,
                              uint32_t* __restrict__ od) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&od[row_idx[e]], 1u);
}

__global__ void k_row_lengths(const uint64_t* __restrict__ ptr, uint32_t V, uint32_t* __restrict__ len) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    len[v] = (uint32_t)(ptr[v + 1] - ptr[v]);
}

// ids with an edge: count, max id, max in-degree
__global__ void k_vertex_stats(const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ od,
                               uint32_t V, const uint32_t* __restrict__ perm, unsigned int* __restrict__ out) {
  unsigned int cnt = 0, mx = 0, mdeg = 0;
  bool any = false;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t indeg = (uint32_t)(col_ptr[v + 1] - col_ptr[v]);
    if (indeg > mdeg) mdeg = indeg;
    if (indeg || od[v]) {
      ++cnt;
      mx = max(mx, perm ? perm[v] : (unsigned)v);
      any = true;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_down(cnt, off, 64);
    mx = max(mx, (unsigned)__shfl_down(mx, off, 64));
    mdeg = max(mdeg, (unsigned)__shfl_down(mdeg, off, 64));
    any = any || __shfl_down((int)any, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (cnt) atomicAdd(&out[0], cnt);
    if (any) atomicMax(&out[1], mx);
    atomicMax(&out[2], mdeg);
  }
}

// Sum of out-weights per source, in CSR order (deterministic), FP64.
__global__ void k_row_weight_sum(const uint64_t* __restrict__ ptr, const float* __restrict__ w,
                                 uint32_t V, double* __restrict__ W) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (uint64_t e = ptr[v]; e < ptr[v + 1]; ++e) s += (double)w[e];
    W[v] = s;
  }
}

// beta = w / W(u) (weighted_pagerank.hpp:119), rounded once to FP32
__global__ void k_normalise_weights(const uint32_t* __restrict__ row_idx, const double* __restrict__ W,
                                    uint64_t nnz, float* __restrict__ w) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x)
    w[e] = (float)(((double)w[e] * 1.0) / W[row_idx[e]]);
}

// Merge-path split of the (row-end, edge) sequence at each tile start
// (Merrill & Garland, "Merge-based parallel sparse matrix-vector
// multiplication", SC16): x rows and D-x edges precede diagonal D.
__global__ void k_tile_starts(const uint64_t* __restrict__ col_ptr, uint32_t V, uint64_t nnz,
                              uint32_t tile_items, uint64_t T, uint32_t* __restrict__ trow,
                              uint64_t* __restrict__ tedge) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t <= T;
       t += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t D = t * (uint64_t)tile_items;
    const uint64_t total = (uint64_t)V + nnz;
    if (D > total) D = total;
    uint64_t lo = D > nnz ? D - nnz : 0, hi = D < V ? D : V;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      // row `mid` ends (its end item precedes edge D-mid-1) iff col_ptr[mid+1] <= D-mid-1
      if (col_ptr[mid + 1] + mid + 1 <= D) lo = mid + 1;
      else hi = mid;
    }
    trow[t] = (uint32_t)lo;
    tedge[t] = D - lo;
  }
}

// A row needs the fix-up when its end lies in tile t but its first edge lies in
// an earlier tile.  Writes t0 (tile of the first edge) or UINT32_MAX.
__global__ void k_tile_fix_flags(const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ trow,
                                 const uint64_t* __restrict__ tedge, uint64_t T,
                                 uint32_t* __restrict__ t0_out) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < T;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r0 = trow[t];
    uint32_t res = 0xFFFFFFFFu;
    if (trow[t + 1] > r0 && tedge[t] > col_ptr[r0]) {
      // last tile whose first edge is <= col_ptr[r0]
      const uint64_t first = col_ptr[r0];
      uint64_t lo = 0, hi = t;  // answer in [0, t)
      while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (tedge[mid] <= first) lo = mid;
        else hi = mid;
      }
      res = (uint32_t)lo;
    }
    t0_out[t] = res;
  }
}

static rmat_params make_rmat(uint32_t scale, uint64_t seed, double a, double b, double c) {
  rmat_params p;
  p.scale = scale;
  // host twin of mix64
  uint64_t z = seed + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  p.sk = z ^ (z >> 31);
  const double pr[3] = {a, a + b, a + b + c};
  uint32_t t[3];
  for (int i = 0; i < 3; ++i) {
    double x = std::floor(pr[i] * 4294967296.0);
    x = std::min(std::max(x, 0.0), 4294967295.0);
    t[i] = (uint32_t)x;
  }
  p.t0 = t[0];
  p.t1 = t[1];
  p.t2 = t[2];
  return p;
}

static int bits_for(uint32_t V) {
  int b = 1;
  while (b < 32 && (1ull << b) < (uint64_t)V) ++b;
  return b;
}

// Radix sort of packed keys (and optional float payload), stable.
void sort_keys(pe_ctx_s* ctx, dbuf<uint64_t>& keys, dbuf<float>* vals, uint64_t n,
                      int end_bit) {
  if (n == 0) return;
  hipStream_t s = ctx->stream;
  dbuf<uint64_t> alt(ctx, n);
  rocprim::double_buffer<uint64_t> kb(keys.p, alt.p);
  size_t temp = 0;
  if (vals) {
    dbuf<float> valt(ctx, n);
    rocprim::double_buffer<float> vb(vals->p, valt.p);
    PE_HIP(rocprim::radix_sort_pairs(nullptr, temp, kb, vb, n, 0u, (unsigned)end_bit, s));
    dbuf<uint8_t> tmp(ctx, temp ? temp : 1);
    PE_HIP(rocprim::radix_sort_pairs(tmp.p, temp, kb, vb, n, 0u, (unsigned)end_bit, s));
    PE_HIP(hipStreamSynchronize(s));
    if (kb.current() != keys.p) std::swap(keys, alt);
    if (vb.current() != vals->p) std::swap(*vals, valt);
  } else {
    PE_HIP(rocprim::radix_sort_keys(nullptr, temp, kb, n, 0u, (unsigned)end_bit, s));
    dbuf<uint8_t> tmp(ctx, temp ? temp : 1);
    PE_HIP(rocprim::radix_sort_keys(tmp.p, temp, kb, n, 0u, (unsigned)end_bit, s));
    PE_HIP(hipStreamSynchronize(s));
    if (kb.current() != keys.p) std::swap(keys, alt);
  }
}

void graph_build_tiles(pe_graph_s* g, uint32_t tile_items) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  PE_REQUIRE(tile_items == 1024 || tile_items == 2048 || tile_items == 4096, PE_ERR_INVALID,
             "tile_items must be 1024, 2048 or 4096");
  const uint64_t total = (uint64_t)g->V + g->nnz;
  const uint64_t T = std::max<uint64_t>(1, (total + tile_items - 1) / tile_items);
  PE_REQUIRE(T < 0xFFFFFFFFull, PE_ERR_RANGE, "too many tiles");
  g->tile_items = tile_items;
  g->num_tiles = T;
  g->tile_row.alloc(ctx, T + 1);
  g->tile_edge.alloc(ctx, T + 1);
  k_tile_starts<<<grid_for(T + 1, 256, 65536), 256, 0, s>>>(g->col_ptr.p, g->V, g->nnz, tile_items, T,
                                                            g->tile_row.p, g->tile_edge.p);
  PE_HIP(hipGetLastError());
  dbuf<uint32_t> t0(ctx, T);
  k_tile_fix_flags<<<grid_for(T, 256, 65536), 256, 0, s>>>(g->col_ptr.p, g->tile_row.p, g->tile_edge.p,
                                                           T, t0.p);
  PE_HIP(hipGetLastError());
  std::vector<uint32_t> h_t0(T), h_row(T + 1);
  PE_HIP(hipMemcpyAsync(h_t0.data(), t0.p, T * 4, hipMemcpyDeviceToHost, s));
  PE_HIP(hipMemcpyAsync(h_row.data(), g->tile_row.p, (T + 1) * 4, hipMemcpyDeviceToHost, s));
  PE_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> fr, f0, f1, lr, l0, l1;
  for (uint64_t t = 0; t < T; ++t) {
    if (h_t0[t] == 0xFFFFFFFFu) continue;
    const bool longspan = (t - h_t0[t]) > 8;
    (longspan ? lr : fr).push_back(h_row[t]);
    (longspan ? l0 : f0).push_back(h_t0[t]);
    (longspan ? l1 : f1).push_back((uint32_t)t);
  }
  g->num_fix_long = lr.size();
  fr.insert(fr.end(), lr.begin(), lr.end());
  f0.insert(f0.end(), l0.begin(), l0.end());
  f1.insert(f1.end(), l1.begin(), l1.end());
  g->num_fix = fr.size();
  g->fix_row.alloc(ctx, std::max<size_t>(1, fr.size()));
  g->fix_t0.alloc(ctx, std::max<size_t>(1, fr.size()));
  g->fix_t1.alloc(ctx, std::max<size_t>(1, fr.size()));
  if (!fr.empty()) {
    PE_HIP(hipMemcpyAsync(g->fix_row.p, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(g->fix_t0.p, f0.data(), f0.size() * 4, hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(g->fix_t1.p, f1.data(), f1.size() * 4, hipMemcpyHostToDevice, s));
  }
  PE_HIP(hipStreamSynchronize(s));
}

// Shared tail of both builders: keys already hold the CSC order.
static void finish_build(pe_graph_s* g, dbuf<uint64_t>& keys, dbuf<float>* w, int b) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t nnz = g->nnz;
  sort_keys(ctx, keys, w, nnz, 2 * b);
  g->col_ptr.alloc(ctx, (uint64_t)g->V + 1);
  g->row_idx.alloc(ctx, std::max<uint64_t>(1, nnz));
  split_keys(s, keys.p, nnz, b, g->V, g->col_ptr.p, g->row_idx.p);
  if (w) g->wn = std::move(*w);
  keys.release();
  g->od.alloc(ctx, std::max<uint32_t>(1, g->V));
  PE_HIP(hipMemsetAsync(g->od.p, 0, (uint64_t)std::max<uint32_t>(1, g->V) * 4, s));
  if (nnz) {
    k_out_degrees<<<grid_for(nnz, 256, 1u << 20), 256, 0, s>>>(g->row_idx.p, nnz, g->od.p);
    PE_HIP(hipGetLastError());
  }
  dbuf<unsigned int> st(ctx, 3);
  PE_HIP(hipMemsetAsync(st.p, 0, 12, s));
  k_vertex_stats<<<grid_for(g->V, 256, 4096), 256, 0, s>>>(g->col_ptr.p, g->od.p, g->V,
                                                          g->relabeled ? g->perm.p : nullptr, st.p);
  PE_HIP(hipGetLastError());
  unsigned int hs[3];
  PE_HIP(hipMemcpyAsync(hs, st.p, 12, hipMemcpyDeviceToHost, s));
  PE_HIP(hipStreamSynchronize(s));
  g->vertices = hs[0];
  g->max_v_i = hs[1];
  g->max_in_degree = hs[2];
  graph_build_tiles(g, 2048);
}

// CSR (out-edges) from freshly generated CSR-ordered keys.
static void build_csr_from_keys(pe_graph_s* g, dbuf<uint64_t>& keys, int b) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  sort_keys(ctx, keys, nullptr, g->nnz, 2 * b);
  g->row_ptr.alloc(ctx, (uint64_t)g->V + 1);
  g->col_idx.alloc(ctx, std::max<uint64_t>(1, g->nnz));
  split_keys(s, keys.p, g->nnz, b, g->V, g->row_ptr.p, g->col_idx.p);
  keys.release();
  PE_HIP(hipStreamSynchronize(s));
}

// W(u) in CSR order, then wn = w / W(u) in CSC order.
static void normalise_weights(pe_graph_s* g, dbuf<uint64_t>& csr_keys, dbuf<float>& csr_w, int b) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  sort_keys(ctx, csr_keys, &csr_w, g->nnz, 2 * b);
  dbuf<uint64_t> rp(ctx, (uint64_t)g->V + 1);
  dbuf<uint32_t> ci(ctx, std::max<uint64_t>(1, g->nnz));
  split_keys(s, csr_keys.p, g->nnz, b, g->V, rp.p, ci.p);
  csr_keys.release();
  dbuf<double> W(ctx, std::max<uint32_t>(1, g->V));
  k_row_weight_sum<<<grid_for(g->V, 256, 1u << 16), 256, 0, s>>>(rp.p, csr_w.p, g->V, W.p);
  if (g->nnz)
    k_normalise_weights<<<grid_for(g->nnz, 256, 1u << 20), 256, 0, s>>>(g->row_idx.p, W.p, g->nnz, g->wn.p);
  PE_HIP(hipGetLastError());
  PE_HIP(hipStreamSynchronize(s));
}

// ---- degree relabeling -------------------------------------------------------------
// New id = rank of the vertex by (out-degree desc, id asc), out-degree counted over
// the CSC entries (reversals included).  Sources with many out-edges get small,
// adjacent ids, so the hot part of the gathered vector packs into few cache lines.
__global__ void k_entry_out_degrees(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                    uint64_t E, int sym, uint32_t* __restrict__ od) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E;
       i += (uint64_t)gridDim.x * blockDim.x) {
    atomicAdd(&od[src[i]], 1u);
    if (sym) atomicAdd(&od[dst[i]], 1u);
  }
}

__global__ void k_degree_keys(const uint32_t* __restrict__ od, uint32_t V, uint64_t* __restrict__ keys) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    keys[v] = ((uint64_t)(0xFFFFFFFFu - od[v]) << 32) | v;
}

__global__ void k_perm_from_keys(const uint64_t* __restrict__ keys, uint32_t V, uint32_t* __restrict__ perm,
                                 uint32_t* __restrict__ rank_of) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t old = (uint32_t)(keys[i] & 0xFFFFFFFFull);
    perm[i] = old;
    rank_of[old] = (uint32_t)i;
  }
}

__global__ void k_relabel(uint32_t* __restrict__ ids, uint64_t n, const uint32_t* __restrict__ rank_of) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    ids[i] = rank_of[ids[i]];
}

static void build_relabel(pe_graph_s* g, dbuf<uint32_t>& ds, dbuf<uint32_t>& dd, bool sym) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint32_t V = g->V;
  dbuf<uint32_t> od(ctx, V);
  PE_HIP(hipMemsetAsync(od.p, 0, (uint64_t)V * 4, s));
  if (g->edges)
    k_entry_out_degrees<<<grid_for(g->edges, 256, 1u << 20), 256, 0, s>>>(ds.p, dd.p, g->edges, sym, od.p);
  dbuf<uint64_t> keys(ctx, V);
  k_degree_keys<<<grid_for(V, 256, 8192), 256, 0, s>>>(od.p, V, keys.p);
  PE_HIP(hipGetLastError());
  od.release();
  sort_keys(ctx, keys, nullptr, V, 64);
  g->perm.alloc(ctx, V);
  g->rank_of.alloc(ctx, V);
  k_perm_from_keys<<<grid_for(V, 256, 8192), 256, 0, s>>>(keys.p, V, g->perm.p, g->rank_of.p);
  if (g->edges) {
    k_relabel<<<grid_for(g->edges, 256, 1u << 20), 256, 0, s>>>(ds.p, g->edges, g->rank_of.p);
    k_relabel<<<grid_for(g->edges, 256, 1u << 20), 256, 0, s>>>(dd.p, g->edges, g->rank_of.p);
  }
  PE_HIP(hipGetLastError());
  g->relabeled = true;
}

// Common builder over a device edge list.
void build_from_edges(pe_graph_s* g, dbuf<uint32_t>& ds, dbuf<uint32_t>& dd, dbuf<float>& dw,
                      uint32_t flags) {
  pe_ctx_s* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const bool sym = flags & PE_GRAPH_SYMMETRIZE;
  const bool weighted = flags & PE_GRAPH_WEIGHTED;
  const uint64_t E = g->edges;
  g->directed = !sym;
  g->nnz = E * (sym ? 2 : 1);
  g->has_weights = weighted;
  const int b = bits_for(g->V);
  if (flags & PE_GRAPH_RELABEL_DEGREE) build_relabel(g, ds, dd, sym);
  auto make_keys = [&](int csr, dbuf<uint64_t>& keys, dbuf<float>* wk) {
    keys.alloc(ctx, std::max<uint64_t>(1, g->nnz));
    if (wk) wk->alloc(ctx, std::max<uint64_t>(1, g->nnz));
    if (E)
      k_edge_keys<<<grid_for(E, 256, 1u << 20), 256, 0, s>>>(ds.p, dd.p, weighted ? dw.p : nullptr, E, b, csr,
                                                             sym, keys.p, wk ? wk->p : nullptr);
    PE_HIP(hipGetLastError());
  };
  {
    dbuf<uint64_t> keys;
    dbuf<float> wk;
    make_keys(0, keys, weighted ? &wk : nullptr);
    finish_build(g, keys, weighted ? &wk : nullptr, b);
  }
  if (weighted) {
    dbuf<uint64_t> keys;
    dbuf<float> wk;
    make_keys(1, keys, &wk);
    normalise_weights(g, keys, wk, b);
  }
  if (sym) {
    g->has_csr = true;
    g->csr_is_csc = true;
  } else if (flags & PE_GRAPH_BUILD_CSR) {
    dbuf<uint64_t> keys;
    make_keys(1, keys, nullptr);
    build_csr_from_keys(g, keys, b);
    g->has_csr = true;
  }
  PE_HIP(hipStreamSynchronize(s));
}

// ---- id mapping helpers (relabeled graphs) -----------------------------------------
template <typename T>
__global__ void k_gather_by_rank(const T* __restrict__ in, const uint32_t* __restrict__ rank_of, uint32_t V,
                                 T* __restrict__ out) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = in[rank_of[v]];
}

void to_external_order(pe_graph_s* g, const void* in, void* out, size_t elem) {
  hipStream_t s = g->ctx->stream;
  if (elem == 8)
    k_gather_by_rank<uint64_t><<<grid_for(g->V, 256, 8192), 256, 0, s>>>((const uint64_t*)in, g->rank_of.p, g->V,
                                                                        (uint64_t*)out);
  else
    k_gather_by_rank<uint32_t><<<grid_for(g->V, 256, 8192), 256, 0, s>>>((const uint32_t*)in, g->rank_of.p, g->V,
                                                                        (uint32_t*)out);
  PE_HIP(hipGetLastError());
}

void to_internal_ids(pe_graph_s* g, const uint32_t* host_ids, uint32_t n, uint32_t* host_out) {
  if (!g->relabeled) {
    std::memcpy(host_out, host_ids, (size_t)n * 4);
    return;
  }
  hipStream_t s = g->ctx->stream;
  dbuf<uint32_t> d(g->ctx, std::max<uint32_t>(1, n));
  PE_HIP(hipMemcpyAsync(d.p, host_ids, (size_t)n * 4, hipMemcpyHostToDevice, s));
  k_relabel<<<grid_for(n, 256, 4096), 256, 0, s>>>(d.p, n, g->rank_of.p);
  PE_HIP(hipGetLastError());
  PE_HIP(hipMemcpyAsync(host_out, d.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  PE_HIP(hipStreamSynchronize(s));
}

// Download a V-long device array in external (input) id order.
void download_external(pe_graph_s* g, const void* dev, void* host, size_t elem) {
  hipStream_t s = g->ctx->stream;
  const uint64_t bytes = (uint64_t)g->V * elem;
  if (g->relabeled) {
    dbuf<uint8_t> tmp(g->ctx, std::max<uint64_t>(1, bytes));
    to_external_order(g, dev, tmp.p, elem);
    PE_HIP(hipMemcpyAsync(host, tmp.p, bytes, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
  } else {
    PE_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace pe

using namespace pe;

extern "C" {

int pe_graph_create_rmat(pe_ctx_t* ctx, uint32_t scale, uint32_t edge_factor, uint64_t seed,
                         double a, double b_, double c, uint32_t flags, pe_graph_t** out) {
  if (!ctx || !out) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&] {
    PE_REQUIRE(scale >= 1 && scale <= 31, PE_ERR_INVALID, "scale must be in [1, 31]");
    PE_REQUIRE(a >= 0 && b_ >= 0 && c >= 0 && a + b_ + c <= 1.0, PE_ERR_INVALID, "bad RMAT probabilities");
    hipStream_t s = ctx->stream;
    auto* g = new pe_graph_s();
    std::unique_ptr<pe_graph_s> guard(g);
    g->ctx = ctx;
    g->V = 1u << scale;
    g->edges = (uint64_t)edge_factor << scale;
    const rmat_params p = make_rmat(scale, seed, a, b_, c);
    const uint64_t E = g->edges;
    dbuf<uint32_t> ds(ctx, std::max<uint64_t>(1, E)), dd(ctx, std::max<uint64_t>(1, E));
    dbuf<float> dw;
    if (flags & PE_GRAPH_WEIGHTED) dw.alloc(ctx, std::max<uint64_t>(1, E));
    if (E)
      k_rmat_edges<<<grid_for(E, 256, 1u << 20), 256, 0, s>>>(p, 0, E, ds.p, dd.p,
                                                              (flags & PE_GRAPH_WEIGHTED) ? dw.p : nullptr);
    PE_HIP(hipGetLastError());
    build_from_edges(g, ds, dd, dw, flags);
    *out = guard.release();
  });
}

int pe_rmat_edges(pe_ctx_t* ctx, uint32_t scale, uint64_t seed, double a, double b, double c,
                  uint64_t e0, uint64_t n, uint32_t* src, uint32_t* dst, float* w) {
  if (!ctx || (n && (!src || !dst))) return PE_ERR_INVALID;
  return guarded(ctx, [&] {
    PE_REQUIRE(scale >= 1 && scale <= 31, PE_ERR_INVALID, "scale must be in [1, 31]");
    if (!n) return;
    hipStream_t s = ctx->stream;
    const rmat_params p = make_rmat(scale, seed, a, b, c);
    dbuf<uint32_t> ds(ctx, n), dd(ctx, n);
    dbuf<float> dw;
    if (w) dw.alloc(ctx, n);
    k_rmat_edges<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(p, e0, n, ds.p, dd.p, w ? dw.p : nullptr);
    PE_HIP(hipGetLastError());
    PE_HIP(hipMemcpyAsync(src, ds.p, n * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(dst, dd.p, n * 4, hipMemcpyDeviceToHost, s));
    if (w) PE_HIP(hipMemcpyAsync(w, dw.p, n * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
  });
}

int pe_graph_create(pe_ctx_t* ctx, const uint32_t* src, const uint32_t* dst, const float* w,
                    uint64_t E, uint32_t V, uint32_t flags, pe_graph_t** out) {
  if (!ctx || !out || (E && (!src || !dst))) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&] {
    PE_REQUIRE(V >= 1, PE_ERR_INVALID, "V must be >= 1");
    for (uint64_t i = 0; i < E; ++i)
      PE_REQUIRE(src[i] < V && dst[i] < V, PE_ERR_RANGE, "edge endpoint >= V");
    const bool weighted = (flags & PE_GRAPH_WEIGHTED) != 0;
    PE_REQUIRE(!weighted || w, PE_ERR_INVALID, "PE_GRAPH_WEIGHTED needs weights");
    hipStream_t s = ctx->stream;
    auto* g = new pe_graph_s();
    std::unique_ptr<pe_graph_s> guard(g);
    g->ctx = ctx;
    g->V = V;
    g->edges = E;
    dbuf<uint32_t> ds(ctx, std::max<uint64_t>(1, E)), dd(ctx, std::max<uint64_t>(1, E));
    dbuf<float> dw;
    if (E) {
      PE_HIP(hipMemcpyAsync(ds.p, src, E * 4, hipMemcpyHostToDevice, s));
      PE_HIP(hipMemcpyAsync(dd.p, dst, E * 4, hipMemcpyHostToDevice, s));
    }
    if (weighted) {
      dw.alloc(ctx, std::max<uint64_t>(1, E));
      if (E) PE_HIP(hipMemcpyAsync(dw.p, w, E * 4, hipMemcpyHostToDevice, s));
    }
    build_from_edges(g, ds, dd, dw, flags);
    *out = guard.release();
  });
}

int pe_graph_permutation(pe_graph_t* g, uint32_t* rank_of) {
  if (!g || !rank_of) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    if (!g->relabeled) {
      for (uint32_t v = 0; v < g->V; ++v) rank_of[v] = v;
      return;
    }
    PE_HIP(hipMemcpyAsync(rank_of, g->rank_of.p, (uint64_t)g->V * 4, hipMemcpyDeviceToHost, g->ctx->stream));
    PE_HIP(hipStreamSynchronize(g->ctx->stream));
  });
}

int pe_graph_destroy(pe_graph_t* g) {
  if (!g) return PE_ERR_INVALID;
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
  return PE_OK;
}

int pe_graph_get_info(const pe_graph_t* g, pe_graph_info_t* info) {
  if (!g || !info) return PE_ERR_INVALID;
  std::memset(info, 0, sizeof(*info));
  info->V = g->V;
  info->max_v_i = g->max_v_i;
  info->vertices = g->vertices;
  info->max_in_degree = g->max_in_degree;
  info->edges = g->edges;
  info->nnz = g->nnz;
  info->is_directed = g->directed ? 1 : 0;
  info->has_csr = g->has_csr ? 1 : 0;
  info->has_weights = g->has_weights ? 1 : 0;
  info->relabeled = g->relabeled ? 1 : 0;
  info->tile_items = g->tile_items;
  info->num_tiles = g->num_tiles;
  info->num_fixup_rows = g->num_fix;
  info->device_bytes = g->device_bytes();
  info->pull_layout = g->pull_layout;
  info->binned_runs = g->pb ? (uint32_t)std::min<uint64_t>(g->pb->nrecords, 0xFFFFFFFFull) : 0;
  info->binned_bytes = g->pb ? g->pb->bytes() : 0;
  return PE_OK;
}

int pe_graph_download_csc(pe_graph_t* g, uint64_t* col_ptr, uint32_t* row_idx, float* wn) {
  if (!g) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    hipStream_t s = g->ctx->stream;
    if (col_ptr) PE_HIP(hipMemcpyAsync(col_ptr, g->col_ptr.p, ((uint64_t)g->V + 1) * 8, hipMemcpyDeviceToHost, s));
    if (row_idx && g->nnz) PE_HIP(hipMemcpyAsync(row_idx, g->row_idx.p, g->nnz * 4, hipMemcpyDeviceToHost, s));
    if (wn) {
      PE_REQUIRE(g->has_weights, PE_ERR_STATE, "graph has no weights");
      if (g->nnz) PE_HIP(hipMemcpyAsync(wn, g->wn.p, g->nnz * 4, hipMemcpyDeviceToHost, s));
    }
    PE_HIP(hipStreamSynchronize(s));
  });
}

int pe_graph_download_csr(pe_graph_t* g, uint64_t* row_ptr, uint32_t* col_idx) {
  if (!g) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(g->has_csr, PE_ERR_STATE, "graph has no CSR (PE_GRAPH_BUILD_CSR)");
    hipStream_t s = g->ctx->stream;
    if (row_ptr) PE_HIP(hipMemcpyAsync(row_ptr, g->csr_ptr(), ((uint64_t)g->V + 1) * 8, hipMemcpyDeviceToHost, s));
    if (col_idx && g->nnz) PE_HIP(hipMemcpyAsync(col_idx, g->csr_idx(), g->nnz * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
  });
}

int pe_graph_set_tile(pe_graph_t* g, uint32_t tile_items) {
  if (!g) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] { graph_build_tiles(g, tile_items); });
}

int pe_graph_out_degrees(pe_graph_t* g, uint32_t* od_host) {
  if (!g || !od_host) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] { download_external(g, g->od.p, od_host, 4); });
}

int pe_graph_in_degrees(pe_graph_t* g, uint32_t* id_host) {
  if (!g || !id_host) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    dbuf<uint32_t> tmp(g->ctx, std::max<uint32_t>(1, g->V));
    k_row_lengths<<<grid_for(g->V, 256, 65536), 256, 0, g->ctx->stream>>>(g->col_ptr.p, g->V, tmp.p);
    PE_HIP(hipGetLastError());
    download_external(g, tmp.p, id_host, 4);
  });
}

}  // extern "C"
