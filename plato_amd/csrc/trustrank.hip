// TrustRank (plato/algo/trustrank/trustrank.hpp:23-259) as a client of the rank
// engine: seed selection (select_good_seeds, :80-145) ranks vertices with the
// pull PageRank on the device -- over the reversed graph for inverse PageRank --
// and picks the top seed_num by (rank desc, id asc) with a stable device radix
// sort; the iteration is the seed-set restart of pe_ppr (:163-259).
#include <random>
#include <unordered_set>
#include <vector>

#include "pe_internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

namespace pe {

// Reversed edges of a directed graph, in input ids: entry (u -> v) of the CSC
// becomes (v -> u).  One thread per column.
__global__ void k_reverse_edges(const uint64_t* __restrict__ col_ptr, const uint32_t* __restrict__ row_idx,
                                uint32_t V, const uint32_t* __restrict__ perm, uint32_t* __restrict__ src,
                                uint32_t* __restrict__ dst) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t ev = perm ? perm[v] : (uint32_t)v;
    for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) {
      const uint32_t u = row_idx[e];
      src[e] = ev;
      dst[e] = perm ? perm[u] : u;
    }
  }
}

// FP64 bit pattern -> unsigned key in the same order as the value
__global__ void k_rank_keys(uint64_t* __restrict__ keys, uint32_t* __restrict__ ids, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = keys[i];
    keys[i] = (b >> 63) ? ~b : (b | (1ull << 63));
    ids[i] = (uint32_t)i;
  }
}

// Ids of the k largest of r[0..n) by (rank desc, id asc): a stable descending radix
// sort of (rank, id) pairs whose ids start ascending.
static std::vector<uint32_t> top_by_rank(pe_ctx_s* ctx, const double* r, uint64_t n, uint32_t k) {
  hipStream_t s = ctx->stream;
  dbuf<uint64_t> keys(ctx, n), kalt(ctx, n);
  dbuf<uint32_t> ids(ctx, n), ialt(ctx, n);
  PE_HIP(hipMemcpyAsync(keys.p, r, n * 8, hipMemcpyHostToDevice, s));
  k_rank_keys<<<grid_for(n, 256, 65536), 256, 0, s>>>(keys.p, ids.p, n);
  PE_HIP(hipGetLastError());
  rocprim::double_buffer<uint64_t> kb(keys.p, kalt.p);
  rocprim::double_buffer<uint32_t> vb(ids.p, ialt.p);
  size_t temp = 0;
  PE_HIP(rocprim::radix_sort_pairs_desc(nullptr, temp, kb, vb, n, 0u, 64u, s));
  dbuf<uint8_t> tmp(ctx, temp ? temp : 1);
  PE_HIP(rocprim::radix_sort_pairs_desc(tmp.p, temp, kb, vb, n, 0u, 64u, s));
  std::vector<uint32_t> out(std::min<uint64_t>(k, n));
  PE_HIP(hipMemcpyAsync(out.data(), vb.current(), out.size() * 4, hipMemcpyDeviceToHost, s));
  PE_HIP(hipStreamSynchronize(s));
  return out;
}

static bool bit(const std::vector<uint64_t>& w, uint32_t v) { return (w[v >> 6] >> (v & 63)) & 1ull; }

static void select_seeds(pe_graph_s* g, pe_graph_s* g_rev, const uint64_t* good_words, const pe_trustrank_opts_t& o,
                         std::vector<uint64_t>& seeds) {
  const uint32_t V = g->V;
  const uint64_t nwords = ((uint64_t)V >> 6) + 1;
  std::vector<uint64_t> good(nwords, 0);
  if (good_words) {
    std::copy(good_words, good_words + nwords, good.begin());
  } else {
    for (uint32_t v = 0; v < V; ++v) good[v >> 6] |= 1ull << (v & 63);
  }
  seeds.assign(nwords, 0);
  const uint64_t n = (uint64_t)g->max_v_i + 1;  // rank pairs cover ids 0..max_v_i (trustrank.hpp:65-68)
  if (n <= o.seed_num) {
    seeds = good;  // trustrank.hpp:95-98
    return;
  }
  if (o.select_method != 2) {
    pe_graph_s* pg = g;
    if (o.select_method == 0 && g->directed) {
      PE_REQUIRE(g_rev != nullptr, PE_ERR_INVALID, "inverse PageRank needs the reversed graph (pe_graph_create_reverse)");
      PE_REQUIRE(g_rev->V == V, PE_ERR_INVALID, "reversed graph has a different vertex count");
      pg = g_rev;
    }
    pe_rank_opts_t ro{};
    ro.iterations = 100;  // pagerank_opts_t defaults (pagerank.hpp:33-41)
    ro.damping = 0.85;
    ro.eps = 0.001;
    ro.precision = 64;
    std::vector<double> r(V);
    const int rc = pe_pagerank(pg, &ro, r.data(), nullptr, nullptr);
    PE_REQUIRE(rc == PE_OK, rc, pg->ctx->last_error);
    for (uint32_t v : top_by_rank(g->ctx, r.data(), n, o.seed_num))
      if (bit(good, v)) seeds[v >> 6] |= 1ull << (v & 63);
  } else {
    std::minstd_rand0 gen((std::minstd_rand0::result_type)o.random_seed);
    std::uniform_int_distribution<uint32_t> dist(0, g->max_v_i);
    std::unordered_set<uint32_t> drawn;
    while (drawn.size() < o.seed_num) {
      const uint32_t v = dist(gen);
      if (drawn.insert(v).second && bit(good, v)) seeds[v >> 6] |= 1ull << (v & 63);
    }
  }
}

}  // namespace pe

using namespace pe;

extern "C" {

int pe_graph_create_reverse(pe_graph_t* g, uint32_t flags, pe_graph_t** out) {
  if (!g || !out) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(g->directed, PE_ERR_STATE, "a symmetrized graph is its own reverse");
    const bool same_ids = flags & PE_GRAPH_SAME_IDS;
    PE_REQUIRE(!(same_ids && (flags & PE_GRAPH_RELABEL_DEGREE)), PE_ERR_INVALID,
               "PE_GRAPH_SAME_IDS keeps g's ids: it excludes PE_GRAPH_RELABEL_DEGREE");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    auto* r = new pe_graph_s();
    std::unique_ptr<pe_graph_s> guard(r);
    r->ctx = ctx;
    r->V = g->V;
    r->edges = g->nnz;
    dbuf<uint32_t> ds(ctx, std::max<uint64_t>(1, g->nnz)), dd(ctx, std::max<uint64_t>(1, g->nnz));
    dbuf<float> dw;
    if (g->nnz) {
      // same ids: the edges stay in g's internal ids; else they go back to input ids
      k_reverse_edges<<<grid_for(g->V, 256, 1u << 20), 256, 0, s>>>(
          g->col_ptr.p, g->row_idx.p, g->V, (g->relabeled && !same_ids) ? g->perm.p : nullptr, ds.p, dd.p);
      PE_HIP(hipGetLastError());
    }
    build_from_edges(r, ds, dd, dw, flags & (PE_GRAPH_BUILD_CSR | PE_GRAPH_RELABEL_DEGREE));
    if (same_ids) {
      r->id_space = g->id_space;
      if (g->relabeled) {
        r->relabeled = true;
        r->perm.alloc(ctx, g->V);
        r->rank_of.alloc(ctx, g->V);
        PE_HIP(hipMemcpyAsync(r->perm.p, g->perm.p, (size_t)g->V * 4, hipMemcpyDeviceToDevice, s));
        PE_HIP(hipMemcpyAsync(r->rank_of.p, g->rank_of.p, (size_t)g->V * 4, hipMemcpyDeviceToDevice, s));
      }
    }
    if (g->encoded) {
      r->encoded = true;
      r->vid_decode.alloc(ctx, g->V);
      PE_HIP(hipMemcpyAsync(r->vid_decode.p, g->vid_decode.p, (size_t)g->V * 8, hipMemcpyDeviceToDevice, s));
    }
    PE_HIP(hipStreamSynchronize(s));
    *out = guard.release();
  });
}

int pe_trustrank_select(pe_graph_t* g, pe_graph_t* g_rev, const uint64_t* good_words,
                        const pe_trustrank_opts_t* opts, uint64_t* seeds_words, uint32_t* nseeds) {
  if (!g || !opts || !seeds_words) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(opts->select_method <= 2, PE_ERR_INVALID, "select_method must be 0, 1 or 2");
    std::vector<uint64_t> seeds;
    select_seeds(g, g_rev, good_words, *opts, seeds);
    uint32_t c = 0;
    for (size_t i = 0; i < seeds.size(); ++i) {
      seeds_words[i] = seeds[i];
      c += (uint32_t)__builtin_popcountll(seeds[i]);
    }
    if (nseeds) *nseeds = c;
  });
}

int pe_trustrank(pe_graph_t* g, pe_graph_t* g_rev, const uint64_t* good_words, const pe_trustrank_opts_t* opts,
                 double* out, uint64_t* seeds_words, pe_rank_stats_t* stats) {
  if (!g || !opts || !out) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(opts->select_method <= 2, PE_ERR_INVALID, "select_method must be 0, 1 or 2");
    std::vector<uint64_t> seeds;
    select_seeds(g, g_rev, good_words, *opts, seeds);
    std::vector<uint32_t> list;
    for (uint32_t v = 0; v < g->V; ++v)
      if ((seeds[v >> 6] >> (v & 63)) & 1ull) list.push_back(v);
    PE_REQUIRE(!list.empty(), PE_ERR_STATE, "no good seeds were selected");
    if (seeds_words) std::copy(seeds.begin(), seeds.end(), seeds_words);
    pe_rank_opts_t ro{};
    ro.iterations = opts->iterations;
    ro.damping = opts->damping;
    ro.eps = opts->eps;
    ro.precision = opts->precision ? opts->precision : 64;
    const int rc = pe_ppr(g, list.data(), (uint32_t)list.size(), &ro, out, nullptr, stats);
    PE_REQUIRE(rc == PE_OK, rc, g->ctx->last_error);
  });
}

}  // extern "C"
