// Rank algorithms on the pull engine: PageRank, weighted PageRank, PPR /
// multi-seed PPR, and the engine-level operators (pull sum, apply, fused
// iterate, active-edge count, push/pull choice).
#include <algorithm>
#include <memory>

#include "pe_internal.hpp"
#include "pull.hpp"

namespace pe {

static constexpr uint64_t PPR_HUB_DEG = 4096;    // out-edges above which the push takes a vertex in chunks
static constexpr uint64_t PPR_HUB_CHUNK = 2048;

// x0(v) per the algorithm's init (pagerank.hpp:90-98, weighted_pagerank.hpp:104-109,
// personalized_pagerank.hpp:71-81, trustrank.hpp:197-206)
template <typename T>
__global__ void k_init_rank(uint32_t V, const uint32_t* __restrict__ od, const uint64_t* __restrict__ seeds,
                            double init_dist, int divide_by_od, T* __restrict__ x) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    double val = seeds ? (((seeds[v >> 6] >> (v & 63)) & 1ull) ? init_dist : 0.0) : init_dist;
    if (divide_by_od && od[v] > 0) val = val / (double)od[v];
    x[v] = (T)val;
  }
}

__global__ void k_init_ctl(iter_ctl* ctl, iter_ctl val) { *ctl = val; }

// Sum of od over the set bits of a bitmap (dualmode.hpp:264-277).
__global__ void k_active_edges(const uint64_t* __restrict__ words, uint64_t nwords, uint32_t V,
                               const uint32_t* __restrict__ od, unsigned long long* __restrict__ out) {
  unsigned long long acc = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nwords;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t w = words[i];
    while (w) {
      const int b = __ffsll((unsigned long long)w) - 1;
      const uint64_t v = i * 64 + b;
      if (v < V) acc += od[v];
      w &= w - 1;
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

uint64_t active_edges_dev(pe_graph_s* g, const uint64_t* words, uint64_t nwords) {
  hipStream_t s = g->ctx->stream;
  dbuf<unsigned long long> c(g->ctx, 1);
  PE_HIP(hipMemsetAsync(c.p, 0, 8, s));
  k_active_edges<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(words, nwords, g->V, g->od.p, c.p);
  PE_HIP(hipGetLastError());
  unsigned long long h = 0;
  PE_HIP(hipMemcpyAsync(&h, c.p, 8, hipMemcpyDeviceToHost, s));
  PE_HIP(hipStreamSynchronize(s));
  return h;
}

// Push scatter of PPR (personalized_pagerank.hpp:225-239): every vertex of the
// accumulated active set adds x(v) to each out-neighbour (FP atomics, the
// reference's write_add) and activates it.  One wave per bitmap word walks the
// word's set bits; lanes stride over each vertex's out-edges.
template <typename T>
__global__ __launch_bounds__(256) void k_ppr_push(const uint64_t* __restrict__ active, uint64_t nwords,
                                                  uint32_t V, const uint64_t* __restrict__ ptr,
                                                  const uint32_t* __restrict__ idx, const T* __restrict__ x,
                                                  T* __restrict__ y, uint64_t* __restrict__ newly) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id();
  const uint64_t nwaves = grid_waves();
  for (uint64_t wi = wave_id; wi < nwords; wi += nwaves) {
    uint64_t w = active[wi];
    while (w) {
      const int b = __ffsll((unsigned long long)w) - 1;
      w &= w - 1;
      const uint64_t v = wi * 64 + b;
      if (v >= V) break;
      if (ptr[v + 1] - ptr[v] > PPR_HUB_DEG) continue;  // k_ppr_push_hubs
      const T xv = x[v];
      for (uint64_t e = ptr[v] + lane; e < ptr[v + 1]; e += 64) {
        const uint32_t dst = idx[e];
        atomicOr((unsigned long long*)&newly[dst >> 6], 1ull << (dst & 63));
        atomicAdd(&y[dst], xv);
      }
    }
  }
}

// Push for the active hubs (more than PPR_HUB_DEG out-edges), one chunk of out-edges per
// wave at a time, so no wave walks a hub alone: the same scatter as k_ppr_push.
template <typename T>
__global__ __launch_bounds__(256) void k_ppr_push_hubs(const uint64_t* __restrict__ active,
                                                       const uint64_t* __restrict__ ptr,
                                                       const uint32_t* __restrict__ idx,
                                                       const uint32_t* __restrict__ hubs,
                                                       const unsigned long long* __restrict__ nhub,
                                                       const uint64_t* __restrict__ start, const T* __restrict__ x,
                                                       T* __restrict__ y, uint64_t* __restrict__ newly) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave_id = grid_wave_id(), nwaves = grid_waves();
  const uint64_t n = *nhub;
  if (n == 0) return;
  const uint64_t total = start[n];
  for (uint64_t c = wave_id; c < total; c += nwaves) {
    uint64_t k0, k1;
    const uint32_t v = hub_chunk(ptr, hubs, start, n, PPR_HUB_CHUNK, c, &k0, &k1);
    if (!((active[v >> 6] >> (v & 63)) & 1ull)) continue;  // uniform
    const T xv = x[v];
    for (uint64_t e = k0 + lane; e < k1; e += 64) {
      const uint32_t dst = idx[e];
      atomicOr((unsigned long long*)&newly[dst >> 6], 1ull << (dst & 63));
      atomicAdd(&y[dst], xv);
    }
  }
}

__global__ void k_or_words(uint64_t* __restrict__ a, const uint64_t* __restrict__ b, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    a[i] |= b[i];
}

template <typename T>
__global__ void k_zero_t(T* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (T)0;
}

// Seed set -> device bitmap (duplicates count once, like a bitmap_t of seeds).
struct seed_set {
  dbuf<uint64_t> words;
  uint32_t count = 0;
  void build(pe_ctx_s* ctx, uint32_t V, const uint32_t* seeds, uint32_t n) {
    std::vector<uint64_t> h(((uint64_t)V >> 6) + 1, 0ull);
    count = 0;
    for (uint32_t i = 0; i < n; ++i) {
      PE_REQUIRE(seeds[i] < V, PE_ERR_RANGE, "seed vertex >= V");
      uint64_t& w = h[seeds[i] >> 6];
      const uint64_t bit = 1ull << (seeds[i] & 63);
      if (!(w & bit)) { w |= bit; ++count; }
    }
    PE_REQUIRE(count > 0, PE_ERR_INVALID, "empty seed set");
    words.alloc(ctx, h.size());
    PE_HIP(hipMemcpyAsync(words.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    PE_HIP(hipStreamSynchronize(ctx->stream));
  }
};

// Device state (FP32 or FP64, internal order) -> host FP64 in input id order.
static void download_state(pe_graph_s* g, int prec, const void* dev, double* out) {
  if (prec == 64) {
    download_external(g, dev, out, 8);
  } else {
    std::vector<float> tmp(g->V);
    download_external(g, dev, tmp.data(), 4);
    for (uint32_t v = 0; v < g->V; ++v) out[v] = tmp[v];
  }
}

// Runs the reference loop on the device (pagerank.hpp:100-146 and variants).
struct rank_run {
  pe_graph_s* g;
  int prec;
  bool weighted;
  rank_args args{};
  dbuf<uint8_t> b0, b1;
  dbuf<iter_ctl> ctl;
  dbuf<double> deltas, rsums;
  rank_work wk;
  iter_ctl host_ctl{};
  double seconds = 0.0;

  rank_run(pe_graph_s* graph, int precision, bool w) : g(graph), prec(precision), weighted(w) {
    PE_REQUIRE(prec == 32 || prec == 64, PE_ERR_INVALID, "precision must be 32 or 64");
    PE_REQUIRE(!w || g->has_weights, PE_ERR_STATE, "graph has no weights");
    const size_t es = prec == 64 ? 8 : 4;
    b0.alloc(g->ctx, (uint64_t)std::max<uint32_t>(1, g->V) * es);
    b1.alloc(g->ctx, (uint64_t)std::max<uint32_t>(1, g->V) * es);
    ctl.alloc(g->ctx, 1);
    wk.init(g, prec);
  }

  void init_state(const uint64_t* seeds, double init_dist, int divide) {
    hipStream_t s = g->ctx->stream;
    if (prec == 64)
      k_init_rank<double><<<grid_for(g->V, 256, 8192), 256, 0, s>>>(g->V, g->od.p, seeds, init_dist, divide,
                                                                  (double*)b0.p);
    else
      k_init_rank<float><<<grid_for(g->V, 256, 8192), 256, 0, s>>>(g->V, g->od.p, seeds, init_dist, divide,
                                                                 (float*)b0.p);
    PE_HIP(hipGetLastError());
  }

  void run(uint32_t iters, double eps, double l1tol) {
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t cap = std::max<uint32_t>(1, iters);
    deltas.alloc(ctx, cap);
    rsums.alloc(ctx, cap);
    iter_ctl c{};
    c.iters = iters;
    c.t = 0;
    c.executed = 0;
    c.done = iters == 0 ? 1u : 0u;
    c.cur = 0;
    c.cap = cap;
    c.eps = eps;
    c.l1tol = l1tol;
    k_init_ctl<<<1, 1, 0, s>>>(ctl.p, c);
    PE_HIP(hipGetLastError());
    PE_HIP(hipEventRecord(ctx->t0, s));
    const bool early = eps > 0.0 || l1tol > 0.0;
    if (iters > 0) {
      if (ctx->timing) {
        // per-kernel events: plain launches
        uint32_t left = iters;
        while (left > 0) {
          const uint32_t chunk = std::min<uint32_t>(left, 8);
          for (uint32_t i = 0; i < chunk; ++i)
            rank_iteration(g, prec, b0.p, b1.p, weighted, args, ctl.p, wk, deltas.p, rsums.p);
          left -= chunk;
          if (early) {
            PE_HIP(hipMemcpyAsync(&host_ctl, ctl.p, sizeof(iter_ctl), hipMemcpyDeviceToHost, s));
            PE_HIP(hipStreamSynchronize(s));
            if (host_ctl.done) break;
          }
        }
      } else {
        // capture a chunk of iterations once, replay it until the controller is done
        const uint32_t chunk = early ? std::min<uint32_t>(iters, 8) : std::min<uint32_t>(iters, 16);
        stream_graph sg;  // destroyed after the stream drains, also on an exception
        sg.begin(s);
        for (uint32_t i = 0; i < chunk; ++i)
          rank_iteration(g, prec, b0.p, b1.p, weighted, args, ctl.p, wk, deltas.p, rsums.p);
        sg.end();
        uint32_t launched = 0;
        while (launched < iters) {
          sg.launch();
          launched += chunk;
          if (early) {
            PE_HIP(hipMemcpyAsync(&host_ctl, ctl.p, sizeof(iter_ctl), hipMemcpyDeviceToHost, s));
            PE_HIP(hipStreamSynchronize(s));
            if (host_ctl.done) break;
          }
        }
        PE_HIP(hipStreamSynchronize(s));
      }
    }
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipMemcpyAsync(&host_ctl, ctl.p, sizeof(iter_ctl), hipMemcpyDeviceToHost, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    PE_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    seconds = ms * 1e-3;
    PE_REQUIRE(host_ctl.done || iters == 0, PE_ERR_INTERNAL, "rank iteration did not finish");
  }

  void download(double* out, double* host_deltas) {
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    void* res = host_ctl.cur ? b1.p : b0.p;
    download_state(g, prec, res, out);
    if (host_deltas && host_ctl.executed)
      PE_HIP(hipMemcpyAsync(host_deltas, deltas.p,
                            std::min(host_ctl.executed, host_ctl.cap) * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
  }

  void fill_stats(pe_rank_stats_t* st) const {
    if (!st) return;
    st->iterations_run = host_ctl.executed;
    st->last_delta = host_ctl.last_delta;
    st->seconds = seconds;
    st->gteps = seconds > 0 ? (double)host_ctl.executed * (double)g->nnz / seconds / 1e9 : 0.0;
  }
};

static rank_args make_args(double damping, double base, double seed_mass, const uint64_t* seeds,
                           int divide, const pe_apply_params_t* p) {
  rank_args a{};
  a.damping = damping;
  a.base = base;
  a.seed_mass = seed_mass;
  a.seeds = seeds;
  a.divide_by_od = divide;
  a.final_iter = p ? p->final_iter : 0;
  a.mode = MODE_APPLY;
  return a;
}

static int vec_prec(const pe_vec_s* v) { return v->dtype == PE_F64 ? 64 : 32; }

}  // namespace pe

using namespace pe;

extern "C" {

int pe_active_edges(pe_graph_t* g, pe_bitmap_t* actives, uint64_t* out) {
  if (!g || !actives || !out) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(actives->bits >= g->V, PE_ERR_INVALID, "bitmap smaller than the vertex set");
    *out = active_edges_dev(g, actives->p, actives->words);
  });
}

int pe_choose_mode(pe_graph_t* g, pe_bitmap_t* actives, double threshold, int* is_push,
                   uint64_t* active_edges) {
  if (!g || !actives || !is_push) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(actives->bits >= g->V, PE_ERR_INVALID, "bitmap smaller than the vertex set");
    const uint64_t A = active_edges_dev(g, actives->p, actives->words);
    const uint64_t Ep = g->directed ? g->edges : g->edges * 2;  // dualmode.hpp:261-262
    *is_push = ((double)A / (double)Ep) < threshold ? 1 : 0;   // dualmode.hpp:279
    if (active_edges) *active_edges = A;
  });
}

int pe_pull_sum(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, int weighted) {
  if (!g || !x || !y) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(x->dtype == y->dtype && (x->dtype == PE_F32 || x->dtype == PE_F64), PE_ERR_INVALID,
               "x and y must both be PE_F32 or both PE_F64");
    PE_REQUIRE(x->n >= g->V && y->n >= g->V, PE_ERR_INVALID, "vectors shorter than V");
    PE_REQUIRE(!weighted || g->has_weights, PE_ERR_STATE, "graph has no weights");
    rank_work wk;
    wk.init(g, vec_prec(x));
    rank_args a = make_args(1.0, 0.0, 0.0, nullptr, 0, nullptr);
    a.mode = MODE_RAW_SUM;
    rank_iteration(g, vec_prec(x), x->p, y->p, weighted != 0, a, nullptr, wk, wk.scalars,
                   wk.scalars + 1);
    PE_HIP(hipStreamSynchronize(g->ctx->stream));
  });
}

int pe_pull_sum_noalloc(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, int weighted) {
  if (!g || !x || !y) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(x->dtype == y->dtype && (x->dtype == PE_F32 || x->dtype == PE_F64), PE_ERR_INVALID,
               "x and y must both be PE_F32 or both PE_F64");
    PE_REQUIRE(x->n >= g->V && y->n >= g->V, PE_ERR_INVALID, "vectors shorter than V");
    PE_REQUIRE(!weighted || g->has_weights, PE_ERR_STATE, "graph has no weights");
    PE_REQUIRE(g->ctx->noalloc_scratch != nullptr, PE_ERR_STATE, "context has no scratch array");
    // the same kernels as pe_pull_sum, their tile partials and per-block sums in this
    // GPU's static scratch array instead of fresh allocations (nothing there needs to
    // start at zero: every word is written before it is read)
    rank_work wk;
    wk.init_in(g, vec_prec(x), static_cast<uint8_t*>(g->ctx->noalloc_scratch), PE_PULL_SUM_NOALLOC_MAX_BYTES);
    rank_args a = make_args(1.0, 0.0, 0.0, nullptr, 0, nullptr);
    a.mode = MODE_RAW_SUM;
    rank_iteration(g, vec_prec(x), x->p, y->p, weighted != 0, a, nullptr, wk, wk.scalars, wk.scalars + 1);
    PE_HIP(hipStreamSynchronize(g->ctx->stream));
  });
}

int pe_rank_apply(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, const pe_apply_params_t* p, double* delta) {
  if (!g || !x || !y || !p) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(x->dtype == y->dtype && (x->dtype == PE_F32 || x->dtype == PE_F64), PE_ERR_INVALID,
               "x and y must both be PE_F32 or both PE_F64");
    PE_REQUIRE(x->n >= g->V && y->n >= g->V, PE_ERR_INVALID, "vectors shorter than V");
    rank_work wk;
    wk.init(g, vec_prec(x));
    const rank_args a = make_args(p->damping, p->base, p->seed_mass, p->seeds ? p->seeds->p : nullptr,
                                  p->divide_by_od, p);
    rank_apply_only(g, vec_prec(x), x->p, y->p, a, wk);
    double h = 0.0;
    PE_HIP(hipMemcpyAsync(&h, wk.scalars, 8, hipMemcpyDeviceToHost, g->ctx->stream));
    PE_HIP(hipStreamSynchronize(g->ctx->stream));
    if (delta) *delta = h;
  });
}

int pe_rank_iterate(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, int weighted, const pe_apply_params_t* p,
                    double* delta) {
  if (!g || !x || !y || !p) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(x->dtype == y->dtype && (x->dtype == PE_F32 || x->dtype == PE_F64), PE_ERR_INVALID,
               "x and y must both be PE_F32 or both PE_F64");
    PE_REQUIRE(x->n >= g->V && y->n >= g->V, PE_ERR_INVALID, "vectors shorter than V");
    PE_REQUIRE(!weighted || g->has_weights, PE_ERR_STATE, "graph has no weights");
    rank_work wk;
    wk.init(g, vec_prec(x));
    rank_args a = make_args(p->damping, p->base, p->seed_mass, p->seeds ? p->seeds->p : nullptr,
                            p->divide_by_od, p);
    if (vec_prec(x) == 32 && g->pull_layout == PE_PULL_BINNED && g->pb)
      a.mass = binned_mass_bound(g, (const float*)x->p, weighted != 0);
    rank_iteration(g, vec_prec(x), x->p, y->p, weighted != 0, a, nullptr, wk, wk.scalars,
                   wk.scalars + 1);
    double h = 0.0;
    PE_HIP(hipMemcpyAsync(&h, wk.scalars, 8, hipMemcpyDeviceToHost, g->ctx->stream));
    PE_HIP(hipStreamSynchronize(g->ctx->stream));
    if (delta) *delta = h;
  });
}

int pe_pagerank(pe_graph_t* g, const pe_rank_opts_t* opts, double* out, double* deltas,
                pe_rank_stats_t* stats) {
  if (!g || !opts || !out) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    const bool w = opts->weighted != 0;
    rank_run run(g, opts->precision ? opts->precision : 32, w);
    // PageRank: x = 1/od (pagerank.hpp:90-98); weighted: x = 1 (weighted_pagerank.hpp:104-109)
    run.init_state(nullptr, 1.0, w ? 0 : 1);
    run.args = make_args(opts->damping, 1.0 - opts->damping, 0.0, nullptr, w ? 0 : 1, nullptr);
    run.args.mass = (double)g->V;  // ranks start at 1; the total stays at most V
    run.run(opts->iterations, opts->eps, opts->l1_norm_tol);
    run.download(out, deltas);
    run.fill_stats(stats);
  });
}

int pe_ppr(pe_graph_t* g, const uint32_t* seeds, uint32_t nseeds, const pe_rank_opts_t* opts,
           double* out, double* deltas, pe_rank_stats_t* stats) {
  if (!g || !opts || !out || !seeds || nseeds == 0) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(!opts->weighted, PE_ERR_INVALID, "PPR is unweighted");
    std::vector<uint32_t> internal(nseeds);
    for (uint32_t i = 0; i < nseeds; ++i) PE_REQUIRE(seeds[i] < g->V, PE_ERR_RANGE, "seed vertex >= V");
    to_internal_ids(g, seeds, nseeds, internal.data());
    seed_set ss;
    ss.build(g->ctx, g->V, internal.data(), nseeds);
    const double init_dist = 1.0 / ss.count;
    rank_run run(g, opts->precision ? opts->precision : 32, false);
    run.init_state(ss.words.p, init_dist, 1);
    run.args = make_args(opts->damping, 0.0, (1.0 - opts->damping) * init_dist, ss.words.p, 1, nullptr);
    run.args.mass = 1.0;  // the restart mass is 1 in total, and so is the starting state
    run.run(opts->iterations, opts->eps, opts->l1_norm_tol);
    run.download(out, deltas);
    run.fill_stats(stats);
  });
}

int pe_rank_init(pe_graph_t* g, pe_vec_t* x, const pe_bitmap_t* seeds, double value, int divide_by_od) {
  if (!g || !x) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(x->dtype == PE_F32 || x->dtype == PE_F64, PE_ERR_INVALID, "x must be PE_F32 or PE_F64");
    PE_REQUIRE(x->n >= g->V, PE_ERR_INVALID, "vector shorter than V");
    PE_REQUIRE(!seeds || seeds->bits >= g->V, PE_ERR_INVALID, "bitmap smaller than the vertex set");
    hipStream_t s = g->ctx->stream;
    const uint64_t* sw = seeds ? seeds->p : nullptr;
    if (x->dtype == PE_F64)
      k_init_rank<double><<<grid_for(g->V, 256, 8192), 256, 0, s>>>(g->V, g->od.p, sw, value, divide_by_od,
                                                                  static_cast<double*>(x->p));
    else
      k_init_rank<float><<<grid_for(g->V, 256, 8192), 256, 0, s>>>(g->V, g->od.p, sw, value, divide_by_od,
                                                                 static_cast<float*>(x->p));
    PE_HIP(hipGetLastError());
    PE_HIP(hipStreamSynchronize(s));
  });
}

int pe_push_add(pe_graph_t* g, pe_bitmap_t* actives, pe_vec_t* x, pe_vec_t* y) {
  if (!g || !actives || !x || !y) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(x->dtype == y->dtype && (x->dtype == PE_F32 || x->dtype == PE_F64), PE_ERR_INVALID,
               "x and y must both be PE_F32 or both PE_F64");
    PE_REQUIRE(x->n >= g->V && y->n >= g->V, PE_ERR_INVALID, "vectors shorter than V");
    PE_REQUIRE(actives->bits >= g->V, PE_ERR_INVALID, "bitmap smaller than the vertex set");
    PE_REQUIRE(g->has_csr, PE_ERR_STATE, "push mode needs out-edges (PE_GRAPH_BUILD_CSR)");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t V = g->V;
    const uint64_t nwords = ((uint64_t)V >> 6) + 1;
    dbuf<uint64_t> newly(ctx, nwords);
    k_zero_t<uint64_t><<<grid_for(nwords, 256, 4096), 256, 0, s>>>(newly.p, nwords);
    hub_table hubs;  // out-edge hubs, pushed in chunks
    hub_table_build(ctx, g->csr_ptr(), V, g->nnz, PPR_HUB_DEG, PPR_HUB_CHUNK, hubs);
    const unsigned pg = grid_for(nwords * 64, 256, 4096);
    const unsigned hg = persistent_grid(ctx, 1u << 20, 4);
    if (x->dtype == PE_F64) {
      k_ppr_push<double><<<pg, 256, 0, s>>>(actives->p, nwords, V, g->csr_ptr(), g->csr_idx(),
                                           static_cast<const double*>(x->p), static_cast<double*>(y->p), newly.p);
      k_ppr_push_hubs<double><<<hg, 256, 0, s>>>(actives->p, g->csr_ptr(), g->csr_idx(), hubs.hubs.p, hubs.count.p,
                                                 hubs.start.p, static_cast<const double*>(x->p),
                                                 static_cast<double*>(y->p), newly.p);
    } else {
      k_ppr_push<float><<<pg, 256, 0, s>>>(actives->p, nwords, V, g->csr_ptr(), g->csr_idx(),
                                          static_cast<const float*>(x->p), static_cast<float*>(y->p), newly.p);
      k_ppr_push_hubs<float><<<hg, 256, 0, s>>>(actives->p, g->csr_ptr(), g->csr_idx(), hubs.hubs.p, hubs.count.p,
                                                hubs.start.p, static_cast<const float*>(x->p),
                                                static_cast<float*>(y->p), newly.p);
    }
    k_or_words<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(actives->p, newly.p, nwords);
    PE_HIP(hipGetLastError());
    PE_HIP(hipStreamSynchronize(s));
  });
}

int pe_ppr_pushpull(pe_graph_t* g, uint32_t src, const pe_rank_opts_t* opts, double* out,
                    double* deltas, uint8_t* modes, pe_rank_stats_t* stats) {
  if (!g || !opts || !out) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    PE_REQUIRE(src < g->V, PE_ERR_RANGE, "src >= V");
    PE_REQUIRE(g->has_csr, PE_ERR_STATE, "push mode needs out-edges (PE_GRAPH_BUILD_CSR)");
    PE_REQUIRE(!opts->weighted, PE_ERR_INVALID, "PPR is unweighted");
    pe_ctx_s* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const int prec = opts->precision ? opts->precision : 32;
    PE_REQUIRE(prec == 32 || prec == 64, PE_ERR_INVALID, "precision must be 32 or 64");
    const size_t es = prec == 64 ? 8 : 4;
    const uint32_t V = g->V;
    uint32_t isrc = src;
    to_internal_ids(g, &src, 1, &isrc);
    seed_set ss;
    ss.build(ctx, V, &isrc, 1);
    dbuf<uint8_t> b0(ctx, (uint64_t)V * es), b1(ctx, (uint64_t)V * es);
    const uint64_t nwords = ((uint64_t)V >> 6) + 1;
    dbuf<uint64_t> active(ctx, nwords), newly(ctx, nwords);
    PE_HIP(hipMemcpyAsync(active.p, ss.words.p, nwords * 8, hipMemcpyDeviceToDevice, s));
    rank_work wk;
    wk.init(g, prec);
    hub_table hubs;  // out-edge hubs, pushed in chunks
    hub_table_build(ctx, g->csr_ptr(), V, g->nnz, PPR_HUB_DEG, PPR_HUB_CHUNK, hubs);
    if (prec == 64)
      k_init_rank<double><<<grid_for(V, 256, 8192), 256, 0, s>>>(V, g->od.p, ss.words.p, 1.0, 1, (double*)b0.p);
    else
      k_init_rank<float><<<grid_for(V, 256, 8192), 256, 0, s>>>(V, g->od.p, ss.words.p, 1.0, 1, (float*)b0.p);
    PE_HIP(hipGetLastError());
    rank_args a = make_args(opts->damping, 0.0, 1.0 - opts->damping, ss.words.p, 1, nullptr);
    a.mass = 1.0;
    const uint64_t Ep = g->directed ? g->edges : g->edges * 2;
    bool do_pull = false;
    uint32_t executed = 0;
    double last = 0.0;
    void* x = b0.p;
    void* y = b1.p;
    PE_HIP(hipEventRecord(ctx->t0, s));
    for (uint32_t t = 0; t < opts->iterations; ++t) {
      const bool final_iter = (opts->iterations - 1 == t);
      a.final_iter = final_iter ? 1 : 0;
      if (!do_pull) do_pull = active_edges_dev(g, active.p, nwords) > Ep / 20;  // (188-205)
      if (do_pull) {
        rank_iteration(g, prec, x, y, false, a, nullptr, wk, wk.scalars, wk.scalars + 1);
      } else {
        const unsigned zg = grid_for(V, 256, 8192);
        if (prec == 64) k_zero_t<double><<<zg, 256, 0, s>>>((double*)y, V);
        else k_zero_t<float><<<zg, 256, 0, s>>>((float*)y, V);
        k_zero_t<uint64_t><<<grid_for(nwords, 256, 4096), 256, 0, s>>>(newly.p, nwords);
        const unsigned pg = grid_for(nwords * 64, 256, 4096);
        const unsigned hg = persistent_grid(ctx, 1u << 20, 4);
        if (prec == 64) {
          k_ppr_push<double><<<pg, 256, 0, s>>>(active.p, nwords, V, g->csr_ptr(), g->csr_idx(),
                                               (const double*)x, (double*)y, newly.p);
          k_ppr_push_hubs<double><<<hg, 256, 0, s>>>(active.p, g->csr_ptr(), g->csr_idx(), hubs.hubs.p,
                                                     hubs.count.p, hubs.start.p, (const double*)x, (double*)y,
                                                     newly.p);
        } else {
          k_ppr_push<float><<<pg, 256, 0, s>>>(active.p, nwords, V, g->csr_ptr(), g->csr_idx(),
                                              (const float*)x, (float*)y, newly.p);
          k_ppr_push_hubs<float><<<hg, 256, 0, s>>>(active.p, g->csr_ptr(), g->csr_idx(), hubs.hubs.p,
                                                    hubs.count.p, hubs.start.p, (const float*)x, (float*)y,
                                                    newly.p);
        }
        k_or_words<<<grid_for(nwords, 256, 4096), 256, 0, s>>>(active.p, newly.p, nwords);
        PE_HIP(hipGetLastError());
        rank_apply_only(g, prec, x, y, a, wk);
      }
      double delta = 0.0;
      PE_HIP(hipMemcpyAsync(&delta, wk.scalars, 8, hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      if (modes) modes[executed] = do_pull ? 0 : 1;
      if (deltas) deltas[executed] = final_iter ? 0.0 : delta;
      last = final_iter ? 0.0 : delta;
      ++executed;
      if (!final_iter && opts->eps > 0.0 && delta < opts->eps) t = opts->iterations - 2;
      std::swap(x, y);
    }
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    download_state(g, prec, x, out);
    if (stats) {
      stats->iterations_run = executed;
      stats->last_delta = last;
      stats->seconds = ms * 1e-3;
      stats->gteps = 0.0;
    }
  });
}

}  // extern "C"
