// Autotuner: times the engine's tunables on the resident graph and keeps the
// fastest.  Tunables and the knob each one sets:
//   pull tile size      merge-path items per tile of the pull gather (1024/2048/4096);
//                       applied to the graph
//   batch task size     edges per wave task of the 64-column PPR block (256 to 16384); longer
//                       columns are split (the block's degree threshold)
//   BFS switch ratio    push iff active edges / E' < ratio (the reference's rule,
//                       0.05 by default; levels do not depend on it)
// Each candidate runs a few real passes; timing is by HIP events.  A choice can be
// persisted with pe_autotune_save / pe_autotune_load keyed by the graph's shape.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pe_internal.hpp"

namespace {

double time_pull(pe_graph_s* g, int iters) {
  pe_rank_opts_t o{};
  o.iterations = (uint32_t)iters;
  o.damping = 0.85;
  o.precision = 32;
  std::vector<double> out(g->V);
  pe_rank_stats_t st{};
  // warm-up then timed
  pe_rank_opts_t w = o;
  w.iterations = 1;
  if (pe_pagerank(g, &w, out.data(), nullptr, &st) != PE_OK) pe::fail(PE_ERR_INTERNAL, g->ctx->last_error);
  if (pe_pagerank(g, &o, out.data(), nullptr, &st) != PE_OK) pe::fail(PE_ERR_INTERNAL, g->ctx->last_error);
  return st.seconds * 1e3 / iters;
}

double time_batch(pe_graph_s* g, uint32_t task_edges, int iters) {
  pe_ppr_batch_opts_t bo{};
  bo.k = 64;
  bo.damping = 0.85;
  bo.tol = 0.0;  // keep every column active
  bo.max_iters = 1000000;
  bo.task_edges = task_edges;
  pe_ppr_batch_t* b = nullptr;
  int rc = pe_ppr_batch_create(g, &bo, &b);
  if (rc != PE_OK) pe::fail(rc, g->ctx->last_error);
  for (uint32_t s = 0; s < 64; ++s) {
    const uint32_t seed = (uint32_t)(((uint64_t)s * 2654435761u) % g->V);
    if (pe_ppr_batch_admit(b, s, &seed, 1) != PE_OK) {
      pe_ppr_batch_destroy(b);
      pe::fail(PE_ERR_INTERNAL, g->ctx->last_error);
    }
  }
  pe_ppr_batch_step(b, 2);  // loading pass + one warm iteration
  pe_ctx_timer_start(g->ctx);
  pe_ppr_batch_step(b, (uint32_t)iters);
  double ms = 0;
  pe_ctx_timer_stop(g->ctx, &ms);
  pe_ppr_batch_destroy(b);
  return ms / iters;
}

double time_bfs(pe_graph_s* g, double ratio, const std::vector<uint32_t>& roots) {
  double total = 0;
  for (uint32_t r : roots) {
    pe_bfs_stats_t st{};
    if (pe_bfs(g, r, ratio, nullptr, nullptr, nullptr, nullptr, 0, &st) != PE_OK)
      pe::fail(PE_ERR_INTERNAL, g->ctx->last_error);
    total += st.seconds;
  }
  return total * 1e3 / std::max<size_t>(1, roots.size());
}

}  // namespace

extern "C" {

int pe_autotune(pe_graph_t* g, uint32_t what, pe_tune_result_t* out) {
  if (!g || !out) return PE_ERR_INVALID;
  return pe::guarded(g->ctx, [&] {
    pe_tune_result_t r{};
    r.tile_items = g->tile_items;
    r.batch_task_edges = 4096;
    r.bfs_push_threshold = 0.05;
    if (what & PE_TUNE_PULL) {
      const uint32_t cand[3] = {1024, 2048, 4096};
      double best = 1e300;
      for (uint32_t t : cand) {
        pe::graph_build_tiles(g, t);
        const double ms = time_pull(g, 3);
        if (ms < best) {
          best = ms;
          r.tile_items = t;
        }
      }
      pe::graph_build_tiles(g, r.tile_items);
      r.pull_ms = best;
    }
    if (what & PE_TUNE_BATCH) {
      const uint32_t cand[4] = {256, 1024, 4096, 16384};
      double best = 1e300;
      for (uint32_t c : cand) {
        const double ms = time_batch(g, c, 2);
        if (ms < best) {
          best = ms;
          r.batch_task_edges = c;
        }
      }
      r.batch_ms = best;
    }
    if (what & PE_TUNE_BFS) {
      PE_REQUIRE(g->has_csr, PE_ERR_STATE, "BFS tuning needs out-edges (PE_GRAPH_BUILD_CSR)");
      std::vector<uint32_t> od(g->V), roots;
      pe_graph_out_degrees(g, od.data());
      for (uint32_t v = 0, step = std::max<uint32_t>(1, g->V / 97); v < g->V && roots.size() < 4; v += step)
        if (od[v] > 0) roots.push_back(v);
      const double cand[5] = {0.01, 0.02, 0.05, 0.1, 0.2};
      double best = 1e300;
      for (double c : cand) {
        const double ms = time_bfs(g, c, roots);
        if (ms < best) {
          best = ms;
          r.bfs_push_threshold = c;
        }
      }
      r.bfs_ms = best;
    }
    *out = r;
  });
}

// One line per graph shape: "<V> <nnz> <directed> <relabeled> tile task ratio".
int pe_autotune_save(pe_graph_t* g, const pe_tune_result_t* r, const char* path) {
  if (!g || !r || !path) return PE_ERR_INVALID;
  return pe::guarded(g->ctx, [&] {
    FILE* f = std::fopen(path, "a");
    PE_REQUIRE(f != nullptr, PE_ERR_INVALID, std::string("cannot open ") + path);
    std::fprintf(f, "%u %llu %d %d %u %u %.6g\n", g->V, (unsigned long long)g->nnz, g->directed ? 1 : 0,
                 g->relabeled ? 1 : 0, r->tile_items, r->batch_task_edges, r->bfs_push_threshold);
    std::fclose(f);
  });
}

int pe_autotune_load(pe_graph_t* g, const char* path, pe_tune_result_t* r, int* found) {
  if (!g || !path || !r || !found) return PE_ERR_INVALID;
  return pe::guarded(g->ctx, [&] {
    *found = 0;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    unsigned V, tile, task;
    unsigned long long nnz;
    int dir, rel;
    double ratio;
    while (std::fscanf(f, "%u %llu %d %d %u %u %lf", &V, &nnz, &dir, &rel, &tile, &task, &ratio) == 7) {
      if (V == g->V && nnz == g->nnz && dir == (g->directed ? 1 : 0) && rel == (g->relabeled ? 1 : 0)) {
        std::memset(r, 0, sizeof(*r));
        r->tile_items = tile;
        r->batch_task_edges = task;
        r->bfs_push_threshold = ratio;
        *found = 1;  // the last matching line wins
      }
    }
    std::fclose(f);
    if (*found) pe::graph_build_tiles(g, r->tile_items);
  });
}

}  // extern "C"
