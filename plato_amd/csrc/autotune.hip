// Autotuner: times the engine's tunables on the resident graph and keeps the
// fastest.  Tunables and the knob each one sets:
//   pull tile size      merge-path items per tile of the pull gather (1024/2048/4096);
//                       applied to the graph
//   binned layout       source chunk and destination bin widths of the binned pull
//                       (2^15 / 2^14 by default); the best layout is left built
//   batch task size     edges per wave task of the 64-column PPR block (256 to 16384); longer
//                       columns are split (the block's degree threshold)
//   batch pass plan     which pass kernel runs passes with 8, 16 and 32 covered columns:
//                       the packed block, or 4, 2 or 1 edges per row load of the wide one
//   BFS switch ratio    push iff active edges / E' < ratio (the reference's rule,
//                       0.05 by default; levels do not depend on it)
// Each candidate runs a few real passes; timing is by HIP events.  Every candidate set holds
// the default, so a tuned choice is never one that measured slower than the default.  A
// choice can be persisted with pe_autotune_save / pe_autotune_load keyed by the graph's shape.
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

// Pass time of a 64-column block with `active` columns in play (tolerance 0: none retires).
double time_batch(pe_graph_s* g, uint32_t task_edges, int iters, uint32_t active = 64,
                  const uint32_t* plan = nullptr) {
  pe_ppr_batch_opts_t bo{};
  bo.k = 64;
  bo.damping = 0.85;
  bo.tol = 0.0;  // keep every column active
  bo.max_iters = 1000000;
  bo.task_edges = task_edges;
  if (plan) {
    bo.plan_packed = plan[0];
    bo.plan_edges4 = plan[1];
    bo.plan_edges2 = plan[2];
  }
  pe_ppr_batch_t* b = nullptr;
  int rc = pe_ppr_batch_create(g, &bo, &b);
  if (rc != PE_OK) pe::fail(rc, g->ctx->last_error);
  for (uint32_t s = 0; s < active; ++s) {
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
      // the tiles belong to the merge-path gather: time it even when a binned layout is set
      const int layout = g->pull_layout;
      g->pull_layout = PE_PULL_MERGE_PATH;
      const uint32_t cand[3] = {2048, 1024, 4096};
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
      g->pull_layout = layout;
      r.pull_ms = best;
    }
    if (what & PE_TUNE_BINNED) {
      // (src_bits, dst_bits): the default first; 2^13-destination bins need half the LDS
      const uint32_t cand[4][2] = {{15, 14}, {14, 14}, {15, 13}, {14, 13}};
      double best = 1e300;
      for (const auto& c : cand) {
        pe_binned_opts_t o{c[0], c[1], 0, 0};
        const int rc = pe_graph_set_pull_layout(g, PE_PULL_BINNED, &o);
        if (rc != PE_OK) pe::fail(rc, g->ctx->last_error);
        const double ms = time_pull(g, 3);
        if (ms < best) {
          best = ms;
          r.binned_src_bits = c[0];
          r.binned_dst_bits = c[1];
        }
      }
      pe_binned_opts_t o{r.binned_src_bits, r.binned_dst_bits, 0, 0};
      const int rc = pe_graph_set_pull_layout(g, PE_PULL_BINNED, &o);
      if (rc != PE_OK) pe::fail(rc, g->ctx->last_error);
      r.binned_ms = best;
    }
    if (what & PE_TUNE_BATCH) {
      const uint32_t cand[4] = {4096, 256, 1024, 16384};
      double best = 1e300;
      for (uint32_t c : cand) {
        const double ms = time_batch(g, c, 2);
        if (ms < best) {
          best = ms;
          r.batch_task_edges = c;
        }
      }
      r.batch_ms = best;
      // pass plan, from the widest kernel down: at its own width (32, 16, 8 covered columns)
      // a kernel is kept if it is at least as fast as what runs that count without it (the
      // wider kernels as already decided); the default (every kernel on) is one of the two
      constexpr uint32_t OFF = PE_BATCH_PLAN_OFF;
      uint32_t plan[3] = {OFF, OFF, OFF};
      const uint32_t counts[3] = {8, 16, 32};
      for (int level = 2; level >= 0; --level) {
        uint32_t on[3] = {plan[0], plan[1], plan[2]}, off[3] = {plan[0], plan[1], plan[2]};
        on[level] = 0;
        off[level] = OFF;
        const double ms_on = time_batch(g, r.batch_task_edges, 2, counts[level], on);
        const double ms_off = time_batch(g, r.batch_task_edges, 2, counts[level], off);
        plan[level] = ms_on <= ms_off ? 0u : OFF;
        r.batch_plan_ms[level] = std::min(ms_on, ms_off);
      }
      r.batch_plan_packed = plan[0];
      r.batch_plan_edges4 = plan[1];
      r.batch_plan_edges2 = plan[2];
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

// One line per graph shape: "<V> <nnz> <directed> <relabeled> tile task ratio src_bits
// dst_bits plan_packed plan_edges4 plan_edges2" (lines with only the first seven fields, from
// before the binned and pass-plan choices, still load).
int pe_autotune_save(pe_graph_t* g, const pe_tune_result_t* r, const char* path) {
  if (!g || !r || !path) return PE_ERR_INVALID;
  return pe::guarded(g->ctx, [&] {
    FILE* f = std::fopen(path, "a");
    PE_REQUIRE(f != nullptr, PE_ERR_INVALID, std::string("cannot open ") + path);
    std::fprintf(f, "%u %llu %d %d %u %u %.6g %u %u %u %u %u\n", g->V, (unsigned long long)g->nnz,
                 g->directed ? 1 : 0, g->relabeled ? 1 : 0, r->tile_items, r->batch_task_edges, r->bfs_push_threshold,
                 r->binned_src_bits, r->binned_dst_bits, r->batch_plan_packed, r->batch_plan_edges4,
                 r->batch_plan_edges2);
    std::fclose(f);
  });
}

int pe_autotune_load(pe_graph_t* g, const char* path, pe_tune_result_t* r, int* found) {
  if (!g || !path || !r || !found) return PE_ERR_INVALID;
  return pe::guarded(g->ctx, [&] {
    *found = 0;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
      unsigned V = 0, tile = 0, task = 0, sb = 0, db = 0, pk = 0, p4 = 0, p2 = 0;
      unsigned long long nnz = 0;
      int dir = 0, rel = 0;
      double ratio = 0;
      const int n = std::sscanf(line, "%u %llu %d %d %u %u %lf %u %u %u %u %u", &V, &nnz, &dir, &rel, &tile, &task,
                                &ratio, &sb, &db, &pk, &p4, &p2);
      if (n != 7 && n != 12) continue;
      if (V == g->V && nnz == g->nnz && dir == (g->directed ? 1 : 0) && rel == (g->relabeled ? 1 : 0)) {
        std::memset(r, 0, sizeof(*r));
        r->tile_items = tile;
        r->batch_task_edges = task;
        r->bfs_push_threshold = ratio;
        r->binned_src_bits = sb;
        r->binned_dst_bits = db;
        r->batch_plan_packed = pk;
        r->batch_plan_edges4 = p4;
        r->batch_plan_edges2 = p2;
        *found = 1;  // the last matching line wins
      }
    }
    std::fclose(f);
    if (*found) {
      pe::graph_build_tiles(g, r->tile_items);
      if (g->pull_layout == PE_PULL_BINNED && r->binned_src_bits && r->binned_dst_bits) {
        pe_binned_opts_t o{r->binned_src_bits, r->binned_dst_bits, 0, 0};
        const int rc = pe_graph_set_pull_layout(g, PE_PULL_BINNED, &o);
        if (rc != PE_OK) pe::fail(rc, g->ctx->last_error);
      }
    }
  });
}

}  // extern "C"
