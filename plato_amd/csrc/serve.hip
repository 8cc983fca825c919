// PPR serving loop: continuous batching of queries into 64-column rank blocks.
//
// Requests (a seed set each) wait in a FIFO until their arrival time; the loop
// admits them into free column slots of the resident blocks, opening another
// block when every slot is taken and free HBM (minus a reserve) still holds one.
// Every pass enqueues a few iterations of every block with work, back to back on
// the one stream, then reads the slot states: finished columns are read out
// (top-k, rank desc / id asc as trustrank.hpp:109-112 orders) and their slots
// are refilled at once, so queries that arrive while others run share the
// device with them.  A column's arithmetic is independent of its neighbours,
// so a query's result does not depend on what else is in flight.
//
// The query semantics are those of pe_ppr_batch (restart mass (1-d)/|S| on each
// seed, trustrank.hpp:197-250; per-column delta < tol then one final iteration,
// pagerank.hpp:137-139).
#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "pe_internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

struct request {
  uint64_t id;
  double arrival;
  std::vector<uint32_t> seeds;
  double admit = -1, done = -1;
  uint32_t iterations = 0, batch = 0, slot = 0;
  std::vector<uint32_t> top_ids;
  std::vector<float> top_vals;
};

}  // namespace

struct pe_server_s {
  pe_graph_s* g = nullptr;
  pe_server_opts_t opts{};
  std::vector<pe_ppr_batch_t*> batches;
  std::vector<std::vector<int64_t>> slot_req;  // per batch, per slot: request index or -1
  std::vector<request> reqs;
  std::deque<size_t> waiting;                   // arrived, not yet admitted (FIFO)
  std::deque<size_t> future;                    // not yet arrived, ordered by arrival
  std::vector<size_t> completed;                // finished, not yet taken
  uint64_t batch_bytes = 0;
  clk::time_point t0;
  bool started = false;
  uint64_t passes = 0;

  ~pe_server_s() {
    for (auto* b : batches) pe_ppr_batch_destroy(b);
  }

  double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }

  bool open_batch() {
    if (batches.size() >= opts.max_batches) return false;
    uint64_t free_b = 0, total_b = 0, ctx_b = 0;
    if (pe_ctx_mem_info(g->ctx, &free_b, &total_b, &ctx_b) != PE_OK) return false;
    if (batch_bytes && free_b < batch_bytes + opts.hbm_reserve_bytes) return false;  // admission by free HBM
    pe_ppr_batch_opts_t bo{};
    bo.k = 64;
    bo.damping = opts.damping;
    bo.tol = opts.tol;
    bo.max_iters = opts.max_iters;
    bo.task_edges = 0;
    pe_ppr_batch_t* b = nullptr;
    const int rc = pe_ppr_batch_create(g, &bo, &b);
    if (rc == PE_ERR_OOM) return false;
    if (rc != PE_OK) pe::fail(rc, std::string("batch create: ") + g->ctx->last_error);
    uint64_t by = 0;
    pe_ppr_batch_bytes(b, &by);
    batch_bytes = by;
    batches.push_back(b);
    slot_req.emplace_back(64, -1);
    return true;
  }

  void admit_waiting() {
    while (!waiting.empty()) {
      bool placed = false;
      for (size_t bi = 0; bi < batches.size() && !placed; ++bi)
        for (uint32_t s = 0; s < 64 && !placed; ++s)
          if (slot_req[bi][s] < 0) {
            const size_t r = waiting.front();
            const int rc = pe_ppr_batch_admit(batches[bi], s, reqs[r].seeds.data(), (uint32_t)reqs[r].seeds.size());
            if (rc != PE_OK) pe::fail(rc, std::string("admit: ") + g->ctx->last_error);
            slot_req[bi][s] = (int64_t)r;
            reqs[r].admit = now();
            reqs[r].batch = (uint32_t)bi;
            reqs[r].slot = s;
            waiting.pop_front();
            placed = true;
          }
      if (!placed && !open_batch()) break;
    }
  }

  // One pass: arrivals, admission, iterations of every busy block, harvest.
  uint32_t pump() {
    if (!started) {
      t0 = clk::now();
      started = true;
    }
    const double t = now();
    while (!future.empty() && reqs[future.front()].arrival <= t) {
      waiting.push_back(future.front());
      future.pop_front();
    }
    admit_waiting();
    std::vector<size_t> busy;
    for (size_t bi = 0; bi < batches.size(); ++bi)
      for (int64_t r : slot_req[bi])
        if (r >= 0) { busy.push_back(bi); break; }
    if (busy.empty()) return 0;
    for (size_t bi : busy) {
      const int rc = pe_ppr_batch_step(batches[bi], opts.iters_per_pass);
      if (rc != PE_OK) pe::fail(rc, std::string("step: ") + g->ctx->last_error);
    }
    ++passes;
    uint32_t finished = 0;
    for (size_t bi : busy) {
      uint32_t st[64], it[64];
      if (pe_ppr_batch_poll(batches[bi], st, it, nullptr) != PE_OK) pe::fail(PE_ERR_INTERNAL, "poll failed");
      std::vector<uint32_t> done;
      for (uint32_t s = 0; s < 64; ++s)
        if (slot_req[bi][s] >= 0 && st[s] == 4) done.push_back(s);
      if (done.empty()) continue;
      const uint32_t k = opts.topk;
      std::vector<uint32_t> ids((size_t)done.size() * k);
      std::vector<float> vals((size_t)done.size() * k);
      if (k) {
        const int rc = pe_ppr_batch_topk(batches[bi], done.data(), (uint32_t)done.size(), k, ids.data(), vals.data());
        if (rc != PE_OK) pe::fail(rc, std::string("topk: ") + g->ctx->last_error);
      }
      const double td = now();
      for (size_t i = 0; i < done.size(); ++i) {
        const uint32_t s = done[i];
        request& rq = reqs[(size_t)slot_req[bi][s]];
        rq.done = td;
        rq.iterations = it[s];
        rq.top_ids.assign(ids.begin() + i * k, ids.begin() + (i + 1) * k);
        rq.top_vals.assign(vals.begin() + i * k, vals.begin() + (i + 1) * k);
        completed.push_back((size_t)slot_req[bi][s]);
        pe_ppr_batch_release(batches[bi], s);
        slot_req[bi][s] = -1;
        ++finished;
      }
    }
    admit_waiting();
    return finished;
  }

  bool idle() const {
    if (!waiting.empty() || !future.empty()) return false;
    for (const auto& v : slot_req)
      for (int64_t r : v)
        if (r >= 0) return false;
    return true;
  }
};

namespace {

void fill_stats(const request& r, pe_request_stats_t* s) {
  s->id = r.id;
  s->arrival_s = r.arrival;
  s->admit_s = r.admit;
  s->done_s = r.done;
  s->iterations = r.iterations;
  s->batch = r.batch;
  s->slot = r.slot;
}

}  // namespace

extern "C" {

int pe_server_create(pe_graph_t* g, const pe_server_opts_t* opts, pe_server_t** out) {
  if (!g || !opts || !out) return PE_ERR_INVALID;
  *out = nullptr;
  return pe::guarded(g->ctx, [&] {
    PE_REQUIRE(opts->max_batches >= 1, PE_ERR_INVALID, "max_batches must be >= 1");
    PE_REQUIRE(opts->iters_per_pass >= 1, PE_ERR_INVALID, "iters_per_pass must be >= 1");
    PE_REQUIRE(opts->topk <= 4096, PE_ERR_INVALID, "topk too large");
    auto s = std::make_unique<pe_server_s>();
    s->g = g;
    s->opts = *opts;
    PE_REQUIRE(s->open_batch(), PE_ERR_OOM, "not enough free HBM for one 64-column block");
    *out = s.release();
  });
}

int pe_server_destroy(pe_server_t* s) {
  if (!s) return PE_ERR_INVALID;
  delete s;
  return PE_OK;
}

int pe_server_submit(pe_server_t* s, uint64_t id, const uint32_t* seeds, uint32_t n, double arrival_s) {
  if (!s || !seeds || n == 0) return PE_ERR_INVALID;
  return pe::guarded(s->g->ctx, [&] {
    PE_REQUIRE(n <= 4096, PE_ERR_INVALID, "too many seeds");
    for (uint32_t i = 0; i < n; ++i) PE_REQUIRE(seeds[i] < s->g->V, PE_ERR_RANGE, "seed vertex >= V");
    request r;
    r.id = id;
    r.arrival = arrival_s;
    r.seeds.assign(seeds, seeds + n);
    s->reqs.push_back(std::move(r));
    const size_t idx = s->reqs.size() - 1;
    // keep future arrivals ordered
    auto it = s->future.begin();
    while (it != s->future.end() && s->reqs[*it].arrival <= arrival_s) ++it;
    s->future.insert(it, idx);
  });
}

int pe_server_pump(pe_server_t* s, uint32_t* completed) {
  if (!s) return PE_ERR_INVALID;
  return pe::guarded(s->g->ctx, [&] {
    const uint32_t c = s->pump();
    if (completed) *completed = c;
  });
}

int pe_server_take(pe_server_t* s, uint32_t max, pe_request_stats_t* stats, uint32_t* topk_ids,
                   float* topk_vals, uint32_t* taken) {
  if (!s || !taken) return PE_ERR_INVALID;
  return pe::guarded(s->g->ctx, [&] {
    const uint32_t n = (uint32_t)std::min<size_t>(max, s->completed.size());
    const uint32_t k = s->opts.topk;
    for (uint32_t i = 0; i < n; ++i) {
      const request& r = s->reqs[s->completed[i]];
      if (stats) fill_stats(r, &stats[i]);
      if (k && topk_ids) std::copy(r.top_ids.begin(), r.top_ids.end(), topk_ids + (size_t)i * k);
      if (k && topk_vals) std::copy(r.top_vals.begin(), r.top_vals.end(), topk_vals + (size_t)i * k);
    }
    s->completed.erase(s->completed.begin(), s->completed.begin() + n);
    *taken = n;
  });
}

int pe_server_info(pe_server_t* s, uint32_t* batches, uint32_t* in_flight, uint32_t* waiting,
                   uint64_t* passes) {
  if (!s) return PE_ERR_INVALID;
  uint32_t f = 0;
  for (const auto& v : s->slot_req)
    for (int64_t r : v) f += r >= 0;
  if (batches) *batches = (uint32_t)s->batches.size();
  if (in_flight) *in_flight = f;
  if (waiting) *waiting = (uint32_t)(s->waiting.size() + s->future.size());
  if (passes) *passes = s->passes;
  return PE_OK;
}

int pe_server_run_trace(pe_server_t* s, uint32_t n, const uint64_t* ids, const double* arrivals,
                        const uint64_t* seed_offsets, const uint32_t* seeds, pe_request_stats_t* stats,
                        uint32_t* topk_ids, float* topk_vals, double* wall_s) {
  if (!s || (n && (!ids || !arrivals || !seed_offsets || !seeds || !stats))) return PE_ERR_INVALID;
  return pe::guarded(s->g->ctx, [&] {
    PE_REQUIRE(s->reqs.empty(), PE_ERR_STATE, "run_trace needs a fresh server");
    for (uint32_t i = 0; i < n; ++i) {
      const uint64_t b = seed_offsets[i], e = seed_offsets[i + 1];
      PE_REQUIRE(e > b, PE_ERR_INVALID, "empty seed list");
      const int rc = pe_server_submit(s, ids[i], seeds + b, (uint32_t)(e - b), arrivals[i]);
      if (rc != PE_OK) pe::fail(rc, s->g->ctx->last_error);
    }
    s->t0 = clk::now();
    s->started = true;
    while (!s->idle()) {
      const uint32_t c = s->pump();
      (void)c;
      bool busy = false;
      for (const auto& v : s->slot_req)
        for (int64_t r : v) busy |= r >= 0;
      if (!busy && s->waiting.empty() && !s->future.empty()) {
        // nothing on the device: sleep until the next arrival
        const double dt = s->reqs[s->future.front()].arrival - s->now();
        if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(dt, 0.05)));
      }
    }
    if (wall_s) *wall_s = s->now();
    // results in request order
    std::vector<size_t> order(s->reqs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    const uint32_t k = s->opts.topk;
    for (uint32_t i = 0; i < n; ++i) {
      const request& r = s->reqs[i];
      fill_stats(r, &stats[i]);
      if (k && topk_ids) std::copy(r.top_ids.begin(), r.top_ids.end(), topk_ids + (size_t)i * k);
      if (k && topk_vals) std::copy(r.top_vals.begin(), r.top_vals.end(), topk_vals + (size_t)i * k);
    }
    s->completed.clear();
  });
}

}  // extern "C"
