// Internal declarations shared by the HIP translation units of libplato_hip.so.
// Not part of the C-ABI (include/plato_hip.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "plato_hip.h"

namespace pe {

// ---- errors ------------------------------------------------------------------
// C-ABI functions never throw: internal code throws pe::error, the boundary
// catches it and turns it into a status code plus the context's last-error text.
struct error : std::runtime_error {
  int code;
  error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw error(code, msg); }

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PE_ERR_OOM : PE_ERR_HIP;
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    throw error(code, buf);
  }
}
#define PE_HIP(x) ::pe::hip_check((x), #x, __FILE__, __LINE__)
#define PE_REQUIRE(cond, code, msg) \
  do { if (!(cond)) ::pe::fail((code), (msg)); } while (0)

// ---- context -------------------------------------------------------------------
struct kernel_timer {
  std::vector<hipEvent_t> begin, end;  // recorded pairs, drained by collect()
  double total_ms = 0.0;
  uint64_t launches = 0;
};

}  // namespace pe

struct pe_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  uint64_t allocated = 0;       // bytes currently held by this context
  uint64_t peak_allocated = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  int timing = 0;               // per-kernel event timing on/off
  pe::kernel_timer timers[PE_KT_COUNT];
  int num_cus = 0;
  // device scalars used by reductions; pinned host mirror for cheap readback
  void* host_pinned = nullptr;  // 4 KiB
  // coherent pinned host words that a one-thread kernel writes at the end of a step and
  // the host polls: a step's counters reach the host without a copy or a stream wait
  unsigned long long* host_mail = nullptr;  // 4 KiB; [0] = sequence number, [1..] = words
  unsigned long long mail_seq = 0;
  // this device's static scratch array of pe_pull_sum_noalloc (its device address)
  void* noalloc_scratch = nullptr;
};

namespace pe {

// Copies n (at most 64) device words to the host and zeroes nclear device words after
// the work queued on the stream, then waits for them: returns the host copy, valid until
// the next call on this context.
const unsigned long long* mail_post(pe_ctx_s* ctx, const unsigned long long* src, uint32_t n,
                                    unsigned long long* clear, uint32_t nclear);

// Allocation through the context keeps a ledger (admission control reads it).
void* dmalloc(pe_ctx_s* ctx, size_t bytes);
void dfree(pe_ctx_s* ctx, void* p, size_t bytes);

template <typename T>
struct dbuf {
  pe_ctx_s* ctx = nullptr;
  T* p = nullptr;
  size_t n = 0;
  dbuf() = default;
  dbuf(pe_ctx_s* c, size_t count) { alloc(c, count); }
  void alloc(pe_ctx_s* c, size_t count) {
    release();
    ctx = c;
    n = count;
    p = count ? static_cast<T*>(dmalloc(c, count * sizeof(T))) : nullptr;
  }
  void release() {
    if (p) dfree(ctx, p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
  ~dbuf() { release(); }
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  dbuf(dbuf&& o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  dbuf& operator=(dbuf&& o) noexcept {
    if (this != &o) { release(); ctx = o.ctx; p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Kernel timing: brackets a launch with events on the context stream when
// timing is enabled.  Used for the per-kernel average launch duration.
struct ktime_scope {
  pe_ctx_s* ctx;
  int which;
  hipEvent_t e1 = nullptr;
  ktime_scope(pe_ctx_s* c, int w);
  ~ktime_scope();
};

// A captured stream segment and its executable.  Capture always ends: if anything
// throws between begin() and end(), the destructor ends the capture and drops the
// partial graph, so the stream is not left in capture mode.  An executable that has
// been launched is destroyed only after its stream has drained, so no launch of it is
// still in flight (or still being traced) when its memory goes.
struct stream_graph {
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool capturing = false, launched = false;

  void begin(hipStream_t st) {
    reset();
    s = st;
    PE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    capturing = true;
  }
  void end() {
    capturing = false;
    PE_HIP(hipStreamEndCapture(s, &graph));
    PE_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  void launch() {
    launched = true;
    PE_HIP(hipGraphLaunch(exec, s));
  }
  void reset() {
    if (capturing) {
      hipGraph_t partial = nullptr;
      if (hipStreamEndCapture(s, &partial) == hipSuccess && partial) (void)hipGraphDestroy(partial);
      (void)hipGetLastError();
      capturing = false;
    }
    if (launched) (void)hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    launched = false;
  }
  ~stream_graph() { reset(); }
};

inline unsigned grid_for(uint64_t work, unsigned per_block, unsigned cap = 1u << 30) {
  uint64_t g = (work + per_block - 1) / per_block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// Persistent grid: enough resident blocks to cover every CU several times.
inline unsigned persistent_grid(pe_ctx_s* ctx, uint64_t tiles, unsigned blocks_per_cu) {
  uint64_t cap = static_cast<uint64_t>(ctx->num_cus > 0 ? ctx->num_cus : 256) * blocks_per_cu;
  if (tiles < cap) cap = tiles;
  return cap ? static_cast<unsigned>(cap) : 1u;
}

// A wave-uniform value marked as such: addresses derived from it are computed in
// scalar registers and their loads go through the scalar path (one load per wave).
__device__ __forceinline__ uint64_t wave_uniform(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Index of the calling wave in the grid, and the number of waves (both wave-uniform).
__device__ __forceinline__ uint64_t grid_wave_id() {
  return wave_uniform((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6);
}
__device__ __forceinline__ uint64_t grid_waves() { return ((uint64_t)gridDim.x * blockDim.x) >> 6; }

// Adds a and b over the block into *da and *db with one atomic each from thread 0 (none
// for a zero sum).  Every thread of the block must call it; blockDim.x is a multiple of
// 64, at most 1024.  Counters that every wave of a large grid bumps are serialized in
// the memory-side atomic unit, so kernels sum per block first.
__device__ __forceinline__ void block_add2(unsigned long long a, unsigned long long b, unsigned long long* da,
                                           unsigned long long* db) {
  __shared__ unsigned long long part[2][16];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][wave] = a;
    part[1][wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sa = 0, sb = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      sa += part[0][w];
      sb += part[1][w];
    }
    if (sa) atomicAdd(da, sa);
    if (sb) atomicAdd(db, sb);
  }
}

// three sums, as block_add2
__device__ __forceinline__ void block_add3(unsigned long long a, unsigned long long b, unsigned long long c,
                                           unsigned long long* da, unsigned long long* db, unsigned long long* dc) {
  __shared__ unsigned long long part3[3][16];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
    c += __shfl_down(c, off, 64);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part3[0][wave] = a;
    part3[1][wave] = b;
    part3[2][wave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sa = 0, sb = 0, sc = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      sa += part3[0][w];
      sb += part3[1][w];
      sc += part3[2][w];
    }
    if (sa) atomicAdd(da, sa);
    if (sb) atomicAdd(db, sb);
    if (sc) atomicAdd(dc, sc);
  }
}

}  // namespace pe

// ---- binned pull layout (pull_binned.hip) -------------------------------------------
namespace pe {
struct pb_unit1 {  // scatter work: run records [r0, r1) of one source chunk
  uint64_t r0, r1;
  uint32_t chunk, pad;
};
struct pb_unit2 {  // gather work: entries [e0, e1) of one bin; slot ~0u = the whole bin
  uint64_t e0, e1;
  uint32_t bin, slot;
};
struct pb_layout {
  uint32_t src_bits = 15, dst_bits = 14, run_cap = 256, unit_log2 = 20;
  uint32_t nchunks = 0, nbins = 0;
  uint64_t nrecords = 0;
  uint64_t nupdates = 0;                // distinct (source, bin) pairs
  uint32_t n1 = 0, n2 = 0, nsplit = 0;
  bool weighted = false;
  dbuf<uint16_t> src_upd;               // [nupdates] source (in its chunk) of each update
  dbuf<uint16_t> dst_local;             // [nnz, padded to 512] bin order: destination in its bin, bit 15 =
                                        // first entry of a new update
  dbuf<uint32_t> gbase;                 // updates before each aligned gather step of 512 entries
  dbuf<float> w;                        // [nnz] bin order, weighted graphs
  dbuf<float> values;                   // [nupdates] contributions of the current iteration
  dbuf<uint64_t> rec_start;             // [nrecords] chunk-major run records (update space)
  dbuf<uint32_t> rec_len;
  dbuf<pb_unit1> units1;
  dbuf<pb_unit2> units2;
  dbuf<uint32_t> split_bin;             // [nsplit] bins summed by several units
  dbuf<unsigned long long> acc;         // [nsplit << dst_bits] fixed-point sums of split bins
  dbuf<uint32_t> work;                  // [4] unit queues: scatter next / finished, gather next /
                                        // finished (zero between launches)
  uint64_t bytes() const {
    return src_upd.bytes() + dst_local.bytes() + gbase.bytes() + w.bytes() + values.bytes() +
           rec_start.bytes() + rec_len.bytes() + units1.bytes() + units2.bytes() + split_bin.bytes() +
           acc.bytes() + work.bytes();
  }
};
}  // namespace pe

// ---- graph ----------------------------------------------------------------------
struct pe_graph_s {
  pe_ctx_s* ctx = nullptr;
  uint32_t V = 0;           // vertex id space [0, V)
  uint64_t edges = 0;       // input edges (Plato graph_info_t::edges_)
  uint64_t nnz = 0;         // CSC entries
  bool directed = true;
  bool has_weights = false;
  bool has_csr = false;
  bool csr_is_csc = false;  // symmetrized graphs share one structure
  uint32_t max_v_i = 0;
  uint32_t vertices = 0;
  uint32_t max_in_degree = 0;

  pe::dbuf<uint64_t> col_ptr;   // [V+1]
  pe::dbuf<uint32_t> row_idx;   // [nnz], ascending source id inside each column
  pe::dbuf<float> wn;           // [nnz] weight / sum of the source's out-weights
  pe::dbuf<uint32_t> od;        // [V] out-degree counted from the in-edges
  pe::dbuf<uint32_t> first_in;  // [V] lowest in-neighbour (0xFFFFFFFF: none); built by the first BFS
  pe::dbuf<uint64_t> row_ptr;   // [V+1] CSR (out-edges)
  pe::dbuf<uint32_t> col_idx;   // [nnz]

  // merge-path tiling of the pull gather (rows + edges, fixed items per tile)
  uint32_t tile_items = 0;
  uint64_t num_tiles = 0;
  pe::dbuf<uint32_t> tile_row;   // [T+1] first row of each tile
  pe::dbuf<uint64_t> tile_edge;  // [T+1] first edge of each tile
  // rows whose edges cross a tile boundary: finished by the fix-up kernel
  uint64_t num_fix = 0;
  pe::dbuf<uint32_t> fix_row;    // [F]
  pe::dbuf<uint32_t> fix_t0;     // [F] first tile holding edges of the row
  pe::dbuf<uint32_t> fix_t1;     // [F] tile holding the row's end
  uint64_t num_fix_long = 0;     // fix rows [num_fix - num_fix_long, num_fix) span > 8 tiles

  // Graphs with one id_space number number their vertices alike (same V, same internal
  // ids): a graph and its reverse built with PE_GRAPH_SAME_IDS.  Device vectors, bitmaps
  // and search states of one are valid for the other.
  uint64_t id_space = next_id_space();
  static uint64_t next_id_space() {
    static std::atomic<uint64_t> n{1};
    return n.fetch_add(1);
  }

  // degree relabeling: internal id = rank_of[input id], input id = perm[internal id]
  bool relabeled = false;
  pe::dbuf<uint32_t> perm;       // [V]
  pe::dbuf<uint32_t> rank_of;    // [V]

  // ids compacted at ingestion (pe_graph_create_csv with PE_GRAPH_ENCODE_IDS):
  // original id of every dense input id
  bool encoded = false;
  pe::dbuf<uint64_t> vid_decode;  // [V]

  // pull layout of the FP32 rank iterations
  int pull_layout = PE_PULL_MERGE_PATH;
  std::unique_ptr<pe::pb_layout> pb;

  const uint64_t* csr_ptr() const { return csr_is_csc ? col_ptr.p : row_ptr.p; }
  const uint32_t* csr_idx() const { return csr_is_csc ? row_idx.p : col_idx.p; }
  uint64_t device_bytes() const {
    return col_ptr.bytes() + row_idx.bytes() + wn.bytes() + od.bytes() + row_ptr.bytes() +
           col_idx.bytes() + perm.bytes() + rank_of.bytes() + tile_row.bytes() + tile_edge.bytes() + fix_row.bytes() +
           fix_t0.bytes() + fix_t1.bytes() + vid_decode.bytes() + (pb ? pb->bytes() : 0);
  }
};

namespace pe {
// graph.hip
void graph_build_tiles(pe_graph_s* g, uint32_t tile_items);
// stable radix sort of packed keys over bits [0, end_bit), optional float payload
void sort_keys(pe_ctx_s* ctx, dbuf<uint64_t>& keys, dbuf<float>* vals, uint64_t n, int end_bit);
// builds the CSC (and CSR, weights, relabeling per flags) from device edge arrays
// (g->V and g->edges set by the caller)
void build_from_edges(pe_graph_s* g, dbuf<uint32_t>& ds, dbuf<uint32_t>& dd, dbuf<float>& dw, uint32_t flags);
// pull_binned.hip: builds g->pb (no-op if built with the same options)
void binned_build(pe_graph_s* g, const pe_binned_opts_t* opts);
// relabeled graphs: input ids -> internal ids (host arrays), and V-long device
// arrays downloaded in input id order (elem = 4 or 8 bytes)
void to_internal_ids(pe_graph_s* g, const uint32_t* host_ids, uint32_t n, uint32_t* host_out);
void download_external(pe_graph_s* g, const void* dev, void* host, size_t elem);
// sum of out-degrees over a bitmap's vertices (rank.hip)
uint64_t active_edges_dev(pe_graph_s* g, const uint64_t* words, uint64_t nwords);
// bitmap -> queue of its set ids below V, appended at *count (bfs.hip)
void frontier_compact(pe_ctx_s* ctx, const uint64_t* words, uint64_t nwords, uint32_t V, uint32_t* queue,
                      unsigned long long* count);
// Hubs of an adjacency (ptr[V+1]): the vertices with more than `deg` entries, and their
// entries cut into chunks of `chunk` (hubs.hip).  A kernel that would give a whole
// adjacency to one wave leaves the hubs to a pass over their chunks instead, spread over
// every wave, so no wave walks a hub alone.  Built on the device (counts stay there).
struct hub_table {
  uint64_t deg = 0, chunk = 0, max_hubs = 0;
  dbuf<uint32_t> hubs;           // hub ids (ascending within each compaction block's slice)
  dbuf<uint64_t> start;          // [max_hubs + 1]: first chunk of hub h; start[*count] = all chunks
  dbuf<unsigned long long> count;  // [1] number of hubs
  dbuf<uint64_t> bits;           // scratch: hub bitmap
};
void hub_table_build(pe_ctx_s* ctx, const uint64_t* ptr, uint32_t V, uint64_t nnz, uint64_t deg, uint64_t chunk,
                     hub_table& t);
// the hub (its id, and its index h in the table) and edge range [k0, k1) of chunk c
// (c below start[n], n = *t.count > 0)
__device__ __forceinline__ uint32_t hub_chunk(const uint64_t* __restrict__ ptr, const uint32_t* __restrict__ hubs,
                                              const uint64_t* __restrict__ start, uint64_t n, uint64_t chunk,
                                              uint64_t c, uint64_t* k0, uint64_t* k1, uint64_t* h = nullptr) {
  uint64_t lo = 0, hi = n - 1;  // last h with start[h] <= c
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (start[mid] <= c) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t v = hubs[lo];
  const uint64_t e = ptr[v + 1];
  if (h) *h = lo;
  *k0 = ptr[v] + (c - start[lo]) * chunk;
  *k1 = *k0 + chunk < e ? *k0 + chunk : e;
  return v;
}
// vectors
struct vec_view {
  int dtype;
  uint64_t n;
  void* p;
};
}  // namespace pe

struct pe_vec_s {
  pe_ctx_s* ctx = nullptr;
  int dtype = PE_F32;
  uint64_t n = 0;
  void* p = nullptr;
  size_t bytes = 0;
};

struct pe_bitmap_s {
  pe_ctx_s* ctx = nullptr;
  uint64_t bits = 0;
  uint64_t words = 0;  // (bits >> 6) + 1, as bitmap.hpp:43-108
  uint64_t* p = nullptr;
};

// ---- C-ABI boundary helper --------------------------------------------------------
namespace pe {
template <typename F>
int guarded(pe_ctx_s* ctx, F&& f) {
  try {
    f();
    return PE_OK;
  } catch (const error& e) {
    if (ctx) ctx->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->last_error = "host allocation failed";
    return PE_ERR_OOM;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return PE_ERR_INTERNAL;
  } catch (...) {
    if (ctx) ctx->last_error = "unknown exception";
    return PE_ERR_INTERNAL;
  }
}
}  // namespace pe
