/*
 * plato_hip.h -- C-ABI of libplato_hip.so, the MI355X (gfx950) dual-mode
 * push/pull edge engine.
 *
 * Conventions
 *  - Every function returns an int status (PE_OK = 0).  Nothing throws or
 *    aborts across this boundary; pe_ctx_last_error() holds the message of the
 *    last failure on a context.
 *  - Host pointers are borrowed for the duration of the call only.  Device
 *    objects (graphs, vectors, bitmaps, batches, servers) are owned by their
 *    handle and freed by the matching *_destroy.
 *  - Work is issued on the context's stream; functions that return results to
 *    host memory synchronize that stream before returning.
 *  - A context is driven by one host thread at a time.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository root).  The reference is a header-only C++
 * template engine whose operators are host lambdas; here the operators are a
 * closed set selected by the entry point (gather-sum, weighted gather-sum,
 * restart-apply, BFS push/pull claim, PPR push-add).
 */
#ifndef PLATO_HIP_H
#define PLATO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define PE_OK 0
#define PE_ERR_INVALID 1   /* bad argument */
#define PE_ERR_HIP 2       /* HIP runtime error */
#define PE_ERR_OOM 3       /* device or host allocation failed */
#define PE_ERR_STATE 4     /* object in the wrong state for the call */
#define PE_ERR_RANGE 5     /* index or size out of range */
#define PE_ERR_INTERNAL 6  /* unexpected internal failure */

/* ---- dtypes of engine state vectors -------------------------------------- */
#define PE_F32 0
#define PE_F64 1
#define PE_U32 2

/* ---- per-kernel timers (pe_ctx_kernel_time) ------------------------------- */
#define PE_KT_PULL_TILE 0   /* pull gather + fused apply (merge-path tiles) */
#define PE_KT_PULL_FIXUP 1  /* rows crossing tile boundaries */
#define PE_KT_BFS_PUSH 2    /* top-down BFS step */
#define PE_KT_BFS_PULL 3    /* bottom-up BFS step */
#define PE_KT_PPR_BLOCK 4   /* batched PPR k-column gather */
#define PE_KT_PULL_BIN_SCATTER 5 /* binned pull: contributions into destination-bin runs */
#define PE_KT_PULL_BIN_GATHER 6  /* binned pull: exact per-bin sums with the apply fused */
#define PE_KT_COUNT 8

typedef struct pe_ctx_s pe_ctx_t;
typedef struct pe_graph_s pe_graph_t;
typedef struct pe_vec_s pe_vec_t;
typedef struct pe_bitmap_s pe_bitmap_t;

/* ===========================================================================
 * Context.  Replaces cluster_info_t (plato/graph/base.hpp:158-233) for one
 * GPU: device, stream, allocation ledger.  Destroy every graph, vector,
 * bitmap, batch and server of a context before the context itself.
 * ========================================================================= */
int pe_ctx_create(int device, pe_ctx_t** out);
int pe_ctx_destroy(pe_ctx_t* ctx);
const char* pe_ctx_last_error(const pe_ctx_t* ctx);
int pe_ctx_sync(pe_ctx_t* ctx);
/* free/total device memory and the bytes this context holds */
int pe_ctx_mem_info(pe_ctx_t* ctx, uint64_t* free_bytes, uint64_t* total_bytes,
                    uint64_t* ctx_bytes);
/* event timer on the context stream (stop synchronizes) */
int pe_ctx_timer_start(pe_ctx_t* ctx);
int pe_ctx_timer_stop(pe_ctx_t* ctx, double* ms);
/* per-kernel timing: when enabled, launches of the kernels named by PE_KT_*
 * are bracketed by events; pe_ctx_kernel_time drains them (synchronizes) and
 * returns the accumulated milliseconds and launch count, then resets. */
int pe_ctx_set_kernel_timing(pe_ctx_t* ctx, int enable);
int pe_ctx_kernel_time(pe_ctx_t* ctx, int which, double* total_ms, uint64_t* launches);
/* multiprocessor count of the device */
int pe_ctx_num_cus(pe_ctx_t* ctx, int* cus);
/* Streaming rates of the device (no counterpart in the reference: the ceiling that
 * roofline figures are read against).  Times `reps` passes of a read-only kernel
 * over `bytes` of device memory and of a copy of `bytes` into a second buffer
 * (16-byte loads, 8 blocks of 256 threads per CU), after one untimed pass of each;
 * read_gbs = bytes / time, copy_gbs = 2 * bytes / time, in 1e9 bytes per second.
 * Allocates and frees 2 * bytes; bytes is a positive multiple of 16. */
int pe_ctx_stream_rates(pe_ctx_t* ctx, uint64_t bytes, int reps, double* read_gbs, double* copy_gbs);

/* ===========================================================================
 * Graph.  Replaces dcsc_t (plato/graph/structure/dcsc.hpp:58-185) and
 * bcsr_t (plato/graph/structure/bcsr.hpp:67-208) with a device CSC over all V
 * ids (uint64 col_ptr[V+1], uint32 row_idx[nnz], ascending source id inside a
 * column) and an optional CSR, built on the device, and the factories
 * create_dcsc_seqs_from_path / create_dualmode_seq_from_path
 * (plato/graph/structure.hpp:473-555, 987-1086) for in-memory edge lists.
 * ========================================================================= */
#define PE_GRAPH_SYMMETRIZE 1u /* add v->u for every u->v (is_directed_ = false) */
#define PE_GRAPH_BUILD_CSR 2u  /* also build out-edges (needed by push mode) */
#define PE_GRAPH_WEIGHTED 4u   /* keep per-edge weights, normalised per source */
#define PE_GRAPH_RELABEL_DEGREE 8u /* renumber ids by out-degree (desc) inside the engine;
                                      algorithm entry points map seeds, roots and results, so
                                      callers keep input ids; structure downloads are internal */

typedef struct {
  uint32_t V;             /* id space [0, V) */
  uint32_t max_v_i;       /* largest id with an edge (graph_info_t::max_v_i_) */
  uint32_t vertices;      /* ids with at least one edge (graph_info_t::vertices_) */
  uint32_t max_in_degree;
  uint64_t edges;         /* input edges (graph_info_t::edges_) */
  uint64_t nnz;           /* CSC entries: edges, or 2*edges when symmetrized */
  int32_t is_directed;    /* graph_info_t::is_directed_ */
  int32_t has_csr;
  int32_t has_weights;
  int32_t relabeled;
  uint32_t tile_items;    /* merge-path items per pull tile */
  uint64_t num_tiles;
  uint64_t num_fixup_rows;
  uint64_t device_bytes;
  int32_t pull_layout;    /* PE_PULL_MERGE_PATH or PE_PULL_BINNED */
  uint32_t binned_runs;   /* run records of the binned layout (0 if not built) */
  uint64_t binned_bytes;  /* device bytes held by the binned layout */
} pe_graph_info_t;

/* Build from a host edge list (src[i] -> dst[i], optional weights w[i]).
 * Ids must be < V.  Self-loops and duplicates are kept (dcsc.hpp:350-394). */
int pe_graph_create(pe_ctx_t* ctx, const uint32_t* src, const uint32_t* dst, const float* w,
                    uint64_t num_edges, uint32_t V, uint32_t flags, pe_graph_t** out);
/* Build a seeded RMAT graph directly on the device: 2^scale ids,
 * edge_factor * 2^scale edges, quadrant probabilities a, b, c (d = 1-a-b-c).
 * Counter-based: edge e depends only on (seed, e); weights (PE_GRAPH_WEIGHTED)
 * are uniform in (0, 1] on a 2^-24 grid. */
int pe_graph_create_rmat(pe_ctx_t* ctx, uint32_t scale, uint32_t edge_factor, uint64_t seed,
                         double a, double b, double c, uint32_t flags, pe_graph_t** out);
/* Generate RMAT edges [e0, e0+n) into host arrays (same generator). */
int pe_rmat_edges(pe_ctx_t* ctx, uint32_t scale, uint64_t seed, double a, double b, double c,
                  uint64_t e0, uint64_t n, uint32_t* src, uint32_t* dst, float* w);
/* Objects built on g (BFS states, PPR blocks, servers, reverses) must not be used after g is
 * destroyed; destroying them afterwards is allowed. */
int pe_graph_destroy(pe_graph_t* g);
int pe_graph_get_info(const pe_graph_t* g, pe_graph_info_t* info);
/* Build from CSV text in host memory, parsed on the device with the reference's
 * csv_parser rules (parsers.hpp:225-293): one `src,dst[,w]` edge per '\n'-separated
 * line, fields split on runs of ',', ids read as strtoul(, 10) reads them (leading
 * whitespace, an optional sign, digits up to the first other character), the weight
 * as float_decoder's strtof (parsers.hpp:72-76); empty lines and lines missing a field
 * are skipped.  flags as pe_graph_create, plus PE_GRAPH_ENCODE_IDS: compact the ids
 * that occur (src and dst, 64-bit) to 0..n-1 in ascending order, replacing
 * vid_encoder_t::encode (vid_encoder.hpp:41-102, which numbers them in hash-table
 * order); pe_graph_vertex_ids returns the original id of each.  Without it, V = the
 * largest id + 1 and an id above 2^32 - 1 fails with PE_ERR_RANGE (the reference's
 * CHECK), naming the line.  Weight literals are decimal with at most 15 significant
 * digits and a normal float value; they round exactly as strtof does (literals
 * further than 10^22 from the integers are rejected when their value lies within a
 * few double ulps of a float rounding midpoint).  Other literals (inf, nan, hex,
 * subnormal or overflowing values) fail with PE_ERR_RANGE, naming the line. */
#define PE_GRAPH_ENCODE_IDS 16u
int pe_graph_create_csv(pe_ctx_t* ctx, const char* text, uint64_t nbytes, uint32_t flags, pe_graph_t** out);
/* original_ids[V]: the id each dense input id stands for (identity unless encoded). */
int pe_graph_vertex_ids(pe_graph_t* g, uint64_t* original_ids);

/* Copy the CSC (col_ptr[V+1], row_idx[nnz], optional normalised weights[nnz]). */
int pe_graph_download_csc(pe_graph_t* g, uint64_t* col_ptr, uint32_t* row_idx, float* wn);
int pe_graph_download_csr(pe_graph_t* g, uint64_t* row_ptr, uint32_t* col_idx);
/* rank_of[input id] = internal id (identity unless PE_GRAPH_RELABEL_DEGREE). */
int pe_graph_permutation(pe_graph_t* g, uint32_t* rank_of);
/* Re-tile the pull gather (merge-path items per tile: 1024, 2048 or 4096). */
int pe_graph_set_tile(pe_graph_t* g, uint32_t tile_items);

/* Pull layout used by the FP32 rank iterations (pe_pagerank, pe_ppr, pe_rank_iterate,
 * weighted or not) -- the same aggregate_message pass (message_passing.hpp:88-144)
 * under two decompositions:
 *   PE_PULL_MERGE_PATH  CSC gather of x[u] per edge, compensated FP32 sums (default);
 *   PE_PULL_BINNED      partition-centric: one contribution per (source, bin) pair
 *                       is streamed into destination bins, then each bin is summed
 *                       in LDS as fixed-point integers (terms truncated at 2^-64,
 *                       order-independent, bit-reproducible).  Builds a second,
 *                       bin-ordered copy of the edges (2.5 B per edge, +4 B with
 *                       weights, + 6 B per (source, bin) pair) on first use.
 * FP64 iterations and raw sums (pe_pull_sum) always use the merge-path gather.
 * opts may be NULL (defaults); its fields tune the binned layout. */
#define PE_PULL_MERGE_PATH 0
#define PE_PULL_BINNED 1
typedef struct {
  uint32_t src_bits;    /* log2 sources per chunk staged in LDS (<= 15; default 15) */
  uint32_t dst_bits;    /* log2 destinations per bin summed in LDS (<= 14; default 14) */
  uint32_t run_cap;     /* max updates per scatter run record, 64-256 (default 256) */
  uint32_t unit_log2;   /* log2 max entries per gather unit; bigger bins are split (default: the
                           largest power of two up to half a CU's share of the entries, >= 10) */
} pe_binned_opts_t;
int pe_graph_set_pull_layout(pe_graph_t* g, int layout, const pe_binned_opts_t* opts);

/* ===========================================================================
 * Engine state.  pe_vec replaces dense_state_t
 * (plato/graph/state/dense_state.hpp:43-175; engine alloc_v_state,
 * plato/engine/dualmode.hpp:231-236); pe_bitmap replaces bitmap_t
 * (plato/util/bitmap.hpp:43-350; engine alloc_v_subset, dualmode.hpp:238-242),
 * same word layout: (bits>>6)+1 uint64 words, bit i of word i>>6 is i&63.
 * ========================================================================= */
int pe_vec_create(pe_ctx_t* ctx, int dtype, uint64_t n, pe_vec_t** out);
int pe_vec_destroy(pe_vec_t* v);
int pe_vec_fill(pe_vec_t* v, double value);                         /* dense_state.hpp:241-254 */
int pe_vec_copy(pe_vec_t* dst, const pe_vec_t* src);                 /* same dtype and length */
int pe_vec_upload(pe_vec_t* v, const void* host, uint64_t n);        /* raw, same dtype */
int pe_vec_download(pe_vec_t* v, void* host, uint64_t n);            /* raw, same dtype */
int pe_vec_download_f64(pe_vec_t* v, double* host, uint64_t n);      /* converted */
int pe_vec_set(pe_vec_t* v, uint64_t i, double value);               /* operator[] store */
/* sum over all entries in FP64 (dense_state_t::foreach reduction) */
int pe_vec_sum(pe_vec_t* v, double* out);

int pe_bitmap_create(pe_ctx_t* ctx, uint64_t bits, pe_bitmap_t** out);
int pe_bitmap_destroy(pe_bitmap_t* b);
int pe_bitmap_clear(pe_bitmap_t* b);                                 /* bitmap.hpp clear */
int pe_bitmap_fill(pe_bitmap_t* b);                                  /* bitmap.hpp:168-179 */
int pe_bitmap_set_bit(pe_bitmap_t* b, uint64_t i);                   /* bitmap.hpp:219-222 */
int pe_bitmap_clr_bit(pe_bitmap_t* b, uint64_t i);
int pe_bitmap_get_bit(pe_bitmap_t* b, uint64_t i, int* out);
int pe_bitmap_count(pe_bitmap_t* b, uint64_t* out);                  /* bitmap.hpp:229-238 */
int pe_bitmap_msb(pe_bitmap_t* b, uint64_t* out);                    /* msb, 0 when empty */
int pe_bitmap_copy_from(pe_bitmap_t* dst, const pe_bitmap_t* src);
int pe_bitmap_upload(pe_bitmap_t* b, const uint64_t* words, uint64_t nwords);
int pe_bitmap_download(pe_bitmap_t* b, uint64_t* words, uint64_t nwords);
/* ascending list of set bits (bitmap.hpp to_vector); *count gets the number */
int pe_bitmap_to_vector(pe_bitmap_t* b, uint32_t* ids, uint64_t cap, uint64_t* count);

/* ===========================================================================
 * Engine operators (plato/engine/dualmode.hpp, plato/graph/message_passing.hpp)
 * ========================================================================= */
/* spread_message (plato/graph/message_passing.hpp:152-249): every vertex v set in
 * `actives` sends one message to target t = key[v] (key: optional PE_U32 vector
 * covering the bitmap; without it t = v), delivered to a sink from a closed set:
 *   PE_SPREAD_COUNT  sink[t] += 1 (PE_U32 sink, no value vector)
 *   PE_SPREAD_ADD    sink[t] += value[v] (PE_F32 or PE_F64 sink, value of the same
 *                    dtype; without a value vector each message adds 1)
 * Every delivered message returns 1, so *result is the number of messages (the
 * reference's sum of sink returns).  A target outside the sink fails with
 * PE_ERR_RANGE after the other messages are delivered.  FP atomics make the order of
 * additions into one target unspecified. */
#define PE_SPREAD_COUNT 0
#define PE_SPREAD_ADD 1
int pe_spread_message(pe_bitmap_t* actives, int op, pe_vec_t* key, pe_vec_t* value, pe_vec_t* sink,
                      uint64_t* result);
/* out-degrees counted from the in-edges: generate_dense_out_degrees_fg
 * (plato/graph/structure.hpp:292-346) */
int pe_graph_out_degrees(pe_graph_t* g, uint32_t* od_host);
int pe_graph_in_degrees(pe_graph_t* g, uint32_t* id_host);
/* sum of out-degrees over the active set: dualmode.hpp:264-277 */
int pe_active_edges(pe_graph_t* g, pe_bitmap_t* actives, uint64_t* out);
/* push/pull decision of foreach_edges: push iff (double)A/(double)E' < threshold
 * (dualmode.hpp:261-279); *is_push = 1 for push */
int pe_choose_mode(pe_graph_t* g, pe_bitmap_t* actives, double threshold, int* is_push,
                   uint64_t* active_edges);

/* Pull-mode aggregate_message (message_passing.hpp:88-144) with the gather-sum
 * signal and the write_add slot of pagerank.hpp:104-116: y[v] = sum of x[u]
 * over in-edges u->v (times the normalised weight when weighted != 0, as
 * weighted_pagerank.hpp:115-128).  x, y: PE_F32 or PE_F64, length V. */
int pe_pull_sum(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, int weighted);

/* pe_pull_sum with no device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The scratch pe_pull_sum allocates
 * per call (compensated partial sums per merge-path tile, per-block partials) comes
 * from a static device array of PE_PULL_SUM_NOALLOC_MAX_BYTES, one per GPU, held until
 * the process ends; a graph needing more (about 2 x 16 B per 2048 (edge + vertex)
 * items in FP64, half that in FP32, so up to ~4 G items in FP64) fails with
 * PE_ERR_RANGE.  Kernel timing is skipped.  Every call on one GPU shares its array:
 * calls that overlap on one GPU (other streams or contexts) race. */
#define PE_PULL_SUM_NOALLOC_MAX_BYTES (64u << 20)
int pe_pull_sum_noalloc(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, int weighted);

/* The initial state of the rank algorithms (the curt_rank.foreach initialisers of
 * pagerank.hpp:92-100 and personalized_pagerank.hpp:73-83): x[v] = value for every v
 * (seeds NULL) or for v in seeds (0 elsewhere), divided by od(v) when divide_by_od and
 * od(v) > 0.  x: PE_F32 or PE_F64, length V. */
int pe_rank_init(pe_graph_t* g, pe_vec_t* x, const pe_bitmap_t* seeds, double value, int divide_by_od);

/* Push-mode broadcast_message with the push operators of push/pull PPR
 * (personalized_pagerank.hpp:222-239): every vertex in `actives` adds x[v] to y[w] for
 * each out-edge v->w (FP atomic adds, the reference's write_add, so the last bits can
 * vary run to run), and each such w joins `actives` once the step is over (the step
 * pushes from the set it started with).  x, y: both PE_F32 or both PE_F64, length V.
 * Needs PE_GRAPH_BUILD_CSR. */
int pe_push_add(pe_graph_t* g, pe_bitmap_t* actives, pe_vec_t* x, pe_vec_t* y);

/* Restart-apply of the rank algorithms (the dense_state_t::foreach lambdas of
 * pagerank.hpp:118-140, weighted_pagerank.hpp:130-148,
 * personalized_pagerank.hpp:101-125, trustrank.hpp:226-250). */
typedef struct {
  double damping;        /* d */
  double base;           /* restart added to every vertex: 1-d for PageRank, else 0 */
  double seed_mass;      /* restart added to seed vertices: (1-d)/|S| */
  const pe_bitmap_t* seeds; /* seed set or NULL */
  int32_t divide_by_od;  /* state holds rank/od (PageRank, PPR) vs rank (weighted) */
  int32_t final_iter;    /* last iteration: store the undivided rank, delta 0 */
} pe_apply_params_t;
/* y holds gathered sums on entry and the new state on exit; x is the current
 * state; *delta gets the reference's delta (FP64). */
int pe_rank_apply(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, const pe_apply_params_t* p,
                  double* delta);
/* Fused pull iteration: pe_pull_sum + pe_rank_apply in one pass. */
int pe_rank_iterate(pe_graph_t* g, pe_vec_t* x, pe_vec_t* y, int weighted,
                    const pe_apply_params_t* p, double* delta);

/* ===========================================================================
 * Algorithms (plato/algo/...).  Whole runs on the device; results returned as
 * FP64 host arrays like the reference's dense_state_t<double>.
 * ========================================================================= */
typedef struct {
  uint32_t iterations;   /* pagerank_opts_t::iteration_ */
  double damping;        /* pagerank_opts_t::damping_ */
  double eps;            /* pagerank_opts_t::eps_ (0: run all iterations) */
  int32_t precision;     /* 32: FP32 state, compensated sums; 64: FP64 state */
  int32_t weighted;      /* weighted_pagerank semantics */
  double l1_norm_tol;    /* >0: also stop when delta/sum(rank) < tol (then one final iteration) */
} pe_rank_opts_t;

typedef struct {
  uint32_t iterations_run;
  double last_delta;
  double seconds;        /* device time of the iterations */
  double gteps;          /* (iterations * nnz) / seconds / 1e9 */
} pe_rank_stats_t;

/* plato::algo::pagerank (plato/algo/pagerank/pagerank.hpp:60-149) and
 * weighted_pagerank (plato/algo/pagerank/weighted_pagerank.hpp:60-157).
 * out: host FP64 [V]; deltas (optional): host FP64 [iterations]. */
int pe_pagerank(pe_graph_t* g, const pe_rank_opts_t* opts, double* out, double* deltas,
                pe_rank_stats_t* stats);
/* plato::algo::personalized_pagerank, pull form
 * (plato/algo/ppr/personalized_pagerank.hpp:40-134) for nseeds == 1, and the
 * TrustRank restart over a seed set (plato/algo/trustrank/trustrank.hpp:197-250)
 * for nseeds > 1 (duplicates count once). */
int pe_ppr(pe_graph_t* g, const uint32_t* seeds, uint32_t nseeds, const pe_rank_opts_t* opts,
           double* out, double* deltas, pe_rank_stats_t* stats);
/* push/pull PPR (personalized_pagerank.hpp:137-275); modes (optional) gets
 * 1 = push, 0 = pull per executed iteration.  Needs PE_GRAPH_BUILD_CSR. */
int pe_ppr_pushpull(pe_graph_t* g, uint32_t src, const pe_rank_opts_t* opts, double* out,
                    double* deltas, uint8_t* modes, pe_rank_stats_t* stats);

/* TrustRank (plato/algo/trustrank/trustrank.hpp:23-259): good seeds chosen among the
 * good vertices (select_good_seeds, :80-145), then the seed-set restart iteration of
 * pe_ppr (:163-259).  select_method: 0 inverse PageRank (PageRank of the reversed
 * graph: pass g_rev from pe_graph_create_reverse, or NULL for a symmetrized g),
 * 1 PageRank, 2 random.  The ranking PageRank uses the reference's defaults (100
 * iterations, damping 0.85, eps 0.001, FP64); the top seed_num vertices by (rank desc,
 * id asc) are kept if good.  Random selection draws with std::minstd_rand0 seeded by
 * random_seed (the reference seeds it from std::random_device, so its draws are not
 * reproducible).  If V <= seed_num every good vertex is a seed. */
typedef struct {
  uint32_t iterations;     /* trustrank_opts_t::iteration_ (100) */
  double damping;          /* damping_ (0.85) */
  uint32_t seed_num;       /* seed_num_ (100) */
  uint32_t select_method;  /* select_method_ (0) */
  double eps;              /* eps_ (0.001) */
  uint64_t random_seed;    /* method 2 */
  int32_t precision;       /* of the TrustRank iteration: 32 or 64 (0 = 64, like the reference) */
} pe_trustrank_opts_t;
/* good_words: host bitmap_t words ((V >> 6) + 1) of the good vertices, or NULL (all).
 * seeds_words: host words receiving the chosen seeds; nseeds: their count. */
int pe_trustrank_select(pe_graph_t* g, pe_graph_t* g_rev, const uint64_t* good_words,
                        const pe_trustrank_opts_t* opts, uint64_t* seeds_words, uint32_t* nseeds);
/* The whole algorithm; out: host FP64 [V]; seeds_words optional. */
int pe_trustrank(pe_graph_t* g, pe_graph_t* g_rev, const uint64_t* good_words, const pe_trustrank_opts_t* opts,
                 double* out, uint64_t* seeds_words, pe_rank_stats_t* stats);

/* ---- k-core --------------------------------------------------------------- */

typedef struct {
  uint32_t epochs;    /* epochs run (the last one modifies nothing) */
  uint64_t modified;  /* coreness decrements over all epochs */
  double seconds;     /* device time of the whole loop, HIP events */
} pe_kcore_stats_t;

/* Replaces plato::algo::kcore_algo_t::compute_shell_index (plato/algo/kcore/kcore.hpp:80-268)
 * over the graph's in-edges (the DCSC "incomings"), or its out-edges when out_edges is
 * set (a BCSR passed as "incomings", as example/kcore_simple.cc does; needs
 * PE_GRAPH_BUILD_CSR).  Writes coreness[V] in input ids (0xFFFFFFFF never remains: vertices
 * without neighbours get 0), and per epoch e < cap the number of modified vertices and
 * the size of the next active set (the reference's modified / need_modified log). */
int pe_kcore(pe_graph_t* g, int out_edges, uint32_t* coreness, uint64_t* epoch_modified,
             uint64_t* epoch_next, uint32_t cap, pe_kcore_stats_t* stats);

/* ---- connected components -------------------------------------------------- */

typedef struct {
  uint32_t epochs;      /* epochs run (the last one lowers no label) */
  uint32_t components;  /* distinct labels over the id space [0, V) */
  double seconds;       /* device time of the whole loop, HIP events */
} pe_cc_stats_t;

/* Replaces plato::algo::connected_component_t::compute (plato/algo/cgm/connected_component.hpp:
 * 170-234): label propagation with the engine's push/pull rule (push needs
 * PE_GRAPH_BUILD_CSR; without it every epoch pulls).  Writes labels[V] in input ids:
 * the smallest input id from which each vertex is reachable (on a symmetrized graph,
 * the smallest id of its component), and per epoch e < cap the mode (1 push, 0 pull)
 * and the sum of the slots' returns (the reference's logged actives; they depend on
 * the order in which labels spread, the labels do not). */
int pe_connected_components(pe_graph_t* g, uint32_t* labels, uint8_t* modes, uint64_t* actives, uint32_t cap,
                            pe_cc_stats_t* stats);

/* ---- HyperANF ---------------------------------------------------------------- */

typedef struct {
  uint32_t iterations;  /* hyperanf_opts_t::iteration_ (20) */
  uint32_t bits;        /* HyperLogLog registers per counter = 2^bits, 6..16 (the BitWidth argument) */
} pe_hyperanf_opts_t;

/* Replaces plato::algo::hyperanf (plato/algo/hyperanf/hyperanf.hpp:80-197) over the
 * in-edges: anfs[0..cap) gets the neighbourhood-function estimates (the initial one
 * first), *nanfs their number, *avg_distance GetAvgDistance of them (the reference's
 * return value), registers (optional, [V << bits] bytes in input-id order) the final
 * counters, *seconds the device time. */
int pe_hyperanf(pe_graph_t* g, const pe_hyperanf_opts_t* opts, double* anfs, uint32_t cap, uint32_t* nanfs,
                double* avg_distance, uint8_t* registers, double* seconds);

/* n-step degrees: plato::algo::nstepdegrees_t::compute (plato/algo/nstepdegrees/
 * nstepdegrees.hpp:166-489).  For each target vertex (target_words: a bitmap over input
 * ids, (V >> 6) + 1 words; NULL = every vertex), the in-degree is the HyperLogLog estimate
 * (2^bits registers, bits in [6, 16]) of the vertices that reach it within `step` hops,
 * minus 1, and the out-degree the same for the vertices it reaches, stored as
 * static_cast<vid_t>(estimate) - 1; other vertices get 0.  step = 0xFFFFFFFF (the driver's
 * -1) runs until no counter grows.  type: PE_NSTEP_IN, PE_NSTEP_OUT or PE_NSTEP_BOTH; the
 * out pass of a directed graph needs PE_GRAPH_BUILD_CSR, and an undirected graph's out
 * degrees equal its in degrees.  Arrays in input id order. */
#define PE_NSTEP_IN 1
#define PE_NSTEP_OUT 2
#define PE_NSTEP_BOTH 3
int pe_nstep_degrees(pe_graph_t* g, const uint64_t* target_words, uint32_t step, uint32_t bits, int type,
                     uint32_t* in_degrees, uint32_t* out_degrees, double* seconds);
/* A new graph with every edge of the directed graph g reversed (same input ids), for
 * pulls over out-edges such as inverse PageRank (the reference runs pagerank over its
 * bcsr_t out-edge structure, trustrank.hpp:101-103).  flags: PE_GRAPH_BUILD_CSR,
 * PE_GRAPH_RELABEL_DEGREE, PE_GRAPH_SAME_IDS.
 * PE_GRAPH_SAME_IDS keeps g's internal ids (and its degree relabeling, if any): device
 * vectors, bitmaps and BFS states of g are then valid for the reverse, which is how the
 * engine's reverse() walks a directed graph's edges backwards (dualmode.hpp:151-153,
 * 299-322).  It excludes PE_GRAPH_RELABEL_DEGREE. */
#define PE_GRAPH_SAME_IDS 32u
int pe_graph_create_reverse(pe_graph_t* g, uint32_t flags, pe_graph_t** out);

typedef struct {
  uint32_t epochs;
  uint32_t visited;       /* |visited| (the reference's return value) */
  uint64_t reached_edges; /* sum of od over reached vertices */
  double seconds;
  /* algorithmic bytes of the search, per level: push 8|F| + 4 sum_F deg + 4|new|; pull
   * 8|U| + 4 * (in-edges a sequential scan reads, each unvisited list up to its first
   * frontier member) + V/8 */
  uint64_t alg_bytes;
} pe_bfs_stats_t;
/* plato::algo::breadth_first_search (plato/algo/bfs/bfs.hpp:55-155) with the
 * engine switch (dualmode.hpp:261-279).  levels: host int32 [V] (root 0,
 * unreached -1).  modes / frontier / active_edges (optional, max_epochs long)
 * get the per-epoch decision (1 push, 0 pull), |F| and sum of od over F.
 * Needs PE_GRAPH_BUILD_CSR for directed graphs. */
int pe_bfs(pe_graph_t* g, uint32_t root, double threshold, int32_t* levels, uint8_t* modes,
           uint64_t* frontier, uint64_t* active_edges, uint32_t max_epochs,
           pe_bfs_stats_t* stats);

/* ===========================================================================
 * Batched PPR.  k = 64 queries share one FP32 rank block [V][64]; each column is
 * personalized_pagerank (plato/algo/ppr/personalized_pagerank.hpp:40-134) with
 * the TrustRank restart over the query's seed set (trustrank.hpp:197-250) and
 * the reference's per-run stopping rule applied per column: after an iteration
 * with delta < tol one final iteration stores the undivided rank
 * (pagerank.hpp:137-139).  Columns retire and are refilled independently:
 * slot states EMPTY 0, LOADING 1, ACTIVE 2, FINAL 3, DONE 4.  A column's
 * arithmetic does not depend on the other columns.
 * ========================================================================= */
typedef struct pe_ppr_batch_s pe_ppr_batch_t;
typedef struct {
  uint32_t k;           /* columns; must be 64 */
  double damping;       /* 1 - alpha */
  double tol;           /* per-column delta threshold (the reference's eps_) */
  uint32_t max_iters;   /* iteration cap per column (the reference's iteration_) */
  uint32_t task_edges;  /* edges per wave task (0: 4096); longer columns are split */
  /* Pass plan: a pass with n covered columns (LOADING, ACTIVE, FINAL, or DONE for one copy
   * pass) runs on the packed block if n <= plan_packed (at most 8), with 4 edges per row
   * load if n <= plan_edges4 (at most 16), 2 if n <= plan_edges2 (at most 32), else 1.
   * 0 takes the widest (8, 16, 32); PE_BATCH_PLAN_OFF never uses that kernel. */
  uint32_t plan_packed, plan_edges4, plan_edges2;
} pe_ppr_batch_opts_t;
#define PE_BATCH_PLAN_OFF 0xFFFFFFFFu

int pe_ppr_batch_create(pe_graph_t* g, const pe_ppr_batch_opts_t* opts, pe_ppr_batch_t** out);
int pe_ppr_batch_destroy(pe_ppr_batch_t* b);
/* device bytes held by the batch (rank blocks, seed masks, task lists) */
int pe_ppr_batch_bytes(const pe_ppr_batch_t* b, uint64_t* bytes);
/* put a query (seed set, input ids) into an EMPTY slot; it starts LOADING */
int pe_ppr_batch_admit(pe_ppr_batch_t* b, uint32_t slot, const uint32_t* seeds, uint32_t nseeds);
/* free a slot (any state) */
int pe_ppr_batch_release(pe_ppr_batch_t* b, uint32_t slot);
/* enqueue `iterations` passes (captured once as a graph and replayed; async) */
int pe_ppr_batch_step(pe_ppr_batch_t* b, uint32_t iterations);
/* slot states, iterations run and last delta per slot (arrays of 64; any may be NULL) */
int pe_ppr_batch_poll(pe_ppr_batch_t* b, uint32_t* status, uint32_t* iters, double* last_delta);
/* full rank vectors of DONE slots, input id order: ranks[n][V] (FP64) */
int pe_ppr_batch_results(pe_ppr_batch_t* b, const uint32_t* slots, uint32_t n, double* ranks);
/* top-k (rank desc, id asc; trustrank.hpp:109-112) of DONE slots: ids/vals [n][k] */
int pe_ppr_batch_topk(pe_ppr_batch_t* b, const uint32_t* slots, uint32_t n, uint32_t k, uint32_t* ids,
                      float* vals);

/* ===========================================================================
 * PPR serving loop: continuous batching of queries into 64-column blocks
 * (pe_ppr_batch), with admission by free HBM.  No reference counterpart: the
 * reference runs one PPR per process launch (example/personalized_pagerank.cc);
 * each query here has that query's semantics (see pe_ppr_batch).
 * ========================================================================= */
typedef struct pe_server_s pe_server_t;
typedef struct {
  double damping;            /* 1 - alpha */
  double tol;                /* per-query delta threshold */
  uint32_t max_iters;        /* per-query iteration cap */
  uint32_t topk;             /* vertices returned per query (rank desc, id asc) */
  uint32_t iters_per_pass;   /* iterations enqueued per block between slot checks */
  uint32_t max_batches;      /* upper bound on resident 64-column blocks */
  uint64_t hbm_reserve_bytes;/* free HBM kept back when opening another block */
} pe_server_opts_t;
typedef struct {
  uint64_t id;
  double arrival_s;          /* seconds since the loop started */
  double admit_s;            /* when the query got a column slot */
  double done_s;             /* when its result was read out */
  uint32_t iterations;
  uint32_t batch;            /* block and slot that ran it */
  uint32_t slot;
} pe_request_stats_t;

int pe_server_create(pe_graph_t* g, const pe_server_opts_t* opts, pe_server_t** out);
int pe_server_destroy(pe_server_t* s);
/* queue a query (seed set in input ids) that arrives at arrival_s */
int pe_server_submit(pe_server_t* s, uint64_t id, const uint32_t* seeds, uint32_t nseeds, double arrival_s);
/* one pass: admit arrived queries, run iterations on busy blocks, harvest finished columns */
int pe_server_pump(pe_server_t* s, uint32_t* completed);
/* move up to max finished queries out: stats[max], topk ids/vals [max][topk] */
int pe_server_take(pe_server_t* s, uint32_t max, pe_request_stats_t* stats, uint32_t* topk_ids,
                   float* topk_vals, uint32_t* taken);
int pe_server_info(pe_server_t* s, uint32_t* batches, uint32_t* in_flight, uint32_t* waiting,
                   uint64_t* passes);
/* replay an arrival trace in real time until every query is answered; seeds of
 * query i are seeds[seed_offsets[i] .. seed_offsets[i+1]); outputs in query order */
int pe_server_run_trace(pe_server_t* s, uint32_t n, const uint64_t* ids, const double* arrivals,
                        const uint64_t* seed_offsets, const uint32_t* seeds, pe_request_stats_t* stats,
                        uint32_t* topk_ids, float* topk_vals, double* wall_s);

/* ===========================================================================
 * BFS engine step: one epoch of breadth_first_search's 5-argument
 * foreach_edges (plato/algo/bfs/bfs.hpp:95-138) with the BFS operators: push
 * claims unvisited out-neighbours of the frontier, pull gives each unvisited
 * vertex its first in-neighbour found in the frontier.  The state holds the
 * levels (the reference keeps parents; levels are the order-free part).
 * Bitmaps are in internal id order; roots and levels use input ids.
 * ========================================================================= */
typedef struct pe_bfs_s pe_bfs_t;
int pe_bfs_create(pe_graph_t* g, pe_bfs_t** out);
int pe_bfs_destroy(pe_bfs_t* b);
/* levels = -1 except root = 0; frontier = {root} */
int pe_bfs_reset(pe_bfs_t* b, uint32_t root, pe_bitmap_t* frontier);
/* one epoch from `current` into `next` (overwritten); push != 0 selects push.
 * *activated = newly reached vertices (the reference's slot-return sum),
 * *next_active_edges = sum of od over them (the next decision's A). */
int pe_bfs_step(pe_bfs_t* b, int push, pe_bitmap_t* current, pe_bitmap_t* next, uint64_t* activated,
                uint64_t* next_active_edges);
int pe_bfs_get_levels(pe_bfs_t* b, int32_t* levels);
/* Later steps of b walk the edges of g, a graph over the same vertex ids as the one b
 * was created on (the graph itself, or its reverse from pe_graph_create_reverse with
 * PE_GRAPH_SAME_IDS; PE_ERR_INVALID otherwise).  Levels and the visited set carry over.
 * The engine binds its search to the reversed edges after reverse() on a directed
 * graph, as the reference's push and pull then walk in_edges_ and out_edges_
 * (dualmode.hpp:299-322). */
int pe_bfs_bind(pe_bfs_t* b, pe_graph_t* g);

/* ===========================================================================
 * Autotuner (no reference counterpart: the reference has no tunables on this
 * path beyond dualmode_engine_opts_t::push_threshold_).  Times candidates on
 * the resident graph and keeps the fastest:
 *   PE_TUNE_PULL    merge-path items per pull tile (applied to the graph);
 *   PE_TUNE_BATCH   the 64-column PPR block's task size, and its pass plan: which
 *                   pass kernel (packed block, 4, 2 or 1 edges per row load) runs
 *                   passes with 8, 16 and 32 covered columns;
 *   PE_TUNE_BFS     the BFS push/pull ratio;
 *   PE_TUNE_BINNED  the binned layout's source chunk and destination bin widths
 *                   (applied to the graph: the best layout is left built).
 * ========================================================================= */
#define PE_TUNE_PULL 1u
#define PE_TUNE_BATCH 2u
#define PE_TUNE_BFS 4u
#define PE_TUNE_BINNED 8u
typedef struct {
  uint32_t tile_items;        /* chosen merge-path items per pull tile */
  uint32_t batch_task_edges;  /* pe_ppr_batch_opts_t::task_edges */
  double bfs_push_threshold;  /* threshold for pe_bfs */
  double pull_ms, batch_ms, bfs_ms;  /* best times seen (per iteration / per search) */
  uint32_t binned_src_bits, binned_dst_bits;  /* pe_binned_opts_t of the best binned layout */
  double binned_ms;                           /* its iteration time */
  uint32_t batch_plan_packed, batch_plan_edges4, batch_plan_edges2;  /* pe_ppr_batch_opts_t plan */
  double batch_plan_ms[3];  /* best pass times with 8, 16 and 32 covered columns */
} pe_tune_result_t;
int pe_autotune(pe_graph_t* g, uint32_t what, pe_tune_result_t* out);
/* append / look up a choice keyed by (V, nnz, directed, relabeled) in a text file */
int pe_autotune_save(pe_graph_t* g, const pe_tune_result_t* r, const char* path);
int pe_autotune_load(pe_graph_t* g, const char* path, pe_tune_result_t* r, int* found);

#ifdef __cplusplus
}
#endif

#endif /* PLATO_HIP_H */
