/*
 * oracle.h -- CPU restatement (FP64, plain C) of the reference algorithms on
 * the dual-mode edge-processing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is the checker used by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.
 *
 * Every function cites the reference file:line it restates.  Each one is
 * pinned by an independent check in tests/test_oracle.py (reference fixtures,
 * closed forms, linear solves, library BFS, invariants); the pin for each
 * function is named next to its declaration.
 */
#ifndef PLATO_ORACLE_H
#define PLATO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic graph ------------------------------------------------------ */

/* 64-bit mixing function used by the counter-based RMAT generator
 * (SplitMix64 finalizer, Steele/Lea/Flood 2014).  Pinned: published test
 * vector in tests/test_oracle.py. */
uint64_t orc_mix64(uint64_t x);

/* Integer quadrant thresholds floor(p * 2^32) for p = a, a+b, a+b+c. */
void orc_rmat_thresholds(double a, double b, double c, uint32_t out[3]);

/* Generate edges [e0, e0+n) of the RMAT graph keyed by (seed, edge id).
 * Pinned: quadrant frequencies / bit marginals (statistical) and the numpy
 * twin in oracle/rmat.py. */
void orc_rmat_edges(uint32_t scale, uint64_t seed, const uint32_t thr[3],
                    uint64_t e0, uint64_t n, uint32_t* src, uint32_t* dst);

/* The edges with ids eids[0..n) (orc_rmat_edges of each id; OpenMP). */
void orc_rmat_edges_at(uint32_t scale, uint64_t seed, const uint32_t thr[3], uint64_t n,
                       const uint64_t* eids, uint32_t* src, uint32_t* dst);

/* Edge weight in (0, 1] of edge e (24-bit uniform grid). */
float orc_rmat_weight(uint64_t seed, uint64_t e);

/* orc_rmat_weight of each edge id in eids[n] (OpenMP). */
void orc_rmat_weights(uint64_t seed, uint64_t n, const uint64_t* eids, float* out);

/* One pass over the generator's edges [0, E) at a scale too large to hold twice:
 * od[2^scale] = out-degree of every source, W (optional) = FP64 sum of each source's
 * edge weights (atomic adds: the last bits may depend on the thread schedule), *digest =
 * the digest of the edge multiset (orc_edges_digest), and the ids of the edges whose
 * destination lies in one of the ranges [lo[r], hi[r]) written to sel_eid (at most cap,
 * in no particular order).  Returns the number of such edges.
 * Pinned: equals orc_rmat_edges + bincount + orc_edges_digest (tests/test_oracle.py). */
int64_t orc_rmat_scan(uint32_t scale, uint64_t seed, const uint32_t thr[3], uint64_t E,
                      uint32_t nranges, const uint32_t* lo, const uint32_t* hi,
                      uint32_t* od, double* W, uint64_t* digest,
                      uint64_t cap, uint64_t* sel_eid);

/* Order-independent digest of an edge multiset: the sum mod 2^64 of
 * mix64(src << 32 | dst) over the edges.  A dropped, added or altered edge, or a
 * swapped direction (src != dst), changes it except with probability about 2^-64. */
uint64_t orc_edges_digest(uint64_t n, const uint32_t* src, const uint32_t* dst);

/* The same digest over the entries of a CSC (entry e of column v is the edge
 * row_idx[e] -> v), with both ends mapped through map[] when it is given (a
 * relabeled graph's internal -> input ids). */
uint64_t orc_csc_digest(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                        const uint32_t* map);

/* ---- structures ----------------------------------------------------------- */

/* In-edge CSC of an edge list: entries are the E input edges, followed (when
 * symmetrize != 0) by their E reversals, as Plato adds reverse entries for
 * undirected input (dcsc.hpp:379-384, bcsr.hpp:411-426).  Columns are sorted by
 * source id, stable in entry order (a parallel LSD radix sort; the output does not
 * depend on the thread count).  col_ptr has V+1 entries; row_idx and
 * w_out (optional) have nnz = E * (symmetrize ? 2 : 1) entries.
 * Pinned: reference DCSC fixtures (dcsc.cc:81-191) and message_passing.cc:48-88. */
int orc_build_csc(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                  const float* w, int symmetrize,
                  uint64_t* col_ptr, uint32_t* row_idx, float* w_out);

/* Out-edge CSR: the same entries grouped by source, sorted by destination.
 * Pinned with orc_build_csc: numpy's stable lexsort of the same entries. */
int orc_build_csr(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                  int symmetrize, uint64_t* row_ptr, uint32_t* col_idx);

/* od(u) = number of in-edge entries whose source is u (structure.hpp:292-338).
 * Pinned: structure.cc:137-218 known degrees (19/59/19). */
void orc_out_degrees(uint32_t V, uint64_t nnz, const uint32_t* row_idx, uint32_t* od);

/* ---- rank algorithms (all FP64) ------------------------------------------ */

/* plato::algo::pagerank (pagerank.hpp:60-149).  Writes the returned state (the
 * undivided rank after the final iteration) to out[V]; deltas[i] (optional,
 * length iters) receives the delta of the i-th executed iteration (0 for the
 * final one).  Returns the number of executed iterations.
 * Pinned: chain closed form, linear solve of the fixed point. */
int orc_pagerank(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                 const uint32_t* od, uint32_t iters, double damping, double eps,
                 double* out, double* deltas, int nthreads);

/* One Plato PageRank iteration over columns [v_lo, v_hi): reads x (state =
 * rank/od), writes y.  Returns the delta contribution of the range.  Used as
 * the timed CPU baseline and by the property tests. */
double orc_pagerank_step(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                         const uint32_t* od, double damping, int final_iter,
                         const double* x, double* y, uint32_t v_lo, uint32_t v_hi,
                         int nthreads);

/* plato::algo::weighted_pagerank (weighted_pagerank.hpp:60-157); w is aligned
 * with row_idx.  Pinned: chain (equals unweighted), linear solve. */
int orc_weighted_pagerank(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                          const float* w, uint32_t iters, double damping, double eps,
                          double* out, double* deltas);

/* One weighted PageRank iteration over columns [v_lo, v_hi) (weighted_pagerank.hpp:
 * 110-143): y(v) = 1 - d + d * sum over in-edges u of x(u) * w / W(u), x the
 * undivided rank state, W(u) the sum of u's out-weights.  Returns the range's delta.
 * Pinned: equals one iteration of orc_weighted_pagerank (tests/test_oracle.py). */
double orc_weighted_pagerank_step(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                                  const float* w, const double* W, double damping,
                                  const double* x, double* y, uint32_t v_lo, uint32_t v_hi);

/* Personalized PageRank with restart mass spread evenly over a seed set:
 * nseeds == 1 is plato::algo::personalized_pagerank, pull form
 * (personalized_pagerank.hpp:40-134); nseeds > 1 is the TrustRank restart
 * (trustrank.hpp:197-250).  Duplicate seeds count once (a seed bitmap).
 * Pinned: chain closed form, linear solve, linearity in the seed set. */
int orc_ppr(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
            const uint32_t* od, const uint32_t* seeds, uint32_t nseeds,
            uint32_t iters, double damping, double eps, double* out, double* deltas);

/* Push/pull PPR (personalized_pagerank.hpp:137-275): sticky switch
 * active_edges > E'/20 (integer division); push scatters x(v) of the
 * accumulated active set over out-edges.  modes[i] = 1 push, 0 pull.
 * Pinned: equals orc_ppr bit for bit (x = 0 outside the active set). */
int orc_ppr_pushpull(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                     const uint64_t* row_ptr, const uint32_t* col_idx,
                     const uint32_t* od, uint64_t edges_prime, uint32_t src,
                     uint32_t iters, double damping, double eps,
                     double* out, double* deltas, uint8_t* modes);

/* ---- BFS ------------------------------------------------------------------- */

/* plato::algo::breadth_first_search (bfs.hpp:55-155) with the engine switch
 * (dualmode.hpp:261-279): push iff (double)A/(double)E' < threshold, where A is
 * the sum of od over the frontier.  Writes levels (root 0, unreached -1), and
 * per epoch: mode (1 push / 0 pull), frontier size |F|, active edges A.
 * Returns the number of epochs; *visited gets |visited|.
 * Pinned: scipy.sparse.csgraph shortest paths, binary tree closed form. */
int orc_bfs(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
            const uint64_t* row_ptr, const uint32_t* col_idx, const uint32_t* od,
            uint64_t edges_prime, double threshold, uint32_t root,
            int32_t* levels, uint8_t* modes, uint64_t* frontier, uint64_t* active_edges,
            uint32_t max_epochs, uint32_t* visited);

/* ---- k-core ----------------------------------------------------------------- */

/* plato::algo::kcore_algo_t::compute_shell_index (kcore.hpp:80-268) on one partition,
 * over the adjacency the caller passes as "incomings" (ptr[V+1], idx): a DCSC gives
 * in-neighbours, a BCSR (as example/kcore_simple.cc passes) out-neighbours.
 * coreness starts at the adjacency length, or 0xFFFFFFFF for vertices without one;
 * every vertex is active in epoch 0.  Per epoch (Jacobi: every active vertex reads the
 * estimates of the previous epoch): an active vertex with no neighbours gets 0; else
 * est = min(len, estimate[v]), count[min(est, estimate[u])] over the neighbours u, and
 * the largest i in [1, est] whose suffix count reaches i replaces coreness[v] when it is
 * smaller, which marks v modified and activates its neighbours for the next epoch
 * (when no such i exists coreness[v] is left as it is, as the reference's loop does).
 * Stops after the epoch that activates nothing.  Returns the number of epochs and, per
 * epoch e < cap, the modified count and the size of the next active set.
 * Pinned: networkx core_number on simple undirected graphs; hand-worked directed
 * cases (including a vertex left unchanged because no i qualifies). */
int orc_kcore(uint32_t V, const uint64_t* ptr, const uint32_t* idx, uint32_t* coreness,
              uint64_t* epoch_modified, uint64_t* epoch_next, uint32_t cap);

/* ---- connected components ---------------------------------------------------- */

/* The labels plato::algo::connected_component_t::compute (cgm/connected_component.hpp:
 * 170-234) converges to: labels start at the vertex ids and only ever take the minimum
 * of a label pushed along an out-edge or pulled over an in-edge, so at the fixed point
 * label(v) = the smallest id u from which v is reachable (u = v included); on an
 * undirected graph, the smallest id of v's component.  Computed here independently of
 * the iteration: a search from every u in ascending order over out-edges labels what it
 * reaches first (searches stop at vertices already labelled, whose reach is labelled).
 * Pinned: the reference test's own BFS labels (cgm/test/test_cgm.cc:83-139) on
 * v100_e2150_ua_c3.csv, scipy connected_components, brute-force reachability. */
void orc_cc_labels(uint32_t V, const uint64_t* row_ptr, const uint32_t* col_idx, uint32_t* labels);

/* ---- HyperANF ---------------------------------------------------------------- */

/* MurmurHash3_x86_32 (A. Appleby, public domain), the hash of plato/util/hash.hpp:33-96.
 * Pinned: the published test vectors of the algorithm. */
uint32_t orc_murmur3_32(const void* key, uint32_t len, uint32_t seed);

/* hyperloglog_t<P>::estimate (plato/util/hyperloglog.hpp:77-116) of 2^P registers. */
double orc_hll_estimate(const uint8_t* regs, uint32_t P);

/* plato::algo::hyperanf (hyperanf.hpp:80-197) with 2^P-register counters
 * (hyperloglog.hpp:32-142): each vertex starts with add(&v, 4) (murmur3 of its 4-byte id,
 * seed 0x5f61767a; register hash >> (32-P), rank 1 + leading zeros of hash << P, at most
 * 33-P); an iteration gives every vertex the register-wise max of its counter and its
 * in-neighbours' counters of the previous iteration (push and pull deliver the same);
 * anf = sum of the estimates; the loop stops after `iters` iterations or when anf grows
 * by less than DBL_EPSILON (that anf is not kept).  Writes the anfs (first: the initial
 * one) and returns their number; regs[V << P] gets the final counters.
 * Pinned: exact BFS average distance within 0.05 on raw_graph_7_7.csv (the reference
 * test), counters equal to the HLL of each vertex's in-reachable set within t steps. */
int orc_hyperanf(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx, uint32_t P, uint32_t iters,
                 uint8_t* regs, double* anfs, uint32_t cap);

/* n-step degrees (plato/algo/nstepdegrees/nstepdegrees.hpp:166-489) over the adjacency
 * the caller passes: in-edges (a CSC: the in-degree pass, whose books gather over in-edges)
 * or out-edges (a CSR: the out-degree pass on the reversed engine).  Every vertex's book is a
 * 2^P-register HyperLogLog started with its own id (as orc_hyperanf); each iteration gives
 * every vertex the register-wise max of its book and its adjacency's books of the previous
 * iteration, for `step` iterations or until no book grows; degree[v] =
 * (uint32)estimate - 1.  With every vertex covered this equals the reference's targets'
 * values (its covered set holds every vertex within step hops of a target).
 * Pinned: the reference test's known answers (test_nstepdegrees.cc: raw_graph_7_7.csv,
 * step 1, bits 6: in(0) = 3, out(0) = 1, in(4) = 1, out(4) = 1) and the HLL of each vertex's
 * step-hop ball built from scratch. */
int orc_nstep_degrees(uint32_t V, const uint64_t* ptr, const uint32_t* idx, uint32_t P, uint32_t step,
                      uint32_t* degree);

/* GetAvgDistance (hyperanf.hpp:58-67) of n anfs. */
double orc_anf_avg_distance(const double* anfs, uint32_t n);

/* ---- edge-list ingestion ---------------------------------------------------- */

/* Plato's csv_parser (parsers.hpp:225-293) over a byte buffer, with the same C
 * library calls: lines as std::istream::getline splits them on '\n' (an empty line
 * is skipped, the text after the last '\n' is a line), fields by strtok_r on ",",
 * src and dst by strtoul(, 10), the optional weight by float_decoder's strtof
 * (parsers.hpp:72-76).  A line without a destination (or, when weighted, without a
 * weight) is skipped.  Writes up to cap edges in file order and returns their number;
 * returns -(line + 1) for the first line whose src or dst exceeds 2^32 - 1 (where the
 * reference's CHECK aborts; src is checked before dst is looked for).
 * Pinned: the reference's data files give their known edges; hand-written lines with
 * whitespace, CRLF, repeated commas, signs, missing fields and trailing text. */
int64_t orc_parse_csv(const char* bytes, uint64_t n, int weighted, uint64_t cap, uint32_t* src,
                      uint32_t* dst, float* w);

/* Writes `src,dst\n` lines (decimal) into out (capacity cap bytes); returns the
 * number of bytes, or -1 if cap is too small.  Test data generator for the parser
 * (OpenMP over line blocks). */
int64_t orc_format_csv(uint64_t E, const uint32_t* src, const uint32_t* dst, char* out, uint64_t cap);

/* ---- top-k ----------------------------------------------------------------- */

/* Top k of r by (rank desc, id asc) (trustrank.hpp:109-112). */
void orc_topk(uint32_t V, const double* r, uint32_t k, uint32_t* ids, double* vals);

#ifdef __cplusplus
}
#endif

#endif
