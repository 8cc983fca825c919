/*
 * oracle.c -- CPU restatement (FP64) of the reference's dual-mode engine
 * clients.  TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never by the product library.
 *
 * See oracle.h for the contract of each function, the reference lines it
 * restates and the test that pins it.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- synthetic graph ------------------------------------------------------ */

uint64_t orc_mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_rmat_thresholds(double a, double b, double c, uint32_t out[3]) {
  const double p[3] = {a, a + b, a + b + c};
  for (int i = 0; i < 3; ++i) {
    double t = floor(p[i] * 4294967296.0);
    if (t < 0) t = 0;
    if (t > 4294967295.0) t = 4294967295.0;
    out[i] = (uint32_t)t;
  }
}

/* draw(e, j) = mix64(mix64(seed) ^ (e*32 + j)); level l uses the low (l even)
 * or high (l odd) half of draw(e, l/2); level l decides bit (scale-1-l).
 * Quadrant: u < t0 -> (0,0); u < t1 -> dst bit; u < t2 -> src bit; else both. */
static inline void rmat_one(uint32_t scale, uint64_t sk, const uint32_t thr[3], uint64_t e,
                            uint32_t* src, uint32_t* dst) {
  uint32_t s = 0, d = 0;
  uint64_t word = 0;
  for (uint32_t l = 0; l < scale; ++l) {
    if ((l & 1u) == 0) word = orc_mix64(sk ^ (e * 32u + (l >> 1)));
    const uint32_t u = (l & 1u) ? (uint32_t)(word >> 32) : (uint32_t)word;
    const uint32_t bit = 1u << (scale - 1u - l);
    if (u >= thr[0]) {
      if (u < thr[1]) d |= bit;
      else if (u < thr[2]) s |= bit;
      else { s |= bit; d |= bit; }
    }
  }
  *src = s;
  *dst = d;
}

void orc_rmat_edges(uint32_t scale, uint64_t seed, const uint32_t thr[3],
                    uint64_t e0, uint64_t n, uint32_t* src, uint32_t* dst) {
  const uint64_t sk = orc_mix64(seed);
  /* every edge is a pure function of (seed, edge id): the threads' split does not
   * change the output */
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 65536)
#endif
  for (uint64_t i = 0; i < n; ++i) rmat_one(scale, sk, thr, e0 + i, src + i, dst + i);
}

void orc_rmat_edges_at(uint32_t scale, uint64_t seed, const uint32_t thr[3], uint64_t n,
                       const uint64_t* eids, uint32_t* src, uint32_t* dst) {
  const uint64_t sk = orc_mix64(seed);
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 4096)
#endif
  for (uint64_t i = 0; i < n; ++i) rmat_one(scale, sk, thr, eids[i], src + i, dst + i);
}

float orc_rmat_weight(uint64_t seed, uint64_t e) {
  const uint64_t r = orc_mix64(orc_mix64(seed) ^ (e * 32u + 31u));
  return (float)((r >> 40) + 1u) * (1.0f / 16777216.0f);
}

void orc_rmat_weights(uint64_t seed, uint64_t n, const uint64_t* eids, float* out) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 65536)
#endif
  for (uint64_t i = 0; i < n; ++i) out[i] = orc_rmat_weight(seed, eids[i]);
}

/* One pass over edges [0, E) of the generator: out-degrees, FP64 out-weight sums,
 * the order-independent digest of the (src, dst) multiset, and the ids of the edges
 * whose destination lies in one of the ranges [lo[r], hi[r]). */
int64_t orc_rmat_scan(uint32_t scale, uint64_t seed, const uint32_t thr[3], uint64_t E,
                      uint32_t nranges, const uint32_t* lo, const uint32_t* hi,
                      uint32_t* od, double* W, uint64_t* digest,
                      uint64_t cap, uint64_t* sel_eid) {
  const uint32_t V = 1u << scale;
  memset(od, 0, sizeof(uint32_t) * V);
  if (W) memset(W, 0, sizeof(double) * V);
  uint64_t dg = 0, nsel = 0;
  const uint64_t B = 1u << 16;
  const uint64_t nblocks = (E + B - 1) / B;
#ifdef _OPENMP
#pragma omp parallel reduction(+ : dg)
#endif
  {
    uint32_t* s = (uint32_t*)malloc(sizeof(uint32_t) * B);
    uint32_t* d = (uint32_t*)malloc(sizeof(uint32_t) * B);
    uint64_t* pick = (uint64_t*)malloc(sizeof(uint64_t) * B);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (uint64_t b = 0; b < nblocks; ++b) {
      const uint64_t e0 = b * B, n = (e0 + B <= E) ? B : E - e0;
      const uint64_t sk = orc_mix64(seed);
      uint64_t np = 0;
      for (uint64_t i = 0; i < n; ++i) {
        const uint64_t e = e0 + i;
        uint32_t ss, dd;
        rmat_one(scale, sk, thr, e, &ss, &dd);
        s[i] = ss;
        d[i] = dd;
        dg += orc_mix64(((uint64_t)ss << 32) | dd);
        for (uint32_t r = 0; r < nranges; ++r)
          if (dd >= lo[r] && dd < hi[r]) { pick[np++] = e; break; }
      }
      for (uint64_t i = 0; i < n; ++i) {
#ifdef _OPENMP
#pragma omp atomic
#endif
        od[s[i]]++;
        if (W) {
          const double w = (double)orc_rmat_weight(seed, e0 + i);
#ifdef _OPENMP
#pragma omp atomic
#endif
          W[s[i]] += w;
        }
      }
      uint64_t at;
#ifdef _OPENMP
#pragma omp atomic capture
#endif
      { at = nsel; nsel += np; }
      for (uint64_t i = 0; i < np && at + i < cap; ++i) sel_eid[at + i] = pick[i];
    }
    free(s); free(d); free(pick);
  }
  *digest = dg;
  return (int64_t)nsel;
}

uint64_t orc_edges_digest(uint64_t n, const uint32_t* src, const uint32_t* dst) {
  uint64_t dg = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : dg) schedule(static, 65536)
#endif
  for (uint64_t i = 0; i < n; ++i) dg += orc_mix64(((uint64_t)src[i] << 32) | dst[i]);
  return dg;
}

uint64_t orc_csc_digest(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                        const uint32_t* map) {
  uint64_t dg = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : dg) schedule(dynamic, 4096)
#endif
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t dv = map ? map[v] : v;
    for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) {
      const uint64_t u = map ? map[row_idx[e]] : row_idx[e];
      dg += orc_mix64((u << 32) | dv);
    }
  }
  return dg;
}

/* ---- structures ----------------------------------------------------------- */

/* entry i of the (possibly symmetrized) entry list */
static inline void entry_at(uint64_t E, const uint32_t* src, const uint32_t* dst,
                            uint64_t i, uint32_t* s, uint32_t* d) {
  if (i < E) { *s = src[i]; *d = dst[i]; }
  else { *s = dst[i - E]; *d = src[i - E]; }
}

/* Entries grouped by major key, sorted by minor key inside a group, equal keys in
 * entry order: a stable LSD radix sort of the packed key (major << b | minor),
 * 11 bits per pass.  Each pass is parallel over contiguous slices of the entries
 * with per-slice digit counts, so it is stable and its output does not depend on
 * the number of threads.  major_is_dst selects CSC (group by dst) or CSR. */
#define ORC_RADIX 11
static int build_grouped(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                         const float* w, int symmetrize, int major_is_dst,
                         uint64_t* ptr, uint32_t* idx, float* w_out) {
  const uint64_t nnz = E * (symmetrize ? 2u : 1u);
  uint32_t b = 1;
  while (b < 32 && ((uint64_t)1 << b) < (uint64_t)V) ++b;
  const uint64_t mask = ((uint64_t)1 << b) - 1;
  const int carry_w = (w && w_out);
  uint64_t* k0 = (uint64_t*)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
  uint64_t* k1 = (uint64_t*)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
  float* w0 = carry_w ? (float*)malloc(sizeof(float) * (nnz ? nnz : 1)) : NULL;
  float* w1 = carry_w ? (float*)malloc(sizeof(float) * (nnz ? nnz : 1)) : NULL;
  if (!k0 || !k1 || (carry_w && (!w0 || !w1))) { free(k0); free(k1); free(w0); free(w1); return -1; }
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 65536)
#endif
  for (uint64_t i = 0; i < nnz; ++i) {
    uint32_t s, d;
    entry_at(E, src, dst, i, &s, &d);
    const uint64_t major = major_is_dst ? d : s, minor = major_is_dst ? s : d;
    k0[i] = (major << b) | minor;
    if (carry_w) w0[i] = w[i < E ? i : i - E];
  }
  const uint32_t bits = 2 * b, nb = 1u << ORC_RADIX;
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  uint64_t* cnt = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nth * nb);
  if (!cnt) { free(k0); free(k1); free(w0); free(w1); return -1; }
  for (uint32_t shift = 0; shift < bits; shift += ORC_RADIX) {
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
    {
      int t = 0, T = 1;
#ifdef _OPENMP
      t = omp_get_thread_num();
      T = omp_get_num_threads();
#endif
      const uint64_t lo = nnz * (uint64_t)t / (uint64_t)T, hi = nnz * (uint64_t)(t + 1) / (uint64_t)T;
      uint64_t* c = cnt + (size_t)t * nb;
      memset(c, 0, sizeof(uint64_t) * nb);
      for (uint64_t i = lo; i < hi; ++i) c[(k0[i] >> shift) & (nb - 1)]++;
#ifdef _OPENMP
#pragma omp barrier
#pragma omp single
#endif
      {
        uint64_t run = 0;
        for (uint32_t dgt = 0; dgt < nb; ++dgt)
          for (int q = 0; q < T; ++q) {
            const uint64_t x = cnt[(size_t)q * nb + dgt];
            cnt[(size_t)q * nb + dgt] = run;
            run += x;
          }
      }
      for (uint64_t i = lo; i < hi; ++i) {
        const uint64_t at = c[(k0[i] >> shift) & (nb - 1)]++;
        k1[at] = k0[i];
        if (carry_w) w1[at] = w0[i];
      }
    }
    uint64_t* tk = k0; k0 = k1; k1 = tk;
    float* tw = w0; w0 = w1; w1 = tw;
  }
  free(cnt);
  /* ptr[v] = first entry whose major is >= v */
  if (nnz == 0) {
    for (uint64_t v = 0; v <= V; ++v) ptr[v] = 0;
  } else {
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 65536)
#endif
    for (uint64_t k = 0; k < nnz; ++k) {
      const uint64_t m = k0[k] >> b;
      const uint64_t prev = k ? (k0[k - 1] >> b) + 1 : 0;
      if (k == 0 || m != (k0[k - 1] >> b))
        for (uint64_t v = prev; v <= m; ++v) ptr[v] = k;
      idx[k] = (uint32_t)(k0[k] & mask);
      if (carry_w) w_out[k] = w0[k];
    }
    for (uint64_t v = (k0[nnz - 1] >> b) + 1; v <= V; ++v) ptr[v] = nnz;
  }
  free(k0); free(k1); free(w0); free(w1);
  return 0;
}

static int ids_in_range(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst) {
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(| : bad) schedule(static, 65536)
#endif
  for (uint64_t i = 0; i < E; ++i) bad |= (src[i] >= V || dst[i] >= V);
  return !bad;
}

int orc_build_csc(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                  const float* w, int symmetrize,
                  uint64_t* col_ptr, uint32_t* row_idx, float* w_out) {
  if (!ids_in_range(V, E, src, dst)) return -2;
  return build_grouped(V, E, src, dst, w, symmetrize, 1, col_ptr, row_idx, w_out);
}

int orc_build_csr(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                  int symmetrize, uint64_t* row_ptr, uint32_t* col_idx) {
  if (!ids_in_range(V, E, src, dst)) return -2;
  return build_grouped(V, E, src, dst, NULL, symmetrize, 0, row_ptr, col_idx, NULL);
}

void orc_out_degrees(uint32_t V, uint64_t nnz, const uint32_t* row_idx, uint32_t* od) {
  memset(od, 0, sizeof(uint32_t) * V);
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 65536)
#endif
  for (uint64_t e = 0; e < nnz; ++e) {
#ifdef _OPENMP
#pragma omp atomic
#endif
    od[row_idx[e]]++;
  }
}

/* ---- rank algorithms ------------------------------------------------------ */

#define ORC_CHUNK 4096u

/* Deterministic parallel sum: fixed 4096-vertex chunks, summed in chunk order. */
static double chunked_sum(const double* part, uint32_t nchunks) {
  double s = 0.0;
  for (uint32_t i = 0; i < nchunks; ++i) s += part[i];
  return s;
}

double orc_pagerank_step(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                         const uint32_t* od, double damping, int final_iter,
                         const double* x, double* y, uint32_t v_lo, uint32_t v_hi,
                         int nthreads) {
  (void)V;
  if (v_hi <= v_lo) return 0.0;
  const uint32_t nchunks = (v_hi - v_lo + ORC_CHUNK - 1) / ORC_CHUNK;
  double* part = (double*)calloc(nchunks, sizeof(double));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (uint32_t c = 0; c < nchunks; ++c) {
    const uint32_t lo = v_lo + c * ORC_CHUNK;
    const uint32_t hi = (lo + ORC_CHUNK < v_hi) ? lo + ORC_CHUNK : v_hi;
    double dl = 0.0;
    for (uint32_t v = lo; v < hi; ++v) {
      /* signal: rank_sum over the column in column order (pagerank.hpp:105-111) */
      double sum = 0.0;
      for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) sum += x[row_idx[e]];
      /* slot + apply (pagerank.hpp:112-135) */
      double r = 1.0 - damping + damping * sum;
      if (final_iter) {
        y[v] = r;
      } else if (od[v] > 0) {
        r = r / od[v];
        y[v] = r;
        dl += fabs(r - x[v]) * od[v];
      } else {
        y[v] = r;
        dl += fabs(r - x[v]);
      }
    }
    part[c] = dl;
  }
  const double delta = chunked_sum(part, nchunks);
  free(part);
  return final_iter ? 0.0 : delta;
}

int orc_pagerank(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                 const uint32_t* od, uint32_t iters, double damping, double eps,
                 double* out, double* deltas, int nthreads) {
  double* x = (double*)malloc(sizeof(double) * (V ? V : 1));
  double* y = (double*)malloc(sizeof(double) * (V ? V : 1));
  if (!x || !y) { free(x); free(y); return -1; }
  /* init (pagerank.hpp:90-98) */
  for (uint32_t v = 0; v < V; ++v) x[v] = od[v] > 0 ? 1.0 / od[v] : 1.0;
  int executed = 0;
  for (uint32_t t = 0; t < iters; ++t) {
    const int final_iter = (iters - 1 == t);
    const double delta = orc_pagerank_step(V, col_ptr, row_idx, od, damping, final_iter,
                                           x, y, 0, V, nthreads);
    if (deltas) deltas[executed] = delta;
    ++executed;
    /* early stop jumps to one more, final, iteration (pagerank.hpp:137-139) */
    if (!final_iter && eps > 0.0 && delta < eps) t = iters - 2;
    double* tmp = x; x = y; y = tmp;
  }
  memcpy(out, x, sizeof(double) * V);
  free(x); free(y);
  return executed;
}

int orc_weighted_pagerank(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                          const float* w, uint32_t iters, double damping, double eps,
                          double* out, double* deltas) {
  double* W = (double*)calloc(V ? V : 1, sizeof(double));
  double* x = (double*)malloc(sizeof(double) * (V ? V : 1));
  double* y = (double*)malloc(sizeof(double) * (V ? V : 1));
  if (!W || !x || !y) { free(W); free(x); free(y); return -1; }
  /* sum of out-weights per source, from the in-edges (weighted_pagerank.hpp:86-99) */
  for (uint32_t v = 0; v < V; ++v)
    for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) W[row_idx[e]] += (double)w[e];
  for (uint32_t v = 0; v < V; ++v) x[v] = 1.0;  /* (104-109) */
  int executed = 0;
  const uint32_t nchunks = (V + ORC_CHUNK - 1) / ORC_CHUNK;
  double* part = (double*)calloc(nchunks ? nchunks : 1, sizeof(double));
  for (uint32_t t = 0; t < iters; ++t) {
    const int final_iter = (iters - 1 == t);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (uint32_t c = 0; c < nchunks; ++c) {
      const uint32_t lo = c * ORC_CHUNK, hi = (lo + ORC_CHUNK < V) ? lo + ORC_CHUNK : V;
      double dl = 0.0;
      for (uint32_t v = lo; v < hi; ++v) {
        double sum = 0.0;
        for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) {
          const uint32_t u = row_idx[e];
          const double beta = ((double)w[e] * 1.0) / W[u];  /* (119-120) */
          sum += x[u] * beta;
        }
        const double r = 1.0 - damping + damping * sum;
        y[v] = r;
        dl += fabs(r - x[v]);  /* (138-143) */
      }
      part[c] = dl;
    }
    const double delta = final_iter ? 0.0 : chunked_sum(part, nchunks);
    if (deltas) deltas[executed] = delta;
    ++executed;
    if (!final_iter && eps > 0.0 && delta < eps) t = iters - 2;
    double* tmp = x; x = y; y = tmp;
  }
  memcpy(out, x, sizeof(double) * V);
  free(part); free(W); free(x); free(y);
  return executed;
}

double orc_weighted_pagerank_step(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                                  const float* w, const double* W, double damping,
                                  const double* x, double* y, uint32_t v_lo, uint32_t v_hi) {
  (void)V;
  double delta = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : delta)
#endif
  for (uint32_t v = v_lo; v < v_hi; ++v) {
    double sum = 0.0;
    for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) {
      const uint32_t u = row_idx[e];
      sum += x[u] * (((double)w[e] * 1.0) / W[u]);
    }
    const double r = 1.0 - damping + damping * sum;
    y[v] = r;
    delta += fabs(r - x[v]);
  }
  return delta;
}

/* seed membership as a byte map (a seed bitmap: duplicates count once) */
static uint8_t* seed_map(uint32_t V, const uint32_t* seeds, uint32_t nseeds, uint32_t* count) {
  uint8_t* m = (uint8_t*)calloc(V ? V : 1, 1);
  uint32_t c = 0;
  for (uint32_t i = 0; i < nseeds; ++i)
    if (seeds[i] < V && !m[seeds[i]]) { m[seeds[i]] = 1; ++c; }
  *count = c;
  return m;
}

/* apply of PPR / TrustRank (personalized_pagerank.hpp:101-120, trustrank.hpp:226-245) */
static double ppr_apply(uint32_t V, const uint32_t* od, const uint8_t* seed, double restart,
                        double damping, int final_iter, const double* x, double* y) {
  double delta = 0.0;
  for (uint32_t v = 0; v < V; ++v) {
    double r = (seed[v] ? restart : 0.0) + damping * y[v];
    if (final_iter) { y[v] = r; continue; }
    if (od[v] > 0) {
      r = r / od[v];
      y[v] = r;
      delta += fabs(r - x[v]) * od[v];
    } else {
      y[v] = r;
      delta += fabs(r - x[v]);
    }
  }
  return final_iter ? 0.0 : delta;
}

int orc_ppr(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
            const uint32_t* od, const uint32_t* seeds, uint32_t nseeds,
            uint32_t iters, double damping, double eps, double* out, double* deltas) {
  uint32_t ns = 0;
  uint8_t* seed = seed_map(V, seeds, nseeds, &ns);
  double* x = (double*)malloc(sizeof(double) * (V ? V : 1));
  double* y = (double*)malloc(sizeof(double) * (V ? V : 1));
  if (!seed || !x || !y || ns == 0) { free(seed); free(x); free(y); return -1; }
  /* seed set: x(v) = [v in S]/|S|/od(v), restart (1-d)/|S| (trustrank.hpp:197-206,
   * 230,238).  With one seed init_dist is exactly 1.0, which is the single-source
   * form x(src) = 1/od(src), restart 1-d (personalized_pagerank.hpp:71-81,105). */
  const double init_dist = 1.0 / ns;
  for (uint32_t v = 0; v < V; ++v) {
    x[v] = seed[v] ? init_dist : 0.0;
    if (od[v] > 0) x[v] = x[v] / od[v];
  }
  const double restart = (1.0 - damping) * init_dist;
  int executed = 0;
  for (uint32_t t = 0; t < iters; ++t) {
    const int final_iter = (iters - 1 == t);
    /* each column is summed by one thread in column order: thread-count independent */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096)
#endif
    for (uint32_t v = 0; v < V; ++v) {
      double sum = 0.0;
      for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) sum += x[row_idx[e]];
      y[v] = sum;
    }
    const double delta = ppr_apply(V, od, seed, restart, damping, final_iter, x, y);
    if (deltas) deltas[executed] = delta;
    ++executed;
    if (!final_iter && eps > 0.0 && delta < eps) t = iters - 2;
    double* tmp = x; x = y; y = tmp;
  }
  memcpy(out, x, sizeof(double) * V);
  free(seed); free(x); free(y);
  return executed;
}

int orc_ppr_pushpull(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
                     const uint64_t* row_ptr, const uint32_t* col_idx,
                     const uint32_t* od, uint64_t edges_prime, uint32_t src,
                     uint32_t iters, double damping, double eps,
                     double* out, double* deltas, uint8_t* modes) {
  if (src >= V) return -1;
  uint8_t* seed = (uint8_t*)calloc(V ? V : 1, 1);
  uint8_t* active = (uint8_t*)calloc(V ? V : 1, 1);
  double* x = (double*)calloc(V ? V : 1, sizeof(double));
  double* y = (double*)malloc(sizeof(double) * (V ? V : 1));
  if (!seed || !active || !x || !y) { free(seed); free(active); free(x); free(y); return -1; }
  seed[src] = 1;
  x[src] = 1.0;
  if (od[src] > 0) x[src] = x[src] / od[src];
  active[src] = 1;
  int do_pull = 0, executed = 0;
  for (uint32_t t = 0; t < iters; ++t) {
    const int final_iter = (iters - 1 == t);
    if (!do_pull) {  /* (188-205) */
      uint64_t ae = 0;
      for (uint32_t v = 0; v < V; ++v) if (active[v]) ae += od[v];
      do_pull = (ae > edges_prime / 20u);
    }
    if (do_pull) {
      for (uint32_t v = 0; v < V; ++v) {
        double sum = 0.0;
        for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) sum += x[row_idx[e]];
        y[v] = sum;
      }
    } else {  /* push over the accumulated active set (222-239) */
      memset(y, 0, sizeof(double) * V);
      /* snapshot: vertices activated in this epoch carry x = 0 and add nothing */
      uint8_t* was = (uint8_t*)malloc(V ? V : 1);
      memcpy(was, active, V);
      for (uint32_t v = 0; v < V; ++v) {
        if (!was[v]) continue;
        for (uint64_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
          const uint32_t w = col_idx[e];
          active[w] = 1;
          y[w] += x[v];
        }
      }
      free(was);
    }
    if (modes) modes[executed] = (uint8_t)(do_pull ? 0 : 1);
    const double delta = ppr_apply(V, od, seed, 1.0 - damping, damping, final_iter, x, y);
    if (deltas) deltas[executed] = delta;
    ++executed;
    if (!final_iter && eps > 0.0 && delta < eps) t = iters - 2;
    double* tmp = x; x = y; y = tmp;
  }
  memcpy(out, x, sizeof(double) * V);
  free(seed); free(active); free(x); free(y);
  return executed;
}

/* ---- BFS -------------------------------------------------------------------- */

int orc_bfs(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx,
            const uint64_t* row_ptr, const uint32_t* col_idx, const uint32_t* od,
            uint64_t edges_prime, double threshold, uint32_t root,
            int32_t* levels, uint8_t* modes, uint64_t* frontier, uint64_t* active_edges,
            uint32_t max_epochs, uint32_t* visited) {
  if (root >= V) return -1;
  uint8_t* cur = (uint8_t*)calloc(V ? V : 1, 1);
  uint8_t* nxt = (uint8_t*)calloc(V ? V : 1, 1);
  if (!cur || !nxt) { free(cur); free(nxt); return -1; }
  for (uint32_t v = 0; v < V; ++v) levels[v] = -1;
  levels[root] = 0;
  cur[root] = 1;
  uint64_t nactive = 1, nvisited = 1;
  int epoch = 0;
  while (nactive != 0) {
    uint64_t A = 0;  /* dualmode.hpp:264-276 */
    for (uint32_t v = 0; v < V; ++v) if (cur[v]) A += od[v];
    const int push = ((double)A / (double)edges_prime) < threshold;  /* (279) */
    memset(nxt, 0, V);
    uint64_t claimed = 0;
    if (push) {  /* bfs.hpp:98-118 */
      for (uint32_t v = 0; v < V; ++v) {
        if (!cur[v]) continue;
        for (uint64_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
          const uint32_t w = col_idx[e];
          if (levels[w] < 0) { levels[w] = epoch + 1; nxt[w] = 1; ++claimed; }
        }
      }
    } else {  /* bfs.hpp:119-136: unvisited columns take their first active in-neighbour */
      for (uint32_t v = 0; v < V; ++v) {
        if (levels[v] >= 0) continue;
        for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) {
          if (cur[row_idx[e]]) { levels[v] = epoch + 1; nxt[v] = 1; ++claimed; break; }
        }
      }
    }
    if ((uint32_t)epoch < max_epochs) {
      if (modes) modes[epoch] = (uint8_t)push;
      if (frontier) frontier[epoch] = nactive;
      if (active_edges) active_edges[epoch] = A;
    }
    nvisited += claimed;
    nactive = claimed;
    uint8_t* t = cur; cur = nxt; nxt = t;
    ++epoch;
  }
  if (visited) *visited = (uint32_t)nvisited;
  free(cur); free(nxt);
  return epoch;
}

/* ---- top-k ----------------------------------------------------------------- */

static const double* g_rank;
static int topk_cmp(const void* a, const void* b) {
  const uint32_t i = *(const uint32_t*)a, j = *(const uint32_t*)b;
  if (g_rank[i] > g_rank[j]) return -1;
  if (g_rank[i] < g_rank[j]) return 1;
  return (i < j) ? -1 : (i > j);
}

void orc_topk(uint32_t V, const double* r, uint32_t k, uint32_t* ids, double* vals) {
  uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
  for (uint32_t v = 0; v < V; ++v) p[v] = v;
  g_rank = r;
  qsort(p, V, sizeof(uint32_t), topk_cmp);
  if (k > V) k = V;
  for (uint32_t i = 0; i < k; ++i) { ids[i] = p[i]; vals[i] = r[p[i]]; }
  free(p);
}

/* ---- k-core (kcore.hpp:80-268) ------------------------------------------------ */
int orc_kcore(uint32_t V, const uint64_t* ptr, const uint32_t* idx, uint32_t* coreness,
              uint64_t* epoch_modified, uint64_t* epoch_next, uint32_t cap) {
  uint32_t* est = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
  uint8_t* act = (uint8_t*)malloc(V ? V : 1);
  uint8_t* nxt = (uint8_t*)malloc(V ? V : 1);
  uint64_t maxlen = 0;
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t len = ptr[v + 1] - ptr[v];
    coreness[v] = len ? (uint32_t)len : 0xFFFFFFFFu;  /* kcore.hpp:128-136 */
    if (len > maxlen) maxlen = len;
    act[v] = 1;                                        /* active_current.fill() */
  }
  uint32_t* count = (uint32_t*)malloc(sizeof(uint32_t) * (maxlen + 1));
  int epoch = 0;
  uint64_t need = 0;
  do {
    /* broadcast: estimate = coreness (kcore.hpp:167-189; on one partition both modes
       leave every estimate equal to its coreness) */
    memcpy(est, coreness, sizeof(uint32_t) * V);
    memset(nxt, 0, V);
    uint64_t modified = 0;
    for (uint32_t v = 0; v < V; ++v) {                 /* kcore.hpp:204-244 */
      if (!act[v]) continue;
      const uint64_t b = ptr[v], e = ptr[v + 1];
      const uint32_t len = (uint32_t)(e - b);
      if (len == 0) {
        coreness[v] = 0;
        continue;
      }
      const uint32_t cap_i = len < est[v] ? len : est[v];
      memset(count, 0, sizeof(uint32_t) * (cap_i + 1));
      for (uint64_t k = b; k < e; ++k) {
        const uint32_t eu = est[idx[k]];
        ++count[eu < cap_i ? eu : cap_i];
      }
      uint32_t sum = 0;
      for (uint32_t i = cap_i; i > 0; --i) {
        sum += count[i];
        if (sum >= i) {
          if (i < coreness[v]) {
            coreness[v] = i;
            for (uint64_t k = b; k < e; ++k) nxt[idx[k]] = 1;
            ++modified;
          }
          break;
        }
      }
    }
    need = 0;
    for (uint32_t v = 0; v < V; ++v) need += nxt[v];
    if ((uint32_t)epoch < cap) {
      epoch_modified[epoch] = modified;
      epoch_next[epoch] = need;
    }
    uint8_t* t = act;
    act = nxt;
    nxt = t;
    ++epoch;
  } while (need);                                      /* kcore.hpp:264 */
  free(count);
  free(nxt);
  free(act);
  free(est);
  return epoch;
}

/* ---- connected components (cgm/connected_component.hpp:170-234) --------------- */
void orc_cc_labels(uint32_t V, const uint64_t* row_ptr, const uint32_t* col_idx, uint32_t* labels) {
  uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
  for (uint32_t v = 0; v < V; ++v) labels[v] = 0xFFFFFFFFu;
  for (uint32_t u = 0; u < V; ++u) {
    if (labels[u] != 0xFFFFFFFFu) continue;
    uint64_t top = 0;
    labels[u] = u;
    stack[top++] = u;
    while (top) {
      const uint32_t w = stack[--top];
      for (uint64_t e = row_ptr[w]; e < row_ptr[w + 1]; ++e) {
        const uint32_t x = col_idx[e];
        if (labels[x] != 0xFFFFFFFFu) continue;
        labels[x] = u;
        stack[top++] = x;
      }
    }
  }
  free(stack);
}

/* ---- HyperANF (hyperanf.hpp:80-197, hyperloglog.hpp:32-142) -------------------- */
static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

uint32_t orc_murmur3_32(const void* key, uint32_t len, uint32_t seed) {
  const uint8_t* data = (const uint8_t*)key;
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  uint32_t h = seed;
  const uint32_t nblocks = len / 4;
  for (uint32_t i = 0; i < nblocks; ++i) {
    uint32_t k;
    memcpy(&k, data + 4 * i, 4);  /* little-endian blocks */
    k *= c1;
    k = rotl32(k, 15);
    k *= c2;
    h ^= k;
    h = rotl32(h, 13);
    h = h * 5 + 0xe6546b64u;
  }
  const uint8_t* tail = data + 4 * nblocks;
  uint32_t k = 0;
  switch (len & 3) {
    case 3: k ^= (uint32_t)tail[2] << 16; /* fall through */
    case 2: k ^= (uint32_t)tail[1] << 8;  /* fall through */
    case 1:
      k ^= tail[0];
      k *= c1;
      k = rotl32(k, 15);
      k *= c2;
      h ^= k;
  }
  h ^= len;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

double orc_hll_estimate(const uint8_t* regs, uint32_t P) {
  const uint32_t m = 1u << P;
  double alpha_mm;
  switch (P) {
    case 4: alpha_mm = 0.673; break;
    case 5: alpha_mm = 0.697; break;
    case 6: alpha_mm = 0.709; break;
    default: alpha_mm = 0.7213 / (1.0 + 1.079 / (double)m); break;
  }
  alpha_mm *= (double)m * (double)m;
  double sum = 0.0;
  for (uint32_t i = 0; i < m; ++i) sum += 1.0 / (double)(1u << regs[i]);
  double est = alpha_mm / sum;
  if (est <= 5.0 / 2.0 * (double)m) {
    uint32_t zeros = 0;
    for (uint32_t i = 0; i < m; ++i) zeros += regs[i] == 0;
    if (zeros) est = (double)m * log((double)m / zeros);
  } else if (est > (1.0 / 30.0) * 4294967296.0) {
    est = -4294967296.0 * log(1.0 - (est / 4294967296.0));
  }
  return est;
}

int orc_hyperanf(uint32_t V, const uint64_t* col_ptr, const uint32_t* row_idx, uint32_t P, uint32_t iters,
                 uint8_t* regs, double* anfs, uint32_t cap) {
  const uint64_t m = 1ull << P;
  uint8_t* nxt = (uint8_t*)malloc((size_t)V * m + 1);
  memset(regs, 0, (size_t)V * m);
  double anf = 0.0;
  for (uint32_t v = 0; v < V; ++v) {
    const uint32_t h = orc_murmur3_32(&v, 4, 0x5f61767au);
    const uint32_t idx = h >> (32 - P);
    uint32_t w = h << P;
    uint8_t rank = 1;
    while (rank <= 32 - P && !(w & 0x80000000u)) {
      ++rank;
      w <<= 1;
    }
    if (rank > regs[(size_t)v * m + idx]) regs[(size_t)v * m + idx] = rank;
    anf += orc_hll_estimate(regs + (size_t)v * m, P);
  }
  int n = 0;
  if ((uint32_t)n < cap) anfs[n] = anf;
  ++n;
  double last = anf;
  for (uint32_t it = 0; it < iters; ++it) {
    memcpy(nxt, regs, (size_t)V * m);
    for (uint32_t v = 0; v < V; ++v) {
      uint8_t* dst = nxt + (size_t)v * m;
      for (uint64_t e = col_ptr[v]; e < col_ptr[v + 1]; ++e) {
        const uint8_t* src = regs + (size_t)row_idx[e] * m;
        for (uint64_t i = 0; i < m; ++i)
          if (src[i] > dst[i]) dst[i] = src[i];
      }
    }
    memcpy(regs, nxt, (size_t)V * m);
    anf = 0.0;
    for (uint32_t v = 0; v < V; ++v) anf += orc_hll_estimate(regs + (size_t)v * m, P);
    if (anf - last < DBL_EPSILON) break;
    if ((uint32_t)n < cap) anfs[n] = anf;
    ++n;
    last = anf;
  }
  free(nxt);
  return n;
}

int orc_nstep_degrees(uint32_t V, const uint64_t* ptr, const uint32_t* idx, uint32_t P, uint32_t step,
                      uint32_t* degree) {
  const uint64_t m = 1ull << P;
  uint8_t* regs = (uint8_t*)calloc((size_t)V * m + 1, 1);
  uint8_t* nxt = (uint8_t*)malloc((size_t)V * m + 1);
  if (!regs || !nxt) { free(regs); free(nxt); return -1; }
  for (uint32_t v = 0; v < V; ++v) {  /* books[v].add(&v, 4) (nstepdegrees.hpp:331-336) */
    const uint32_t h = orc_murmur3_32(&v, 4, 0x5f61767au);
    const uint32_t r0 = h >> (32 - P);
    uint32_t w = h << P;
    uint8_t rank = 1;
    while (rank <= 32 - P && !(w & 0x80000000u)) { ++rank; w <<= 1; }
    regs[(size_t)v * m + r0] = rank;
  }
  int executed = 0;
  for (uint32_t it = 0; it < step; ++it) {  /* (343-409): stops when no counter grows */
    memcpy(nxt, regs, (size_t)V * m);
    int changed = 0;
    for (uint32_t v = 0; v < V; ++v) {
      uint8_t* dst = nxt + (size_t)v * m;
      for (uint64_t e = ptr[v]; e < ptr[v + 1]; ++e) {
        const uint8_t* src = regs + (size_t)idx[e] * m;
        for (uint64_t i = 0; i < m; ++i)
          if (src[i] > dst[i]) { dst[i] = src[i]; changed = 1; }
      }
    }
    uint8_t* t = regs; regs = nxt; nxt = t;
    ++executed;
    if (!changed) break;
  }
  for (uint32_t v = 0; v < V; ++v)  /* static_cast<vid_t>(estimate()) - 1 (412-421) */
    degree[v] = (uint32_t)orc_hll_estimate(regs + (size_t)v * m, P) - 1u;
  free(regs); free(nxt);
  return executed;
}

double orc_anf_avg_distance(const double* anfs, uint32_t n) {
  double avg = 0.0;
  for (uint32_t i = 1; i < n; ++i) avg += i * (anfs[i] - anfs[i - 1]);
  return avg / (anfs[n - 1] - anfs[0]);
}

/* ---- edge-list ingestion (parsers.hpp:225-293) ------------------------------ */
int64_t orc_parse_csv(const char* bytes, uint64_t n, int weighted, uint64_t cap, uint32_t* src,
                      uint32_t* dst, float* w) {
  uint64_t pos = 0, count = 0, line = 0;
  char* buf = (char*)malloc(n + 1);
  while (pos < n) {
    uint64_t e = pos;
    while (e < n && bytes[e] != '\n') ++e;
    const uint64_t len = e - pos;
    memcpy(buf, bytes + pos, len);
    buf[len] = '\0';
    pos = e + 1;  /* past the '\n' (or past the end) */
    const uint64_t this_line = line++;
    if (buf[0] == '\0') continue;
    char* save = NULL;
    char* tok = strtok_r(buf, ",", &save);
    if (!tok) continue;
    const unsigned long s = strtoul(tok, NULL, 10);
    if (s > 0xFFFFFFFFul) { free(buf); return -(int64_t)this_line - 1; }
    tok = strtok_r(NULL, ",", &save);
    if (!tok) continue;
    const unsigned long d = strtoul(tok, NULL, 10);
    if (d > 0xFFFFFFFFul) { free(buf); return -(int64_t)this_line - 1; }
    float wt = 0.0f;
    if (weighted) {
      tok = strtok_r(NULL, ",", &save);
      if (!tok) continue;
      wt = strtof(tok, NULL);
    }
    if (count < cap) {
      src[count] = (uint32_t)s;
      dst[count] = (uint32_t)d;
      if (weighted && w) w[count] = wt;
    }
    ++count;
  }
  free(buf);
  return (int64_t)count;
}

static int orc_u32_digits(uint32_t x, char* tmp) {
  int n = 0;
  do { tmp[n++] = (char)('0' + x % 10u); x /= 10u; } while (x);
  return n;
}

int64_t orc_format_csv(uint64_t E, const uint32_t* src, const uint32_t* dst, char* out, uint64_t cap) {
  const uint64_t B = 1u << 16;
  const uint64_t nb = (E + B - 1) / B;
  uint64_t* len = (uint64_t*)calloc(nb + 1, sizeof(uint64_t));
  #pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < (int64_t)nb; ++b) {
    char tmp[12];
    uint64_t L = 0;
    for (uint64_t i = (uint64_t)b * B; i < E && i < ((uint64_t)b + 1) * B; ++i)
      L += (uint64_t)orc_u32_digits(src[i], tmp) + (uint64_t)orc_u32_digits(dst[i], tmp) + 2;
    len[b + 1] = L;
  }
  for (uint64_t b = 0; b < nb; ++b) len[b + 1] += len[b];
  if (len[nb] > cap) { free(len); return -1; }
  #pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < (int64_t)nb; ++b) {
    char tmp[12];
    char* o = out + len[b];
    for (uint64_t i = (uint64_t)b * B; i < E && i < ((uint64_t)b + 1) * B; ++i) {
      int n = orc_u32_digits(src[i], tmp);
      while (n) *o++ = tmp[--n];
      *o++ = ',';
      n = orc_u32_digits(dst[i], tmp);
      while (n) *o++ = tmp[--n];
      *o++ = '\n';
    }
  }
  const int64_t total = (int64_t)len[nb];
  free(len);
  return total;
}
