// Edge-list ingestion on the device: CSV text -> edge arrays -> (optional id
// compaction) -> CSC/CSR build.  Replaces the reference's csv_parser
// (plato/graph/parsers.hpp:225-293) feeding the two BSP builds, and vid_encoder_t
// (plato/graph/structure/vid_encoder.hpp:41-102).
//
// Lines: '\n' positions are counted per 64 KiB segment, scanned, and written in
// order; line i runs from the byte after newline i-1 to newline i (the text after
// the last newline is a line when non-empty), as std::istream::getline splits.
// Each line is parsed by one thread with the reference's rules: fields are the
// maximal runs of non-',' bytes (strtok_r), ids are read like strtoul(, 10) and
// weights like strtof; an empty line, or a line missing a field, is skipped.
// Valid lines are compacted in file order.
#include <algorithm>
#include <string>
#include <vector>

#include "pe_internal.hpp"

namespace pe {

static constexpr uint64_t CSV_SEG = 1ull << 16;  // bytes per newline-count segment

// ---- lines ----------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_csv_count(const char* __restrict__ b, uint64_t n,
                                                    uint32_t* __restrict__ seg_count) {
  __shared__ uint32_t s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const uint64_t p0 = (uint64_t)blockIdx.x * CSV_SEG, p1 = std::min<uint64_t>(n, p0 + CSV_SEG);
  uint32_t c = 0;
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) c += b[p] == '\n' ? 1u : 0u;
  atomicAdd(&s_cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) seg_count[blockIdx.x] = s_cnt;
}

// Positions of the newlines in order; seg_off is the exclusive scan of seg_count.
__global__ __launch_bounds__(1024) void k_csv_newlines(const char* __restrict__ b, uint64_t n,
                                                       const uint64_t* __restrict__ seg_off,
                                                       uint64_t* __restrict__ nl) {
  __shared__ uint32_t s_wave[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t p0 = (uint64_t)blockIdx.x * CSV_SEG, p1 = std::min<uint64_t>(n, p0 + CSV_SEG);
  uint64_t out = seg_off[blockIdx.x];
  for (uint64_t base = p0; base < p1; base += 1024) {
    const uint64_t p = base + threadIdx.x;
    const bool hit = p < p1 && b[p] == '\n';
    const uint64_t m = __ballot(hit);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int i = 0; i < 16; ++i) {
      if (i < wave) before += s_wave[i];
      total += s_wave[i];
    }
    if (hit) nl[out + before + (uint64_t)__popcll(m & ((1ull << lane) - 1ull))] = p;
    out += total;
    __syncthreads();
  }
}

// ---- fields ---------------------------------------------------------------------------------
__device__ __forceinline__ bool csv_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}
__device__ __forceinline__ bool csv_digit(char c) { return c >= '0' && c <= '9'; }

// Next field of [p, e): skips ',' bytes, returns false when none is left; the field
// is [*fs, *fe).
__device__ __forceinline__ bool csv_field(const char* b, uint64_t& p, uint64_t e, uint64_t& fs, uint64_t& fe) {
  while (p < e && b[p] == ',') ++p;
  if (p >= e) return false;
  fs = p;
  while (p < e && b[p] != ',') ++p;
  fe = p;
  return true;
}

// strtoul(field, nullptr, 10): leading whitespace, optional sign, decimal digits up to
// the first other byte; no digits -> 0; overflow -> 2^64 - 1; '-' negates modulo 2^64.
__device__ uint64_t csv_strtoul(const char* b, uint64_t p, uint64_t e) {
  while (p < e && csv_space(b[p])) ++p;
  bool neg = false;
  if (p < e && (b[p] == '+' || b[p] == '-')) {
    neg = b[p] == '-';
    ++p;
  }
  uint64_t v = 0;
  bool any = false, ovf = false;
  for (; p < e && csv_digit(b[p]); ++p) {
    any = true;
    const uint64_t d = (uint64_t)(b[p] - '0');
    if (v > (~0ull - d) / 10ull) ovf = true;
    else v = v * 10ull + d;
  }
  if (!any) return 0ull;
  if (ovf) return ~0ull;
  return neg ? 0ull - v : v;
}

__constant__ double c_pow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                   1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
__constant__ uint64_t c_pow10u[20] = {1ull,
                                      10ull,
                                      100ull,
                                      1000ull,
                                      10000ull,
                                      100000ull,
                                      1000000ull,
                                      10000000ull,
                                      100000000ull,
                                      1000000000ull,
                                      10000000000ull,
                                      100000000000ull,
                                      1000000000000ull,
                                      10000000000000ull,
                                      100000000000000ull,
                                      1000000000000000ull,
                                      10000000000000000ull,
                                      100000000000000000ull,
                                      1000000000000000000ull,
                                      10000000000000000000ull};

using u128i = unsigned __int128;

// 10^k as a 128-bit integer (k <= 38)
__device__ __forceinline__ u128i pow10_128(int k) {
  u128i r = 1;
  for (int i = 0; i < k; ++i) r *= 10u;
  return r;
}

// strtof(field) for decimal literals (strtof semantics: leading whitespace, sign,
// digits with an optional '.', an optional exponent; no digits -> 0).  The value is
// D * 10^e10 with D < 2^53 and |e10| <= 22, so one double multiply or divide of exact
// operands is correctly rounded (Clinger, "How to read floating point numbers
// accurately", PLDI 1990); rounding that double to float is then correct unless the
// double sits exactly on a float midpoint, which is settled by comparing D * 10^e10
// with the midpoint in 128-bit integers.  Exponents beyond 22 take up to four
// correctly rounded operations and are accepted when the result is clear of a float
// midpoint.
// inf, nan, hex, more than 15 significant digits, an exponent beyond that range, a
// result next to a midpoint, or a float result that is subnormal or infinite set *bad.
__device__ float csv_strtof(const char* b, uint64_t p, uint64_t e, bool& bad) {
  while (p < e && csv_space(b[p])) ++p;
  bool neg = false;
  if (p < e && (b[p] == '+' || b[p] == '-')) {
    neg = b[p] == '-';
    ++p;
  }
  if (p < e && (b[p] == 'i' || b[p] == 'I' || b[p] == 'n' || b[p] == 'N')) {
    bad = true;
    return 0.0f;
  }
  if (p + 1 < e && b[p] == '0' && (b[p + 1] == 'x' || b[p + 1] == 'X')) {
    bad = true;
    return 0.0f;
  }
  uint64_t D = 0;
  int64_t e10 = 0, pending = 0;
  bool any = false;
  auto digit = [&](uint32_t c, bool frac) {
    any = true;
    if (frac) e10 -= 1;
    if (c == 0) {
      if (D) ++pending;
      return;
    }
    if (pending + 1 > 15) {
      bad = true;
      return;
    }
    const uint64_t mul = c_pow10u[pending + 1];
    if (D > ((1ull << 53) - c) / mul) {
      bad = true;
      return;
    }
    D = D * mul + c;
    pending = 0;
  };
  for (; p < e && csv_digit(b[p]); ++p) digit((uint32_t)(b[p] - '0'), false);
  if (p < e && b[p] == '.') {
    ++p;
    for (; p < e && csv_digit(b[p]); ++p) digit((uint32_t)(b[p] - '0'), true);
  }
  if (!any) return 0.0f;  // no conversion
  if (p < e && (b[p] == 'e' || b[p] == 'E')) {
    uint64_t q = p + 1;
    bool eneg = false;
    if (q < e && (b[q] == '+' || b[q] == '-')) {
      eneg = b[q] == '-';
      ++q;
    }
    if (q < e && csv_digit(b[q])) {
      int64_t x = 0;
      for (; q < e && csv_digit(b[q]); ++q) x = x < 100000 ? x * 10 + (b[q] - '0') : x;
      e10 += eneg ? -x : x;
    }
  }
  if (bad) return 0.0f;
  e10 += pending;
  if (D == 0) return neg ? -0.0f : 0.0f;
  if (e10 > 22 || e10 < -22) {
    // at most four correctly rounded double operations by powers of ten <= 10^22:
    // |d - x| <= 2 ulp(d), which decides the float rounding unless d lies within a few
    // double ulps of a float midpoint
    if (e10 > 60 || e10 < -66) {
      bad = true;
      return 0.0f;
    }
    double d = (double)D;
    int64_t k = e10;
    while (k > 22) {
      d *= c_pow10[22];
      k -= 22;
    }
    while (k < -22) {
      d /= c_pow10[22];
      k += 22;
    }
    d = k >= 0 ? d * c_pow10[k] : d / c_pow10[-k];
    const float f = (float)d;
    const int64_t low = (int64_t)((uint64_t)__double_as_longlong(d) & ((1ull << 29) - 1ull)) - (1ll << 28);
    if (!(fabsf(f) >= 1.17549435e-38f) || isinf(f) || (low >= -16 && low <= 16)) {
      bad = true;
      return 0.0f;
    }
    return neg ? -f : f;
  }
  const double d = e10 >= 0 ? (double)D * c_pow10[e10] : (double)D / c_pow10[-e10];
  float f = (float)d;
  if (!(fabsf(f) >= 1.17549435e-38f) || isinf(f)) {  // subnormal or out of float range
    bad = true;
    return 0.0f;
  }
  // exact float midpoint? (the 29 low mantissa bits of d equal 1 << 28)
  const uint64_t db = (uint64_t)__double_as_longlong(d);
  if ((db & ((1ull << 29) - 1ull)) == (1ull << 28)) {
    const int ex = (int)((db >> 52) & 0x7FF) - 1075;  // d = Mm * 2^ex
    const u128i Mm = (u128i)((db & ((1ull << 52) - 1ull)) | (1ull << 52));
    int cmp;  // sign of (D * 10^e10 - d)
    if (e10 >= 0) {
      const u128i X = (u128i)D * pow10_128((int)e10);
      const u128i M = ex >= 0 ? (Mm << ex) : Mm;
      const u128i Xs = ex >= 0 ? X : (X << (-ex));
      cmp = Xs > M ? 1 : (Xs < M ? -1 : 0);
    } else {
      const u128i R = Mm * pow10_128((int)-e10);  // d * 10^k = Mm * 10^k * 2^ex
      const u128i L = ex >= 0 ? (u128i)D : ((u128i)D << (-ex));
      const u128i Rs = ex >= 0 ? (R << ex) : R;
      cmp = L > Rs ? 1 : (L < Rs ? -1 : 0);
    }
    if (cmp != 0) {
      const uint32_t lo = (uint32_t)((db >> 29) & 0x7FFFFFull);  // float just below d
      const uint32_t fexp = (uint32_t)(((db >> 52) & 0x7FF) - 1023 + 127);
      const uint32_t below = (fexp << 23) | lo;
      f = __uint_as_float(cmp > 0 ? below + 1u : below);
    }
  }
  return neg ? -f : f;
}

// One line per thread.  status: 0 skipped, 1 edge, 2 src or dst above the id limit,
// 3 weight literal outside the exact range.
__global__ void k_csv_parse(const char* __restrict__ b, uint64_t n, const uint64_t* __restrict__ nl, uint64_t nnl,
                            uint64_t nlines, int weighted, uint64_t id_limit, uint8_t* __restrict__ status,
                            uint64_t* __restrict__ src, uint64_t* __restrict__ dst, float* __restrict__ w) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nlines;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t p = i ? nl[i - 1] + 1 : 0;
    uint64_t e = i < nnl ? nl[i] : n;
    // a NUL byte ends the line's string
    for (uint64_t q = p; q < e; ++q)
      if (b[q] == '\0') {
        e = q;
        break;
      }
    uint8_t st = 0;
    uint64_t s = 0, d = 0;
    float wt = 0.0f;
    uint64_t fs, fe;
    if (p < e && csv_field(b, p, e, fs, fe)) {
      s = csv_strtoul(b, fs, fe);
      if (s > id_limit) {
        st = 2;
      } else if (csv_field(b, p, e, fs, fe)) {
        d = csv_strtoul(b, fs, fe);
        if (d > id_limit) {
          st = 2;
        } else if (!weighted) {
          st = 1;
        } else if (csv_field(b, p, e, fs, fe)) {
          bool bad = false;
          wt = csv_strtof(b, fs, fe, bad);
          st = bad ? 3 : 1;
        }
      }
    }
    status[i] = st;
    src[i] = s;
    dst[i] = d;
    if (weighted) w[i] = wt;
  }
}

// ---- compaction of the edges in file order ---------------------------------------------------
__global__ __launch_bounds__(1024) void k_csv_count_ok(const uint8_t* __restrict__ status, uint64_t nlines,
                                                       uint32_t* __restrict__ seg_count) {
  __shared__ uint32_t s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const uint64_t p0 = (uint64_t)blockIdx.x * CSV_SEG, p1 = std::min<uint64_t>(nlines, p0 + CSV_SEG);
  uint32_t c = 0;
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) c += status[p] == 1 ? 1u : 0u;
  atomicAdd(&s_cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) seg_count[blockIdx.x] = s_cnt;
}

__global__ __launch_bounds__(1024) void k_csv_compact(const uint8_t* __restrict__ status, uint64_t nlines,
                                                      const uint64_t* __restrict__ seg_off,
                                                      const uint64_t* __restrict__ src,
                                                      const uint64_t* __restrict__ dst,
                                                      const float* __restrict__ w, uint64_t* __restrict__ osrc,
                                                      uint64_t* __restrict__ odst, float* __restrict__ ow) {
  __shared__ uint32_t s_wave[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t p0 = (uint64_t)blockIdx.x * CSV_SEG, p1 = std::min<uint64_t>(nlines, p0 + CSV_SEG);
  uint64_t out = seg_off[blockIdx.x];
  for (uint64_t base = p0; base < p1; base += 1024) {
    const uint64_t p = base + threadIdx.x;
    const bool ok = p < p1 && status[p] == 1;
    const uint64_t m = __ballot(ok);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int i = 0; i < 16; ++i) {
      if (i < wave) before += s_wave[i];
      total += s_wave[i];
    }
    if (ok) {
      const uint64_t o = out + before + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
      osrc[o] = src[p];
      if (odst) odst[o] = dst[p];
      if (w) ow[o] = w[p];
    }
    out += total;
    __syncthreads();
  }
}

// ---- ids ----------------------------------------------------------------------------------------
__global__ void k_u64_concat(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, uint64_t n,
                             uint64_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    out[i] = a[i];
    out[n + i] = b[i];
  }
}

// distinct values of a sorted array: flag the first of each run
__global__ void k_u64_heads(const uint64_t* __restrict__ k, uint64_t n, uint8_t* __restrict__ head) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    head[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

// dense id of each original id: its index in the sorted distinct ids
__global__ void k_encode(const uint64_t* __restrict__ ids, uint64_t n, const uint64_t* __restrict__ uniq,
                         uint64_t nu, uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = ids[i];
    uint64_t lo = 0, hi = nu - 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (uniq[mid] < x) lo = mid + 1;
      else hi = mid;
    }
    out[i] = (uint32_t)lo;
  }
}

__global__ void k_u64_to_u32(const uint64_t* __restrict__ a, uint64_t n, uint32_t* __restrict__ out,
                             unsigned long long* __restrict__ maxv) {
  uint64_t m = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    out[i] = (uint32_t)a[i];
    m = a[i] > m ? a[i] : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(maxv, (unsigned long long)m);
}

// Two-level ordered compaction helper on the host side of a segment count.
static std::vector<uint64_t> exclusive_offsets(pe_ctx_s* ctx, const uint32_t* dcount, uint64_t nseg, uint64_t& total) {
  std::vector<uint32_t> h(nseg);
  PE_HIP(hipMemcpyAsync(h.data(), dcount, nseg * 4, hipMemcpyDeviceToHost, ctx->stream));
  PE_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<uint64_t> off(nseg);
  total = 0;
  for (uint64_t i = 0; i < nseg; ++i) {
    off[i] = total;
    total += h[i];
  }
  return off;
}

}  // namespace pe

using namespace pe;

extern "C" {

int pe_graph_create_csv(pe_ctx_t* ctx, const char* text, uint64_t nbytes, uint32_t flags, pe_graph_t** out) {
  if (!ctx || !out || (nbytes && !text)) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&] {
    hipStream_t s = ctx->stream;
    const bool weighted = (flags & PE_GRAPH_WEIGHTED) != 0;
    const bool encode = (flags & PE_GRAPH_ENCODE_IDS) != 0;
    // text on the device
    dbuf<char> db(ctx, std::max<uint64_t>(1, nbytes));
    if (nbytes) PE_HIP(hipMemcpyAsync(db.p, text, nbytes, hipMemcpyHostToDevice, s));
    // newlines
    const uint64_t nseg = std::max<uint64_t>(1, (nbytes + CSV_SEG - 1) / CSV_SEG);
    PE_REQUIRE(nseg < 0x7FFFFFFFull, PE_ERR_RANGE, "text too large");
    dbuf<uint32_t> cnt(ctx, nseg);
    uint64_t nnl = 0;
    if (nbytes) {
      k_csv_count<<<(unsigned)nseg, 1024, 0, s>>>(db.p, nbytes, cnt.p);
      PE_HIP(hipGetLastError());
    } else {
      PE_HIP(hipMemsetAsync(cnt.p, 0, 4, s));
    }
    auto off = exclusive_offsets(ctx, cnt.p, nseg, nnl);
    dbuf<uint64_t> doff(ctx, nseg), nl(ctx, std::max<uint64_t>(1, nnl));
    PE_HIP(hipMemcpyAsync(doff.p, off.data(), nseg * 8, hipMemcpyHostToDevice, s));
    if (nbytes) {
      k_csv_newlines<<<(unsigned)nseg, 1024, 0, s>>>(db.p, nbytes, doff.p, nl.p);
      PE_HIP(hipGetLastError());
    }
    // lines: one per newline, plus the text after the last newline when non-empty
    uint64_t last_nl = 0;
    if (nnl) PE_HIP(hipMemcpyAsync(&last_nl, nl.p + nnl - 1, 8, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    const uint64_t tail_start = nnl ? last_nl + 1 : 0;
    const uint64_t nlines = nnl + (tail_start < nbytes ? 1 : 0);
    // parse
    const uint64_t id_limit = encode ? ~0ull : 0xFFFFFFFFull;
    dbuf<uint8_t> status(ctx, std::max<uint64_t>(1, nlines));
    dbuf<uint64_t> ls(ctx, std::max<uint64_t>(1, nlines)), ld(ctx, std::max<uint64_t>(1, nlines));
    dbuf<float> lw;
    if (weighted) lw.alloc(ctx, std::max<uint64_t>(1, nlines));
    if (nlines) {
      k_csv_parse<<<grid_for(nlines, 256, 1u << 20), 256, 0, s>>>(db.p, nbytes, nl.p, nnl, nlines, weighted ? 1 : 0,
                                                                  id_limit, status.p, ls.p, ld.p, lw.p);
      PE_HIP(hipGetLastError());
    }
    db.release();
    nl.release();
    // errors: first offending line (host scan of the status bytes is enough: 1 B per line)
    std::vector<uint8_t> hst(nlines);
    if (nlines) PE_HIP(hipMemcpyAsync(hst.data(), status.p, nlines, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    for (uint64_t i = 0; i < nlines; ++i) {
      if (hst[i] == 2)
        fail(PE_ERR_RANGE, "line " + std::to_string(i + 1) + ": vertex id above 2^32 - 1 (use PE_GRAPH_ENCODE_IDS)");
      if (hst[i] == 3)
        fail(PE_ERR_RANGE, "line " + std::to_string(i + 1) +
                               ": weight literal outside the exactly-rounded decimal range");
    }
    // compact the edges in file order
    const uint64_t lseg = std::max<uint64_t>(1, (nlines + CSV_SEG - 1) / CSV_SEG);
    dbuf<uint32_t> lcnt(ctx, lseg);
    uint64_t E = 0;
    if (nlines) {
      k_csv_count_ok<<<(unsigned)lseg, 1024, 0, s>>>(status.p, nlines, lcnt.p);
      PE_HIP(hipGetLastError());
    } else {
      PE_HIP(hipMemsetAsync(lcnt.p, 0, 4, s));
    }
    auto loff = exclusive_offsets(ctx, lcnt.p, lseg, E);
    dbuf<uint64_t> dloff(ctx, lseg), es(ctx, std::max<uint64_t>(1, E)), ed(ctx, std::max<uint64_t>(1, E));
    dbuf<float> ew;
    if (weighted) ew.alloc(ctx, std::max<uint64_t>(1, E));
    PE_HIP(hipMemcpyAsync(dloff.p, loff.data(), lseg * 8, hipMemcpyHostToDevice, s));
    if (nlines) {
      k_csv_compact<<<(unsigned)lseg, 1024, 0, s>>>(status.p, nlines, dloff.p, ls.p, ld.p, weighted ? lw.p : nullptr,
                                                    es.p, ed.p, weighted ? ew.p : nullptr);
      PE_HIP(hipGetLastError());
    }
    status.release();
    ls.release();
    ld.release();
    lw.release();
    PE_REQUIRE(E > 0, PE_ERR_INVALID, "no edges in the text");
    // ids
    auto* g = new pe_graph_s();
    std::unique_ptr<pe_graph_s> guard(g);
    g->ctx = ctx;
    g->edges = E;
    dbuf<uint32_t> ds(ctx, E), dd(ctx, E);
    if (encode) {
      dbuf<uint64_t> all(ctx, 2 * E);
      k_u64_concat<<<grid_for(E, 256, 1u << 20), 256, 0, s>>>(es.p, ed.p, E, all.p);
      PE_HIP(hipGetLastError());
      sort_keys(ctx, all, nullptr, 2 * E, 64);
      dbuf<uint8_t> head(ctx, 2 * E);
      k_u64_heads<<<grid_for(2 * E, 256, 1u << 20), 256, 0, s>>>(all.p, 2 * E, head.p);
      PE_HIP(hipGetLastError());
      const uint64_t hseg = (2 * E + CSV_SEG - 1) / CSV_SEG;
      dbuf<uint32_t> hcnt(ctx, hseg);
      k_csv_count_ok<<<(unsigned)hseg, 1024, 0, s>>>(head.p, 2 * E, hcnt.p);
      PE_HIP(hipGetLastError());
      uint64_t nu = 0;
      auto hoff = exclusive_offsets(ctx, hcnt.p, hseg, nu);
      PE_REQUIRE(nu <= 0xFFFFFFFFull, PE_ERR_RANGE, "more than 2^32 - 1 distinct vertex ids");
      dbuf<uint64_t> dhoff(ctx, hseg);
      PE_HIP(hipMemcpyAsync(dhoff.p, hoff.data(), hseg * 8, hipMemcpyHostToDevice, s));
      g->vid_decode.alloc(ctx, nu);
      k_csv_compact<<<(unsigned)hseg, 1024, 0, s>>>(head.p, 2 * E, dhoff.p, all.p, nullptr, nullptr, g->vid_decode.p,
                                                    nullptr, nullptr);
      PE_HIP(hipGetLastError());
      k_encode<<<grid_for(E, 256, 1u << 20), 256, 0, s>>>(es.p, E, g->vid_decode.p, nu, ds.p);
      k_encode<<<grid_for(E, 256, 1u << 20), 256, 0, s>>>(ed.p, E, g->vid_decode.p, nu, dd.p);
      PE_HIP(hipGetLastError());
      g->V = (uint32_t)nu;
      g->encoded = true;
    } else {
      dbuf<unsigned long long> mx(ctx, 1);
      PE_HIP(hipMemsetAsync(mx.p, 0, 8, s));
      k_u64_to_u32<<<grid_for(E, 256, 65536), 256, 0, s>>>(es.p, E, ds.p, mx.p);
      k_u64_to_u32<<<grid_for(E, 256, 65536), 256, 0, s>>>(ed.p, E, dd.p, mx.p);
      PE_HIP(hipGetLastError());
      unsigned long long m = 0;
      PE_HIP(hipMemcpyAsync(&m, mx.p, 8, hipMemcpyDeviceToHost, s));
      PE_HIP(hipStreamSynchronize(s));
      PE_REQUIRE(m < 0xFFFFFFFFull, PE_ERR_RANGE, "vertex id 2^32 - 1 leaves no room for V");
      g->V = (uint32_t)(m + 1);
    }
    es.release();
    ed.release();
    build_from_edges(g, ds, dd, ew, flags & ~PE_GRAPH_ENCODE_IDS);
    *out = guard.release();
  });
}

int pe_graph_vertex_ids(pe_graph_t* g, uint64_t* original_ids) {
  if (!g || !original_ids) return PE_ERR_INVALID;
  return guarded(g->ctx, [&] {
    if (!g->encoded) {
      for (uint32_t v = 0; v < g->V; ++v) original_ids[v] = v;
      return;
    }
    PE_HIP(hipMemcpyAsync(original_ids, g->vid_decode.p, (uint64_t)g->V * 8, hipMemcpyDeviceToHost, g->ctx->stream));
    PE_HIP(hipStreamSynchronize(g->ctx->stream));
  });
}

}  // extern "C"
