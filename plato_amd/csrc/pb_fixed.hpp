// Fixed-point terms of the binned pull (pull_binned.hip), host and device.
//
// The kernels use the single-word form: a term x (non-negative, finite) is the integer
// floor(x * 2^F) in one 64-bit word (FP32 terms travel as the 32-bit code of
// fixed_code_f32 and are widened by fixed_code_term; fixed64_f32 is the same term
// computed directly, kept for the host check; weighted terms use fixed64_f64), with F
// fixed for a whole iteration by pb_frac_bits(M), M a bound on the total of all terms
// (the rank mass that flows along the edges): every per-destination sum of terms then stays below
// 2^61, so 64-bit integer sums cannot overflow and are exact and order-independent.
// Each term is truncated at 2^-F (at most one unit of 2^-F per term).
//
// The two-word form below keeps 96 fractional bits whatever the magnitude; it is
// kept as the reference the single-word form is checked against on the host.
//
// A term x (non-negative, finite, below 2^32) is the integer R = floor(x * 2^PB_FRAC)
// with PB_FRAC = 96, kept as two accumulator parts: `hi` = bits 64..127 of R and
// `d1` = bits 32..63 of R (carried in u128::lo as d1 << 32); bits below 2^32 of R
// (2^-64 in value) are truncated.  Per destination, the `hi` parts are summed in one
// 64-bit word (their total stays below 2^64 while the sum of values stays below 2^32)
// and the `d1` parts in another (each below 2^32, so no overflow before 2^32 terms):
// no carry ever crosses between the words, and the sum is exact for the truncated
// terms in any order.  The value of a destination is hi * 2^-32 + lo * 2^-64.
//
// Compiled by hipcc for the kernels and by the host compiler for the CPU check that
// compares every FP32 term against a 128-bit integer reference.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PB_HD __host__ __device__ __forceinline__
#else
#define PB_HD inline
#endif

namespace pe {

static constexpr int PB_FRAC = 96;  // fractional bits of the fixed-point terms

struct u128 {
  uint64_t lo, hi;
};

PB_HD uint32_t pb_bits32(float v) {
  uint32_t b;
  __builtin_memcpy(&b, &v, sizeof(b));
  return b;
}

PB_HD uint64_t pb_bits64(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, sizeof(b));
  return b;
}

// m * 2^k as a 128-bit integer, bits below 2^0 dropped, branch-free; exact for
// k + bits(m) <= 128.  64-bit shifts are written with their amount mod 64, each
// candidate selected by the range of k where it is the right one.
PB_HD u128 shl128(uint64_t m, int k) {
  const uint64_t l_left = m << (k & 63), l_right = m >> ((-k) & 63);
  const uint64_t h_left = m << ((k - 64) & 63), h_right = m >> ((64 - k) & 63);
  u128 r;
  r.lo = (k >= 0) ? ((k < 64) ? l_left : 0ull) : ((k > -64) ? l_right : 0ull);
  r.hi = (k >= 64) ? ((k < 128) ? h_left : 0ull) : ((k > 0) ? h_right : 0ull);
  return r;
}

// FP32 term -> (d1 << 32, hi).  x = m * 2^(e - 150); R = m * 2^k with k = e - 54.
// Zero and subnormals give 0 (they lie below the 2^-64 resolution).
PB_HD u128 fixed_f32_acc(float v) {
  const uint32_t b = pb_bits32(v);
  const uint32_t e = (b >> 23) & 0xFFu;
  const uint32_t m = e ? ((b & 0x7FFFFFu) | 0x800000u) : 0u;
  const int k = (int)e - 150 + PB_FRAC;
  const int s1 = k - 32;  // d1 = m << s1 (s1 in [0, 32)) or m >> -s1 (s1 in (-32, 0))
  const uint32_t d1l = m << (s1 & 31), d1r = m >> ((-s1) & 31);
  const uint32_t d1 = ((uint32_t)s1 < 32u) ? d1l : (((uint32_t)(-s1 - 1) < 31u) ? d1r : 0u);
  const uint64_t hl = (uint64_t)m << ((k - 64) & 63);  // k in [64, 128)
  const uint32_t hr = m >> ((64 - k) & 31);            // k in (32, 64): bits that reach 2^64
  const uint64_t hi = ((uint32_t)(k - 64) < 64u) ? hl : (((uint32_t)(k - 33) < 31u) ? (uint64_t)hr : 0ull);
  return u128{(uint64_t)d1 << 32, hi};
}

// FP64 term (weighted graphs: the exact product w * x) -> the full 128-bit R; the
// accumulator keeps its bits 32..127 like fixed_f32_acc.
PB_HD u128 fixed_f64(double v) {
  const uint64_t b = pb_bits64(v);
  const int e = (int)((b >> 52) & 0x7FFull);
  if (e == 0 || (b >> 63)) return u128{0ull, 0ull};
  const uint64_t m = (b & ((1ull << 52) - 1ull)) | (1ull << 52);
  return shl128(m, e - 1075 + PB_FRAC);
}

// floor(m * 2^k) for k in (-64, 64) (0 for k <= -64), branch-free; the caller keeps
// the result below 2^64.
PB_HD uint64_t pb_shl64(uint64_t m, int k) {
  const uint64_t l = m << (k & 63), r = m >> ((-k) & 63);
  return k >= 0 ? (k < 64 ? l : 0ull) : (k > -64 ? r : 0ull);
}

// Fractional bits F for terms whose total is at most `mass`: mass * 2^F <= 2^61.
PB_HD int pb_frac_bits(double mass) {
  if (!(mass >= 0x1p-60)) mass = 0x1p-60;
  int e = 0;
  (void)frexp(mass, &e);  // mass < 2^e
  return 61 - e;
}

// FP32 term -> floor(x * 2^F) for x non-negative and finite with x * 2^F < 2^64; zero
// and subnormals of FP32 give their exact floor too.  x = m * 2^(k - F) with the 24-bit
// significand m, so the term is m shifted left by k >= 0 or right by -k (a right shift
// by 31 already clears any m); this takes about half the instructions of pb_shl64.
PB_HD uint64_t fixed64_f32(float v, int F) {
  const uint32_t b = pb_bits32(v);  // no sign bit
  const uint32_t e = b >> 23;
  const uint32_t m = (b & 0x7FFFFFu) | ((e ? 1u : 0u) << 23);
  const int k = (e ? (int)e : 1) - 150 + F;
  const int s = k > 0 ? k : 0;
  const int t = k >= 0 ? 0 : (-k < 31 ? -k : 31);
  return (uint64_t)(m >> t) << s;
}

// The same term in 32 bits: the shifted significand (below 2^24) in bits 0..23 and the
// left shift still to apply (0..40, since the term is below 2^64) above it.  The binned
// scatter codes each source once while staging it, so the gather decodes a term with a
// mask and two shifts.
PB_HD uint32_t fixed_code_f32(float v, int F) {
  const uint32_t b = pb_bits32(v);  // no sign bit
  const uint32_t e = b >> 23;
  const uint32_t m = (b & 0x7FFFFFu) | ((e ? 1u : 0u) << 23);
  const int k = (e ? (int)e : 1) - 150 + F;
  const int s = k > 0 ? k : 0;
  const int t = k >= 0 ? 0 : (-k < 31 ? -k : 31);
  return ((uint32_t)s << 24) | (m >> t);
}

PB_HD uint64_t fixed_code_term(uint32_t c) { return (uint64_t)(c & 0xFFFFFFu) << (c >> 24); }

// FP64 term (weighted graphs: the exact product w * x) -> floor(x * 2^F).
PB_HD uint64_t fixed64_f64(double v, int F) {
  const uint64_t b = pb_bits64(v);
  if (b >> 63) return 0ull;
  const int e = (int)((b >> 52) & 0x7FFull);
  const uint64_t m = e ? ((b & ((1ull << 52) - 1ull)) | (1ull << 52)) : (b & ((1ull << 52) - 1ull));
  return pb_shl64(m, (e ? e : 1) - 1075 + F);
}

// Value of a single-word accumulator.
PB_HD double from_fixed64(uint64_t acc, int F) { return ldexp((double)acc, -F); }

// Value of a destination's accumulator words (lo = sum of d1 parts, hi = sum of hi parts).
PB_HD double from_fixed(uint64_t lo, uint64_t hi) {
  return ldexp((double)hi, 64 - PB_FRAC) + ldexp((double)lo, 32 - PB_FRAC);
}

}  // namespace pe
