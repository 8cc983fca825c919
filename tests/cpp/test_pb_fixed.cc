// CPU check of the binned pull's fixed-point terms (plato_amd/csrc/pb_fixed.hpp)
// against an independent 128-bit integer reference (unsigned __int128):
//   1. every non-negative FP32 value below 2^32 (all mantissas of exponents 0..158):
//      fixed_f32_acc == bits 32..127 of floor(x * 2^96);
//   2. 2^22 FP64 products of two random FP32 values (weighted terms): fixed_f64 ==
//      floor(x * 2^96);
//   3. sums: 2^20 random terms accumulated in the two-word form (any order) equal the
//      __int128 sum of the truncated terms, and from_fixed of the words is within one
//      rounding of that sum;
//   4. the single-word form the kernels use: fixed64_f32 and the decoded 32-bit code
//      fixed_code_term(fixed_code_f32) == floor(x * 2^F) for every FP32 value x with
//      x * 2^F < 2^64, at F = 33 (PageRank on 2^27 vertices) and F = 60 (PPR);
//      fixed64_f64 on 2^22 FP64 products at both;
//   5. pb_frac_bits(M) puts M * 2^F in [2^60, 2^61), and 2^20 terms whose values add
//      up to M sum in one 64-bit word exactly as in __int128, below 2^61.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "pb_fixed.hpp"

using u128i = unsigned __int128;

static u128i ref_fixed(uint64_t m, int k) {  // floor(m * 2^k), k may be negative
  if (k >= 0) return k >= 128 ? 0 : (u128i)m << k;
  return -k >= 64 ? 0 : (u128i)(m >> -k);
}

int main() {
  long bad = 0;
  // 1. exhaustive FP32
  for (uint32_t e = 0; e <= 158; ++e) {
    for (uint32_t f = 0; f < (1u << 23); ++f) {
      const uint32_t bits = (e << 23) | f;
      float x;
      __builtin_memcpy(&x, &bits, 4);
      const pe::u128 r = pe::fixed_f32_acc(x);
      u128i R = 0;
      if (e) R = ref_fixed((uint64_t)(f | (1u << 23)), (int)e - 150 + pe::PB_FRAC);
      const uint64_t d1 = (uint64_t)(R >> 32) & 0xFFFFFFFFull, hi = (uint64_t)(R >> 64);
      if (r.lo != (d1 << 32) || r.hi != hi) {
        if (bad < 5) std::printf("fp32 mismatch bits=%08x\n", bits);
        ++bad;
      }
    }
  }
  std::printf("fp32 exhaustive: %ld mismatches\n", bad);
  // 2. FP64 products
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  std::uniform_int_distribution<int> ex(-40, 20);
  long bad64 = 0;
  for (int i = 0; i < (1 << 22); ++i) {
    const float w = U(rng), xv = std::ldexp(U(rng), ex(rng));
    const double p = (double)w * (double)xv;
    const pe::u128 r = pe::fixed_f64(p);
    u128i R = 0;
    uint64_t b;
    __builtin_memcpy(&b, &p, 8);
    const int e = (int)((b >> 52) & 0x7FF);
    if (e) R = ref_fixed((b & ((1ull << 52) - 1)) | (1ull << 52), e - 1075 + pe::PB_FRAC);
    if (r.lo != (uint64_t)R || r.hi != (uint64_t)(R >> 64)) ++bad64;
  }
  std::printf("fp64 products: %ld mismatches\n", bad64);
  // 3. sums in two words against the __int128 sum of truncated terms
  long badsum = 0;
  for (int trial = 0; trial < 4; ++trial) {
    uint64_t lo = 0, hi = 0;
    u128i S = 0;
    const int n = 1 << 20;
    for (int i = 0; i < n; ++i) {
      const float xv = std::ldexp(U(rng), ex(rng));
      const pe::u128 r = pe::fixed_f32_acc(xv);
      lo += r.lo >> 32;
      hi += r.hi;
      S += ((u128i)r.hi << 64) | (u128i)r.lo;
    }
    const u128i T = ((u128i)hi << 64) + ((u128i)lo << 32);
    if (T != S) ++badsum;
    const double v = pe::from_fixed(lo, hi);
    const long double exact = (long double)hi * 0x1p-32L + (long double)lo * 0x1p-64L;
    if (std::fabs((long double)v - exact) > exact * 0x1p-51L) ++badsum;
  }
  std::printf("sums: %ld mismatches\n", badsum);
  // 4. single-word terms
  long bad1 = 0;
  for (int F : {33, 60}) {
    for (uint32_t e = 0; (int)e - 127 + 1 + F <= 64 && e < 255; ++e) {
      for (uint32_t f = 0; f < (1u << 23); ++f) {
        const uint32_t bits = (e << 23) | f;
        float x;
        __builtin_memcpy(&x, &bits, 4);
        const uint64_t m = e ? (uint64_t)(f | (1u << 23)) : (uint64_t)f;
        const u128i R = ref_fixed(m, (e ? (int)e : 1) - 150 + F);
        if (pe::fixed64_f32(x, F) != (uint64_t)R || (R >> 64) != 0) {
          if (bad1 < 5) std::printf("fixed64 mismatch F=%d bits=%08x\n", F, bits);
          ++bad1;
        }
        if (pe::fixed_code_term(pe::fixed_code_f32(x, F)) != (uint64_t)R) {
          if (bad1 < 5) std::printf("fixed code mismatch F=%d bits=%08x\n", F, bits);
          ++bad1;
        }
      }
    }
    for (int i = 0; i < (1 << 22); ++i) {
      const float w = U(rng), xv = std::ldexp(U(rng), ex(rng) - (F > 40 ? 22 : 0));
      const double p = (double)w * (double)xv;
      uint64_t b;
      __builtin_memcpy(&b, &p, 8);
      const int e = (int)((b >> 52) & 0x7FF);
      const uint64_t m = e ? ((b & ((1ull << 52) - 1)) | (1ull << 52)) : (b & ((1ull << 52) - 1));
      const u128i R = ref_fixed(m, (e ? e : 1) - 1075 + F);
      if (pe::fixed64_f64(p, F) != (uint64_t)R) ++bad1;
    }
  }
  std::printf("single-word terms: %ld mismatches\n", bad1);
  // 5. scale choice and single-word sums
  long badf = 0;
  for (int i = -60; i <= 40; ++i) {
    for (double frac : {1.0, 1.37, 1.999}) {
      const double M = std::ldexp(frac, i);
      const int F = pe::pb_frac_bits(M);
      const long double s = std::ldexp((long double)M, F);
      if (!(s >= 0x1p60L && s < 0x1p61L)) ++badf;
    }
  }
  for (double M : {1.0, 134217728.0}) {
    const int F = pe::pb_frac_bits(M);
    const int n = 1 << 20;
    uint64_t acc = 0;
    u128i S = 0;
    double total = 0.0;
    std::vector<float> xs(n);
    for (int i = 0; i < n; ++i) {
      xs[i] = (float)(U(rng) * 2.0 * M / n);
      total += xs[i];
    }
    const float scale = (float)(M / total);  // values add up to about M
    for (int i = 0; i < n; ++i) {
      const float xv = xs[i] * scale;
      const uint64_t t = pe::fixed64_f32(xv, F);
      acc += t;
      S += ref_fixed(xv == 0.0f ? 0 : ((pe::pb_bits32(xv) & 0x7FFFFFu) | 0x800000u),
                     (int)((pe::pb_bits32(xv) >> 23) & 0xFF) - 150 + F);
    }
    if ((u128i)acc != S || S >= ((u128i)1 << 61)) ++badf;
    if (std::fabs(pe::from_fixed64(acc, F) - (double)M) > 1e-6 * M) ++badf;
  }
  std::printf("scale and single-word sums: %ld mismatches\n", badf);
  return (bad || bad64 || badsum || bad1 || badf) ? 1 : 0;
}
