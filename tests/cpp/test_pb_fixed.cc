// CPU check of the binned pull's fixed-point terms (plato_amd/csrc/pb_fixed.hpp)
// against an independent 128-bit integer reference (unsigned __int128):
//   1. every non-negative FP32 value below 2^32 (all mantissas of exponents 0..158):
//      fixed_f32_acc == bits 32..127 of floor(x * 2^96);
//   2. 2^22 FP64 products of two random FP32 values (weighted terms): fixed_f64 ==
//      floor(x * 2^96);
//   3. sums: 2^20 random terms accumulated as the kernels do (two words, any order)
//      equal the __int128 sum of the truncated terms, and from_fixed of the words is
//      within one rounding of that sum.
#include <cstdio>
#include <cstdlib>
#include <random>

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
  return (bad || bad64 || badsum) ? 1 : 0;
}
