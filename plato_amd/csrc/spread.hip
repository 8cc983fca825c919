// spread_message (plato/graph/message_passing.hpp:152-249) on the device.
//
// The reference walks the active bitmap, lets a user functor send messages to
// explicit ranks, and sums the return values of a sink functor over the delivered
// messages.  On one device every message goes to rank 0, so the operator is a pass
// over the active bitmap: each active vertex v sends one message, target t = key[v]
// (or v itself), and the sink is one of a closed set:
//   PE_SPREAD_COUNT  counts[t] += 1                    (uint32 atomics)
//   PE_SPREAD_ADD    acc[t] += value[v]                (FP32 / FP64 atomics)
// Each delivered message returns 1 to the reduction, so the result is the number of
// messages, as in the reference's tests (message_passing.cc:90-169).  A target beyond
// the sink's vector is an error (PE_ERR_RANGE), reported after the pass.
//
// Layout: one lane per bitmap word, so a wave covers 4096 consecutive ids; the bits of
// a word are visited in ascending order.  Message counts are summed per block.
#include "pe_internal.hpp"

namespace pe {

template <int OP, typename T>
__global__ __launch_bounds__(256) void k_spread(const uint64_t* __restrict__ words, uint64_t nwords, uint64_t nbits,
                                                const uint32_t* __restrict__ key, const T* __restrict__ value,
                                                void* __restrict__ sink, uint64_t nsink,
                                                unsigned long long* __restrict__ counters) {
  unsigned long long sent = 0, dropped = 0;
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t bits = words[w];
    while (bits) {
      const uint64_t v = (w << 6) + (uint64_t)__builtin_ctzll(bits);
      bits &= bits - 1;
      if (v >= nbits) break;
      const uint64_t t = key ? key[v] : v;
      if (t >= nsink) {
        ++dropped;
        continue;
      }
      if (OP == PE_SPREAD_COUNT) atomicAdd(static_cast<uint32_t*>(sink) + t, 1u);
      else atomicAdd(static_cast<T*>(sink) + t, value ? value[v] : (T)1);
      ++sent;
    }
  }
  block_add2(sent, dropped, &counters[0], &counters[1]);
}

__global__ void k_spread_zero(unsigned long long* c) {
  c[0] = 0ull;
  c[1] = 0ull;
}

}  // namespace pe

using namespace pe;

extern "C" int pe_spread_message(pe_bitmap_t* actives, int op, pe_vec_t* key, pe_vec_t* value, pe_vec_t* sink,
                                 uint64_t* result) {
  if (!actives || !sink) return PE_ERR_INVALID;
  return guarded(actives->ctx, [&] {
    PE_REQUIRE(op == PE_SPREAD_COUNT || op == PE_SPREAD_ADD, PE_ERR_INVALID, "op must be PE_SPREAD_COUNT or PE_SPREAD_ADD");
    PE_REQUIRE(!key || key->dtype == PE_U32, PE_ERR_INVALID, "key must be a PE_U32 vector");
    PE_REQUIRE(!key || key->n >= actives->bits, PE_ERR_INVALID, "key shorter than the bitmap");
    PE_REQUIRE(sink->ctx == actives->ctx && (!key || key->ctx == actives->ctx) && (!value || value->ctx == actives->ctx),
               PE_ERR_INVALID, "objects of different contexts");
    if (op == PE_SPREAD_COUNT) {
      PE_REQUIRE(sink->dtype == PE_U32, PE_ERR_INVALID, "PE_SPREAD_COUNT needs a PE_U32 sink");
      PE_REQUIRE(!value, PE_ERR_INVALID, "PE_SPREAD_COUNT takes no value vector");
    } else {
      PE_REQUIRE(sink->dtype == PE_F32 || sink->dtype == PE_F64, PE_ERR_INVALID, "PE_SPREAD_ADD needs a PE_F32 or PE_F64 sink");
      PE_REQUIRE(!value || (value->dtype == sink->dtype && value->n >= actives->bits), PE_ERR_INVALID,
                 "value must have the sink's dtype and cover the bitmap");
    }
    pe_ctx_s* ctx = actives->ctx;
    hipStream_t s = ctx->stream;
    dbuf<unsigned long long> c(ctx, 2);
    k_spread_zero<<<1, 1, 0, s>>>(c.p);
    const unsigned grid = persistent_grid(ctx, grid_for(actives->words, 256), 8);
    const uint32_t* kp = key ? static_cast<const uint32_t*>(key->p) : nullptr;
    if (op == PE_SPREAD_COUNT)
      k_spread<PE_SPREAD_COUNT, uint32_t><<<grid, 256, 0, s>>>(actives->p, actives->words, actives->bits, kp, nullptr,
                                                               sink->p, sink->n, c.p);
    else if (sink->dtype == PE_F32)
      k_spread<PE_SPREAD_ADD, float><<<grid, 256, 0, s>>>(actives->p, actives->words, actives->bits, kp,
                                                          value ? static_cast<const float*>(value->p) : nullptr,
                                                          sink->p, sink->n, c.p);
    else
      k_spread<PE_SPREAD_ADD, double><<<grid, 256, 0, s>>>(actives->p, actives->words, actives->bits, kp,
                                                           value ? static_cast<const double*>(value->p) : nullptr,
                                                           sink->p, sink->n, c.p);
    PE_HIP(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    PE_HIP(hipMemcpyAsync(h, c.p, sizeof(h), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    PE_REQUIRE(h[1] == 0, PE_ERR_RANGE, "a message's target lies beyond the sink vector");
    if (result) *result = h[0];
  });
}
