// Hub tables: the vertices of an adjacency with more than `deg` entries and the chunks of
// their entries, for passes that spread a hub's entries over every wave (see hub_table).
#include "pe_internal.hpp"

namespace pe {

// a bit per vertex with more than deg entries (one word per wave)
__global__ __launch_bounds__(256) void k_hub_bits(const uint64_t* __restrict__ ptr, uint32_t V, uint64_t nwords,
                                                  uint64_t deg, uint64_t* __restrict__ bits) {
  const int lane = threadIdx.x & 63;
  for (uint64_t word = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; word < nwords;
       word += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t v = word * 64 + lane;
    const bool hub = v < V && ptr[v + 1] - ptr[v] > deg;
    const uint64_t m = __ballot(hub);
    if (lane == 0) bits[word] = m;
  }
}

// One block: start[h] = first chunk of hub h, start[n] = the number of chunks.
__global__ __launch_bounds__(256) void k_hub_chunks(const uint64_t* __restrict__ ptr, const uint32_t* __restrict__ hubs,
                                                    const unsigned long long* __restrict__ nhub, uint64_t chunk,
                                                    uint64_t* __restrict__ start) {
  __shared__ uint64_t s_wave[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t n = *nhub;
  uint64_t carry = 0;
  for (uint64_t h0 = 0; h0 < n; h0 += 256) {
    const uint64_t h = h0 + threadIdx.x;
    uint64_t c = 0;
    if (h < n) {
      const uint32_t v = hubs[h];
      c = (ptr[v + 1] - ptr[v] + chunk - 1) / chunk;
    }
    uint64_t incl = c;
    for (int off = 1; off < 64; off <<= 1) {
      const uint64_t t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t before = 0, sum = 0;
    for (int k = 0; k < 4; ++k) {
      before += k < wave ? s_wave[k] : 0ull;
      sum += s_wave[k];
    }
    __syncthreads();
    if (h < n) start[h] = carry + before + incl - c;
    carry += sum;
  }
  if (threadIdx.x == 0) start[n] = carry;
}

void hub_table_build(pe_ctx_s* ctx, const uint64_t* ptr, uint32_t V, uint64_t nnz, uint64_t deg, uint64_t chunk,
                     hub_table& t) {
  hipStream_t s = ctx->stream;
  const uint64_t nwords = ((uint64_t)V >> 6) + 1;
  t.deg = deg;
  t.chunk = chunk;
  t.max_hubs = nnz / (deg + 1) + 1;  // every hub holds more than deg of the nnz entries
  t.hubs.alloc(ctx, t.max_hubs);
  t.start.alloc(ctx, t.max_hubs + 1);
  t.count.alloc(ctx, 1);
  t.bits.alloc(ctx, nwords);
  PE_HIP(hipMemsetAsync(t.count.p, 0, sizeof(unsigned long long), s));
  k_hub_bits<<<grid_for(nwords * 64, 256, 8192), 256, 0, s>>>(ptr, V, nwords, deg, t.bits.p);
  frontier_compact(ctx, t.bits.p, nwords, V, t.hubs.p, t.count.p);
  k_hub_chunks<<<1, 256, 0, s>>>(ptr, t.hubs.p, t.count.p, chunk, t.start.p);
  PE_HIP(hipGetLastError());
}

}  // namespace pe
