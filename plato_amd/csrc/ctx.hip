// Context, allocation ledger, timers, engine state vectors and bitmaps.
#include "pe_internal.hpp"
#include "pull.hpp"

namespace pe {

void* dmalloc(pe_ctx_s* ctx, size_t bytes) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    char buf[160];
    std::snprintf(buf, sizeof(buf), "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    fail(PE_ERR_OOM, buf);
  }
  ctx->allocated += bytes;
  if (ctx->allocated > ctx->peak_allocated) ctx->peak_allocated = ctx->allocated;
  return p;
}

void dfree(pe_ctx_s* ctx, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  ctx->allocated -= bytes;
}

ktime_scope::ktime_scope(pe_ctx_s* c, int w) : ctx(c), which(w) {
  if (!ctx || !ctx->timing) return;
  hipEvent_t e0;
  PE_HIP(hipEventCreate(&e0));
  PE_HIP(hipEventCreate(&e1));
  PE_HIP(hipEventRecord(e0, ctx->stream));
  ctx->timers[which].begin.push_back(e0);
}

ktime_scope::~ktime_scope() {
  if (!e1) return;
  if (hipEventRecord(e1, ctx->stream) == hipSuccess) ctx->timers[which].end.push_back(e1);
  else (void)hipEventDestroy(e1);
}

// ---- small kernels --------------------------------------------------------------

template <typename T>
__global__ void k_fill(T* __restrict__ p, uint64_t n, T v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
__global__ void k_to_f64(const T* __restrict__ p, double* __restrict__ out, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (double)p[i];
}

// Deterministic two-level FP64 sum: fixed grid, fixed per-block order.
template <typename T>
__global__ void k_sum_partial(const T* __restrict__ p, uint64_t n, double* __restrict__ part) {
  __shared__ double s[256];
  double acc = 0.0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    acc += (double)p[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_sum_final(const double* __restrict__ part, unsigned n, double* __restrict__ out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// plain streaming copy (also the calibration kernel for traffic counters)
__global__ __launch_bounds__(256) void k_copy_words(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   uint64_t n16) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// read-only stream (the read rate): every word is loaded and folded into a value that
// is stored only if it equals a constant, so the loads cannot be dropped
__global__ __launch_bounds__(256) void k_read_words(const uint4* __restrict__ src, uint64_t n16,
                                                   uint32_t* __restrict__ sink) {
  uint32_t acc = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 v = src[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;
}

__global__ void k_copy_tail(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_bitmap_fill(uint64_t* __restrict__ w, uint64_t words, uint64_t bits) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x) {
    // bits [0, bits) set; the tail of the last word masked (bitmap.hpp:168-179)
    uint64_t lo = i * 64;
    uint64_t v;
    if (lo + 64 <= bits) v = ~0ull;
    else if (lo >= bits) v = 0ull;
    else v = (1ull << (bits - lo)) - 1ull;
    w[i] = v;
  }
}

__global__ void k_bitmap_count(const uint64_t* __restrict__ w, uint64_t words,
                               unsigned long long* __restrict__ out) {
  unsigned long long c = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += __popcll(w[i]);
  // wave reduce then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

__global__ void k_bitmap_msb(const uint64_t* __restrict__ w, uint64_t words,
                             unsigned long long* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t v = w[i];
    if (v) {
      unsigned long long bit = i * 64 + (63 - __clzll(v));
      atomicMax(out, bit + 1);  // 0 means empty
    }
  }
}

__global__ void k_bitmap_set(uint64_t* w, uint64_t i, int set) {
  if (set) atomicOr((unsigned long long*)&w[i >> 6], 1ull << (i & 63));
  else atomicAnd((unsigned long long*)&w[i >> 6], ~(1ull << (i & 63)));
}

// Compaction of set bits into ascending ids: one wave per 64 words; per-wave
// offsets come from a prefix over word popcounts (two passes).
__global__ void k_bitmap_wave_counts(const uint64_t* __restrict__ w, uint64_t words,
                                     uint32_t* __restrict__ cnt, uint64_t nwaves) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint32_t c = i < words ? __popcll(w[i]) : 0u;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0 && (i >> 6) < nwaves) cnt[i >> 6] = c;
}

__global__ void k_bitmap_emit(const uint64_t* __restrict__ w, uint64_t words,
                              const uint64_t* __restrict__ base, uint32_t* __restrict__ ids,
                              uint64_t cap, uint64_t nwaves) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  if ((i >> 6) >= nwaves) return;  // whole wave out of range
  uint64_t v = i < words ? w[i] : 0ull;
  uint32_t c = __popcll(v);
  // exclusive prefix of popcounts inside the wave
  uint32_t incl = c;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  uint64_t pos = base[i >> 6] + (incl - c);
  while (v) {
    int b = __ffsll((unsigned long long)v) - 1;
    if (pos < cap) ids[pos] = (uint32_t)(i * 64 + b);
    ++pos;
    v &= v - 1;
  }
}

// One thread: the step's counters to the host words, the next step's counters zeroed,
// then the sequence number, released at system scope after the words.
__global__ void k_mail_post(const unsigned long long* __restrict__ src, uint32_t n,
                            unsigned long long* __restrict__ clear, uint32_t nclear,
                            unsigned long long* __restrict__ mail, unsigned long long seq) {
  for (uint32_t i = 0; i < n; ++i) mail[1 + i] = src[i];
  for (uint32_t i = 0; i < nclear; ++i) clear[i] = 0ull;
  __threadfence_system();
  __hip_atomic_store(&mail[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

const unsigned long long* mail_post(pe_ctx_s* ctx, const unsigned long long* src, uint32_t n,
                                    unsigned long long* clear, uint32_t nclear) {
  PE_REQUIRE(n <= 64, PE_ERR_INTERNAL, "mail_post: more than 64 words");
  const unsigned long long seq = ++ctx->mail_seq;
  k_mail_post<<<1, 1, 0, ctx->stream>>>(src, n, clear, nclear, ctx->host_mail, seq);
  PE_HIP(hipGetLastError());
  // poll the sequence number; every 1024 polls ask the stream, so that a failed kernel
  // ends the wait with its error instead of spinning
  for (uint64_t k = 1;; ++k) {
    if (__atomic_load_n(&ctx->host_mail[0], __ATOMIC_ACQUIRE) == seq) break;
    if ((k & 1023) == 0) {
      const hipError_t e = hipStreamQuery(ctx->stream);
      if (e == hipSuccess) {
        PE_REQUIRE(__atomic_load_n(&ctx->host_mail[0], __ATOMIC_ACQUIRE) == seq, PE_ERR_INTERNAL,
                   "mail_post: stream finished without the sequence number");
        break;
      }
      if (e != hipErrorNotReady) PE_HIP(e);
    }
  }
  return ctx->host_mail + 1;
}

}  // namespace pe

using namespace pe;

// ================================================================================
// C-ABI
// ================================================================================

extern "C" {

int pe_ctx_create(int device, pe_ctx_t** out) {
  if (!out) return PE_ERR_INVALID;
  *out = nullptr;
  pe_ctx_s* ctx = new (std::nothrow) pe_ctx_s();
  if (!ctx) return PE_ERR_OOM;
  int rc = guarded(ctx, [&] {
    int n = 0;
    PE_HIP(hipGetDeviceCount(&n));
    PE_REQUIRE(device >= 0 && device < n, PE_ERR_INVALID, "device index out of range");
    ctx->device = device;
    PE_HIP(hipSetDevice(device));
    PE_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    PE_HIP(hipEventCreate(&ctx->t0));
    PE_HIP(hipEventCreate(&ctx->t1));
    int cus = 0;
    PE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    ctx->num_cus = cus;
    PE_HIP(hipHostMalloc(&ctx->host_pinned, 4096, hipHostMallocDefault));
    PE_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_mail), 4096, hipHostMallocCoherent));
    ctx->host_mail[0] = 0ull;
    ctx->noalloc_scratch = pull_scratch_setup(device);  // with the device current
  });
  if (rc != PE_OK) {
    std::fprintf(stderr, "pe_ctx_create: %s\n", ctx->last_error.c_str());
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return PE_OK;
}

int pe_ctx_destroy(pe_ctx_t* ctx) {
  if (!ctx) return PE_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->timers) {
    for (auto e : t.begin) (void)hipEventDestroy(e);
    for (auto e : t.end) (void)hipEventDestroy(e);
  }
  if (ctx->t0) (void)hipEventDestroy(ctx->t0);
  if (ctx->t1) (void)hipEventDestroy(ctx->t1);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->host_mail) (void)hipHostFree(ctx->host_mail);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PE_OK;
}

const char* pe_ctx_last_error(const pe_ctx_t* ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

int pe_ctx_sync(pe_ctx_t* ctx) {
  if (!ctx) return PE_ERR_INVALID;
  return guarded(ctx, [&] { PE_HIP(hipStreamSynchronize(ctx->stream)); });
}

int pe_ctx_mem_info(pe_ctx_t* ctx, uint64_t* free_bytes, uint64_t* total_bytes,
                    uint64_t* ctx_bytes) {
  if (!ctx) return PE_ERR_INVALID;
  return guarded(ctx, [&] {
    size_t f = 0, t = 0;
    PE_HIP(hipSetDevice(ctx->device));
    PE_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    if (ctx_bytes) *ctx_bytes = ctx->allocated;
  });
}

int pe_ctx_timer_start(pe_ctx_t* ctx) {
  if (!ctx) return PE_ERR_INVALID;
  return guarded(ctx, [&] { PE_HIP(hipEventRecord(ctx->t0, ctx->stream)); });
}

int pe_ctx_timer_stop(pe_ctx_t* ctx, double* ms) {
  if (!ctx || !ms) return PE_ERR_INVALID;
  return guarded(ctx, [&] {
    PE_HIP(hipEventRecord(ctx->t1, ctx->stream));
    PE_HIP(hipEventSynchronize(ctx->t1));
    float f = 0.f;
    PE_HIP(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
    *ms = f;
  });
}

int pe_ctx_set_kernel_timing(pe_ctx_t* ctx, int enable) {
  if (!ctx) return PE_ERR_INVALID;
  ctx->timing = enable ? 1 : 0;
  return PE_OK;
}

int pe_ctx_stream_rates(pe_ctx_t* ctx, uint64_t bytes, int reps, double* read_gbs, double* copy_gbs) {
  if (!ctx || bytes == 0 || bytes % 16 || reps < 1) return PE_ERR_INVALID;
  return guarded(ctx, [&] {
    const uint64_t n16 = bytes / 16;
    dbuf<uint4> a(ctx, n16), b(ctx, n16);
    dbuf<uint32_t> sink(ctx, 1);
    hipStream_t s = ctx->stream;
    PE_HIP(hipMemsetAsync(a.p, 1, bytes, s));
    PE_HIP(hipMemsetAsync(b.p, 0, bytes, s));
    const unsigned grid = (unsigned)std::max(1, ctx->num_cus) * 8u;
    float ms = 0.f;
    k_read_words<<<grid, 256, 0, s>>>(a.p, n16, sink.p);
    PE_HIP(hipEventRecord(ctx->t0, s));
    for (int r = 0; r < reps; ++r) k_read_words<<<grid, 256, 0, s>>>(a.p, n16, sink.p);
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    if (read_gbs) *read_gbs = (double)bytes * reps / (ms * 1e-3) / 1e9;
    k_copy_words<<<grid, 256, 0, s>>>(a.p, b.p, n16);
    PE_HIP(hipEventRecord(ctx->t0, s));
    for (int r = 0; r < reps; ++r) k_copy_words<<<grid, 256, 0, s>>>(a.p, b.p, n16);
    PE_HIP(hipEventRecord(ctx->t1, s));
    PE_HIP(hipEventSynchronize(ctx->t1));
    PE_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    if (copy_gbs) *copy_gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    PE_HIP(hipGetLastError());
  });
}

int pe_ctx_kernel_time(pe_ctx_t* ctx, int which, double* total_ms, uint64_t* launches) {
  if (!ctx || which < 0 || which >= PE_KT_COUNT) return PE_ERR_INVALID;
  return guarded(ctx, [&] {
    auto& t = ctx->timers[which];
    PE_HIP(hipStreamSynchronize(ctx->stream));
    size_t n = t.begin.size() < t.end.size() ? t.begin.size() : t.end.size();
    for (size_t i = 0; i < n; ++i) {
      float f = 0.f;
      PE_HIP(hipEventElapsedTime(&f, t.begin[i], t.end[i]));
      t.total_ms += f;
      t.launches += 1;
    }
    for (auto e : t.begin) (void)hipEventDestroy(e);
    for (auto e : t.end) (void)hipEventDestroy(e);
    t.begin.clear();
    t.end.clear();
    if (total_ms) *total_ms = t.total_ms;
    if (launches) *launches = t.launches;
    t.total_ms = 0.0;
    t.launches = 0;
  });
}

int pe_ctx_num_cus(pe_ctx_t* ctx, int* cus) {
  if (!ctx || !cus) return PE_ERR_INVALID;
  *cus = ctx->num_cus;
  return PE_OK;
}

// ---- vectors ----------------------------------------------------------------------

static size_t dtype_size(int dt) {
  switch (dt) {
    case PE_F32: return 4;
    case PE_F64: return 8;
    case PE_U32: return 4;
    default: return 0;
  }
}

int pe_vec_create(pe_ctx_t* ctx, int dtype, uint64_t n, pe_vec_t** out) {
  if (!ctx || !out) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&] {
    size_t es = dtype_size(dtype);
    PE_REQUIRE(es != 0, PE_ERR_INVALID, "unknown dtype");
    auto* v = new pe_vec_s();
    v->ctx = ctx;
    v->dtype = dtype;
    v->n = n;
    v->bytes = es * (n ? n : 1);
    try {
      v->p = dmalloc(ctx, v->bytes);
    } catch (...) {
      delete v;
      throw;
    }
    PE_HIP(hipMemsetAsync(v->p, 0, v->bytes, ctx->stream));
    *out = v;
  });
}

int pe_vec_destroy(pe_vec_t* v) {
  if (!v) return PE_ERR_INVALID;
  (void)hipStreamSynchronize(v->ctx->stream);
  dfree(v->ctx, v->p, v->bytes);
  delete v;
  return PE_OK;
}

int pe_vec_fill(pe_vec_t* v, double value) {
  if (!v) return PE_ERR_INVALID;
  return guarded(v->ctx, [&] {
    unsigned g = grid_for(v->n, 256, 4096);
    if (v->dtype == PE_F32)
      k_fill<float><<<g, 256, 0, v->ctx->stream>>>((float*)v->p, v->n, (float)value);
    else if (v->dtype == PE_F64)
      k_fill<double><<<g, 256, 0, v->ctx->stream>>>((double*)v->p, v->n, value);
    else
      k_fill<uint32_t><<<g, 256, 0, v->ctx->stream>>>((uint32_t*)v->p, v->n, (uint32_t)value);
    PE_HIP(hipGetLastError());
  });
}

int pe_vec_copy(pe_vec_t* dst, const pe_vec_t* src) {
  if (!dst || !src) return PE_ERR_INVALID;
  return guarded(dst->ctx, [&] {
    PE_REQUIRE(dst->dtype == src->dtype && dst->n == src->n, PE_ERR_INVALID, "vectors differ in dtype or length");
    const uint64_t bytes = src->n * dtype_size(src->dtype);
    const uint64_t n16 = bytes / 16;
    hipStream_t s = dst->ctx->stream;
    if (n16)
      k_copy_words<<<persistent_grid(dst->ctx, grid_for(n16, 256), 16), 256, 0, s>>>(
          (const uint4*)src->p, (uint4*)dst->p, n16);
    if (bytes % 16)
      k_copy_tail<<<1, 64, 0, s>>>((const uint8_t*)src->p + n16 * 16, (uint8_t*)dst->p + n16 * 16, bytes % 16);
    PE_HIP(hipGetLastError());
  });
}

int pe_vec_upload(pe_vec_t* v, const void* host, uint64_t n) {
  if (!v || (!host && n)) return PE_ERR_INVALID;
  return guarded(v->ctx, [&] {
    PE_REQUIRE(n <= v->n, PE_ERR_RANGE, "upload longer than vector");
    PE_HIP(hipMemcpyAsync(v->p, host, n * dtype_size(v->dtype), hipMemcpyHostToDevice,
                          v->ctx->stream));
    PE_HIP(hipStreamSynchronize(v->ctx->stream));
  });
}

int pe_vec_download(pe_vec_t* v, void* host, uint64_t n) {
  if (!v || (!host && n)) return PE_ERR_INVALID;
  return guarded(v->ctx, [&] {
    PE_REQUIRE(n <= v->n, PE_ERR_RANGE, "download longer than vector");
    PE_HIP(hipMemcpyAsync(host, v->p, n * dtype_size(v->dtype), hipMemcpyDeviceToHost,
                          v->ctx->stream));
    PE_HIP(hipStreamSynchronize(v->ctx->stream));
  });
}

int pe_vec_download_f64(pe_vec_t* v, double* host, uint64_t n) {
  if (!v || (!host && n)) return PE_ERR_INVALID;
  return guarded(v->ctx, [&] {
    PE_REQUIRE(n <= v->n, PE_ERR_RANGE, "download longer than vector");
    if (v->dtype == PE_F64) {
      PE_HIP(hipMemcpyAsync(host, v->p, n * 8, hipMemcpyDeviceToHost, v->ctx->stream));
    } else {
      dbuf<double> tmp(v->ctx, n ? n : 1);
      unsigned g = grid_for(n, 256, 4096);
      if (v->dtype == PE_F32)
        k_to_f64<float><<<g, 256, 0, v->ctx->stream>>>((const float*)v->p, tmp.p, n);
      else
        k_to_f64<uint32_t><<<g, 256, 0, v->ctx->stream>>>((const uint32_t*)v->p, tmp.p, n);
      PE_HIP(hipGetLastError());
      PE_HIP(hipMemcpyAsync(host, tmp.p, n * 8, hipMemcpyDeviceToHost, v->ctx->stream));
      PE_HIP(hipStreamSynchronize(v->ctx->stream));
    }
    PE_HIP(hipStreamSynchronize(v->ctx->stream));
  });
}

int pe_vec_set(pe_vec_t* v, uint64_t i, double value) {
  if (!v) return PE_ERR_INVALID;
  return guarded(v->ctx, [&] {
    PE_REQUIRE(i < v->n, PE_ERR_RANGE, "index out of range");
    if (v->dtype == PE_F32) {
      float f = (float)value;
      PE_HIP(hipMemcpyAsync((float*)v->p + i, &f, 4, hipMemcpyHostToDevice, v->ctx->stream));
      PE_HIP(hipStreamSynchronize(v->ctx->stream));
    } else if (v->dtype == PE_F64) {
      PE_HIP(hipMemcpyAsync((double*)v->p + i, &value, 8, hipMemcpyHostToDevice, v->ctx->stream));
      PE_HIP(hipStreamSynchronize(v->ctx->stream));
    } else {
      uint32_t u = (uint32_t)value;
      PE_HIP(hipMemcpyAsync((uint32_t*)v->p + i, &u, 4, hipMemcpyHostToDevice, v->ctx->stream));
      PE_HIP(hipStreamSynchronize(v->ctx->stream));
    }
  });
}

int pe_vec_sum(pe_vec_t* v, double* out) {
  if (!v || !out) return PE_ERR_INVALID;
  return guarded(v->ctx, [&] {
    const unsigned G = 512;
    dbuf<double> part(v->ctx, G + 1);
    if (v->dtype == PE_F32)
      k_sum_partial<float><<<G, 256, 0, v->ctx->stream>>>((const float*)v->p, v->n, part.p);
    else if (v->dtype == PE_F64)
      k_sum_partial<double><<<G, 256, 0, v->ctx->stream>>>((const double*)v->p, v->n, part.p);
    else
      k_sum_partial<uint32_t><<<G, 256, 0, v->ctx->stream>>>((const uint32_t*)v->p, v->n, part.p);
    k_sum_final<<<1, 256, 0, v->ctx->stream>>>(part.p, G, part.p + G);
    PE_HIP(hipGetLastError());
    PE_HIP(hipMemcpyAsync(out, part.p + G, 8, hipMemcpyDeviceToHost, v->ctx->stream));
    PE_HIP(hipStreamSynchronize(v->ctx->stream));
  });
}

// ---- bitmaps ------------------------------------------------------------------------

int pe_bitmap_create(pe_ctx_t* ctx, uint64_t bits, pe_bitmap_t** out) {
  if (!ctx || !out) return PE_ERR_INVALID;
  *out = nullptr;
  return guarded(ctx, [&] {
    auto* b = new pe_bitmap_s();
    b->ctx = ctx;
    b->bits = bits;
    b->words = (bits >> 6) + 1;
    try {
      b->p = static_cast<uint64_t*>(dmalloc(ctx, b->words * 8));
    } catch (...) {
      delete b;
      throw;
    }
    PE_HIP(hipMemsetAsync(b->p, 0, b->words * 8, ctx->stream));
    *out = b;
  });
}

int pe_bitmap_destroy(pe_bitmap_t* b) {
  if (!b) return PE_ERR_INVALID;
  (void)hipStreamSynchronize(b->ctx->stream);
  dfree(b->ctx, b->p, b->words * 8);
  delete b;
  return PE_OK;
}

int pe_bitmap_clear(pe_bitmap_t* b) {
  if (!b) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    k_fill<uint64_t><<<grid_for(b->words, 256, 2048), 256, 0, b->ctx->stream>>>(b->p, b->words, 0ull);
    PE_HIP(hipGetLastError());
  });
}

int pe_bitmap_fill(pe_bitmap_t* b) {
  if (!b) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    k_bitmap_fill<<<grid_for(b->words, 256, 2048), 256, 0, b->ctx->stream>>>(b->p, b->words, b->bits);
    PE_HIP(hipGetLastError());
  });
}

int pe_bitmap_set_bit(pe_bitmap_t* b, uint64_t i) {
  if (!b) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    PE_REQUIRE(i < b->bits, PE_ERR_RANGE, "bit index out of range");
    k_bitmap_set<<<1, 1, 0, b->ctx->stream>>>(b->p, i, 1);
    PE_HIP(hipGetLastError());
  });
}

int pe_bitmap_clr_bit(pe_bitmap_t* b, uint64_t i) {
  if (!b) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    PE_REQUIRE(i < b->bits, PE_ERR_RANGE, "bit index out of range");
    k_bitmap_set<<<1, 1, 0, b->ctx->stream>>>(b->p, i, 0);
    PE_HIP(hipGetLastError());
  });
}

int pe_bitmap_get_bit(pe_bitmap_t* b, uint64_t i, int* out) {
  if (!b || !out) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    PE_REQUIRE(i < b->bits, PE_ERR_RANGE, "bit index out of range");
    uint64_t w = 0;
    PE_HIP(hipMemcpyAsync(&w, b->p + (i >> 6), 8, hipMemcpyDeviceToHost, b->ctx->stream));
    PE_HIP(hipStreamSynchronize(b->ctx->stream));
    *out = (int)((w >> (i & 63)) & 1ull);
  });
}

int pe_bitmap_count(pe_bitmap_t* b, uint64_t* out) {
  if (!b || !out) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    dbuf<unsigned long long> c(b->ctx, 1);
    PE_HIP(hipMemsetAsync(c.p, 0, 8, b->ctx->stream));
    k_bitmap_count<<<grid_for(b->words, 256, 2048), 256, 0, b->ctx->stream>>>(b->p, b->words, c.p);
    PE_HIP(hipGetLastError());
    unsigned long long h = 0;
    PE_HIP(hipMemcpyAsync(&h, c.p, 8, hipMemcpyDeviceToHost, b->ctx->stream));
    PE_HIP(hipStreamSynchronize(b->ctx->stream));
    *out = h;
  });
}

int pe_bitmap_msb(pe_bitmap_t* b, uint64_t* out) {
  if (!b || !out) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    dbuf<unsigned long long> c(b->ctx, 1);
    PE_HIP(hipMemsetAsync(c.p, 0, 8, b->ctx->stream));
    k_bitmap_msb<<<grid_for(b->words, 256, 2048), 256, 0, b->ctx->stream>>>(b->p, b->words, c.p);
    PE_HIP(hipGetLastError());
    unsigned long long h = 0;
    PE_HIP(hipMemcpyAsync(&h, c.p, 8, hipMemcpyDeviceToHost, b->ctx->stream));
    PE_HIP(hipStreamSynchronize(b->ctx->stream));
    *out = h ? h - 1 : 0;
  });
}

int pe_bitmap_copy_from(pe_bitmap_t* dst, const pe_bitmap_t* src) {
  if (!dst || !src) return PE_ERR_INVALID;
  return guarded(dst->ctx, [&] {
    PE_REQUIRE(dst->words == src->words, PE_ERR_INVALID, "bitmap sizes differ");
    PE_HIP(hipMemcpyAsync(dst->p, src->p, dst->words * 8, hipMemcpyDeviceToDevice, dst->ctx->stream));
  });
}

int pe_bitmap_upload(pe_bitmap_t* b, const uint64_t* words, uint64_t nwords) {
  if (!b || (!words && nwords)) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    PE_REQUIRE(nwords <= b->words, PE_ERR_RANGE, "too many words");
    PE_HIP(hipMemcpyAsync(b->p, words, nwords * 8, hipMemcpyHostToDevice, b->ctx->stream));
    PE_HIP(hipStreamSynchronize(b->ctx->stream));
  });
}

int pe_bitmap_download(pe_bitmap_t* b, uint64_t* words, uint64_t nwords) {
  if (!b || (!words && nwords)) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    PE_REQUIRE(nwords <= b->words, PE_ERR_RANGE, "too many words");
    PE_HIP(hipMemcpyAsync(words, b->p, nwords * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    PE_HIP(hipStreamSynchronize(b->ctx->stream));
  });
}

int pe_bitmap_to_vector(pe_bitmap_t* b, uint32_t* ids, uint64_t cap, uint64_t* count) {
  if (!b || !count || (!ids && cap)) return PE_ERR_INVALID;
  return guarded(b->ctx, [&] {
    hipStream_t s = b->ctx->stream;
    uint64_t waves = (b->words + 63) / 64;
    dbuf<uint32_t> cnt(b->ctx, waves);
    dbuf<uint64_t> base(b->ctx, waves);
    k_bitmap_wave_counts<<<grid_for(waves * 64, 256), 256, 0, s>>>(b->p, b->words, cnt.p, waves);
    PE_HIP(hipGetLastError());
    std::vector<uint32_t> hc(waves);
    std::vector<uint64_t> hb(waves);
    PE_HIP(hipMemcpyAsync(hc.data(), cnt.p, waves * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    uint64_t total = 0;
    for (uint64_t i = 0; i < waves; ++i) { hb[i] = total; total += hc[i]; }
    *count = total;
    if (!cap) return;
    uint64_t n = total < cap ? total : cap;
    PE_HIP(hipMemcpyAsync(base.p, hb.data(), waves * 8, hipMemcpyHostToDevice, s));
    dbuf<uint32_t> out(b->ctx, n ? n : 1);
    k_bitmap_emit<<<grid_for(waves * 64, 256), 256, 0, s>>>(b->p, b->words, base.p, out.p, n, waves);
    PE_HIP(hipGetLastError());
    PE_HIP(hipMemcpyAsync(ids, out.p, n * 4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
