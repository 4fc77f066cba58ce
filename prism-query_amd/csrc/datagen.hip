#include <hipcub/hipcub.hpp>
// Seeded synthetic columns for tests and benchmarks.  Counter-based and
// integer-only, so the device output is bit-identical to the numpy twin
// (prism_executor/datagen.py) at any size.
#include "common.hpp"

namespace prism {
namespace gen {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

__host__ __device__ __forceinline__ uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t stream_base(uint64_t seed, uint32_t stream) {
  return mix(seed * kGolden + static_cast<uint64_t>(stream) + 1);
}
__host__ __device__ __forceinline__ uint64_t rnd(uint64_t base, uint64_t i) {
  return mix(base + (i + 1) * kGolden);
}
__host__ __device__ __forceinline__ uint64_t uniform(uint64_t r, uint64_t range) {
  return ((r >> 32) * range) >> 32;
}

struct Feistel {
  uint64_t n;
  uint32_t half;
  uint64_t mask;
  uint64_t keys[4];
};

inline Feistel make_feistel(uint64_t n, uint64_t seed, uint32_t stream) {
  Feistel f;
  f.n = n;
  uint32_t bits = 2;
  while ((1ULL << bits) < n) ++bits;
  if (bits & 1) ++bits;
  f.half = bits / 2;
  f.mask = (1ULL << f.half) - 1;
  uint64_t base = stream_base(seed, stream);
  for (int r = 0; r < 4; ++r) f.keys[r] = rnd(base, static_cast<uint64_t>(r));
  return f;
}

__host__ __device__ __forceinline__ uint64_t feistel_once(const Feistel& f, uint64_t x) {
  uint64_t L = x >> f.half, R = x & f.mask;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    uint64_t F = mix(R ^ f.keys[r]) & f.mask;
    uint64_t nl = R;
    R = L ^ F;
    L = nl;
  }
  return (L << f.half) | R;
}
// Cycle walking keeps the bijection inside [0, n): the orbit of x < n returns
// below n no later than at x itself.
__host__ __device__ __forceinline__ uint64_t permute(const Feistel& f, uint64_t x) {
  do {
    x = feistel_once(f, x);
  } while (x >= f.n);
  return x;
}

}  // namespace gen

__global__ void k_uniform_i64(int64_t* out, uint64_t n, uint64_t base, int64_t lo, uint64_t range) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = lo + static_cast<int64_t>(gen::uniform(gen::rnd(base, i), range));
}
__global__ void k_uniform_i32(int32_t* out, uint64_t n, uint64_t base, int64_t lo, uint64_t range) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = static_cast<int32_t>(lo + static_cast<int64_t>(gen::uniform(gen::rnd(base, i), range)));
}
__global__ void k_bits_i64(int64_t* out, uint64_t n, uint64_t base) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = static_cast<int64_t>(gen::rnd(base, i));
}
__global__ void k_permutation(int64_t* out, uint64_t n, gen::Feistel f) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = static_cast<int64_t>(gen::permute(f, i));
}
__global__ void k_zipf(int64_t* out, uint64_t n, uint64_t base, const uint64_t* __restrict__ cdf,
                       uint64_t V, gen::Feistel f) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t u = gen::rnd(base, i);
    // first j with cdf[j] > u (clamped to V-1)
    uint64_t lo = 0, hi = V - 1;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    out[i] = static_cast<int64_t>(gen::permute(f, lo));
  }
}
struct Alphabet { uint8_t c[16]; };
__global__ void k_flag_utf8(uint8_t* bytes, int32_t* offsets, uint64_t n, uint64_t base, Alphabet a, uint32_t na) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
    offsets[i] = static_cast<int32_t>(i);
    if (i < n) bytes[i] = a.c[gen::uniform(gen::rnd(base, i), na)];
  }
}

// Random lowercase strings: row lengths first (plus a trailing 0, so that an
// exclusive scan gives the n + 1 offsets), then byte p of the data buffer.
__global__ void k_lower_lengths(int32_t* len, uint64_t n, uint64_t base, uint32_t min_len, uint32_t span) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
    len[i] = i < n ? static_cast<int32_t>(min_len + gen::uniform(gen::rnd(base, i), span)) : 0;
}

__global__ void k_lower_bytes(uint8_t* bytes, uint64_t total, uint64_t base) {
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < total; p += (uint64_t)gridDim.x * blockDim.x)
    bytes[p] = static_cast<uint8_t>('a' + gen::uniform(gen::rnd(base, p), 26));
}

static dim3 gen_grid(prism_ctx* ctx, uint64_t n) {
  uint64_t blocks = ceil_div(n ? n : 1, 256);
  uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 16;
  return dim3(static_cast<unsigned>(blocks < cap ? blocks : cap));
}

static void check_range(uint64_t range) {
  if (range == 0 || range > (1ULL << 32))
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: range must be in [1, 2^32]");
}

}  // namespace prism

using namespace prism;

extern "C" {

prism_status prism_datagen_uniform_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n, uint64_t seed,
                                       uint32_t stream, int64_t lo, uint64_t range) {
  return guarded([&] {
    check_range(range);
    use_device(ctx);
    hipLaunchKernelGGL(k_uniform_i64, gen_grid(ctx, n), dim3(256), 0, ctx->stream, d_out, n,
                       gen::stream_base(seed, stream), lo, range);
    PRISM_CHECK_LAUNCH();
  });
}

prism_status prism_datagen_uniform_i32(prism_ctx* ctx, int32_t* d_out, uint64_t n, uint64_t seed,
                                       uint32_t stream, int64_t lo, uint64_t range) {
  return guarded([&] {
    check_range(range);
    use_device(ctx);
    hipLaunchKernelGGL(k_uniform_i32, gen_grid(ctx, n), dim3(256), 0, ctx->stream, d_out, n,
                       gen::stream_base(seed, stream), lo, range);
    PRISM_CHECK_LAUNCH();
  });
}

prism_status prism_datagen_bits_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n, uint64_t seed,
                                    uint32_t stream) {
  return guarded([&] {
    use_device(ctx);
    hipLaunchKernelGGL(k_bits_i64, gen_grid(ctx, n), dim3(256), 0, ctx->stream, d_out, n,
                       gen::stream_base(seed, stream));
    PRISM_CHECK_LAUNCH();
  });
}

prism_status prism_datagen_permutation_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n,
                                           uint64_t seed, uint32_t stream) {
  return guarded([&] {
    if (n == 0) return;
    use_device(ctx);
    hipLaunchKernelGGL(k_permutation, gen_grid(ctx, n), dim3(256), 0, ctx->stream, d_out, n,
                       gen::make_feistel(n, seed, stream));
    PRISM_CHECK_LAUNCH();
  });
}

prism_status prism_datagen_zipf_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n, uint64_t seed,
                                    uint32_t stream, const uint64_t* d_cdf, uint64_t n_values,
                                    uint32_t perm_stream) {
  return guarded([&] {
    if (n_values == 0 || !d_cdf) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: empty CDF");
    use_device(ctx);
    hipLaunchKernelGGL(k_zipf, gen_grid(ctx, n), dim3(256), 0, ctx->stream, d_out, n,
                       gen::stream_base(seed, stream), d_cdf, n_values,
                       gen::make_feistel(n_values, seed, perm_stream));
    PRISM_CHECK_LAUNCH();
  });
}

prism_status prism_datagen_flag_utf8(prism_ctx* ctx, uint8_t* d_bytes, int32_t* d_offsets,
                                     uint64_t n, uint64_t seed, uint32_t stream,
                                     const char* alphabet, uint32_t n_alpha) {
  return guarded([&] {
    if (n_alpha == 0 || n_alpha > 16)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: alphabet must hold 1..16 bytes");
    if (n >= (1ULL << 31)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: utf8 column too long");
    Alphabet a{};
    memcpy(a.c, alphabet, n_alpha);
    use_device(ctx);
    hipLaunchKernelGGL(k_flag_utf8, gen_grid(ctx, n + 1), dim3(256), 0, ctx->stream, d_bytes,
                       d_offsets, n, gen::stream_base(seed, stream), a, n_alpha);
    PRISM_CHECK_LAUNCH();
  });
}

prism_status prism_datagen_lower_utf8(prism_ctx* ctx, uint8_t* d_bytes, uint64_t bytes_capacity, int32_t* d_offsets,
                                      uint64_t n, uint64_t seed, uint32_t stream, uint32_t min_len, uint32_t max_len,
                                      uint64_t* total_bytes) {
  return guarded([&] {
    if (!ctx || !d_offsets || !total_bytes) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL argument");
    if (min_len > max_len) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: min_len > max_len");
    if (static_cast<double>(n) * max_len >= 2147483647.0)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: utf8 data would exceed i32 offsets");
    use_device(ctx);
    Ctx* c = as_ctx(ctx);
    hipStream_t st = ctx->stream;
    DeviceBuffer len(c, (n + 1) * 4);
    hipLaunchKernelGGL(k_lower_lengths, gen_grid(ctx, n + 1), dim3(256), 0, st, len.as<int32_t>(), n,
                       gen::stream_base(seed, stream), min_len, max_len - min_len + 1);
    PRISM_CHECK_LAUNCH();
    size_t tmp = 0;
    PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, len.as<int32_t>(), d_offsets, static_cast<int>(n + 1), st));
    DeviceBuffer t(c, tmp ? tmp : 1);
    PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(t.ptr, tmp, len.as<int32_t>(), d_offsets, static_cast<int>(n + 1), st));
    int32_t total = 0;
    PRISM_HIP(hipMemcpyAsync(&total, d_offsets + n, 4, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    *total_bytes = static_cast<uint64_t>(total);
    if (static_cast<uint64_t>(total) > bytes_capacity)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: byte buffer too small (" + std::to_string(total) + " needed)");
    if (total) {
      hipLaunchKernelGGL(k_lower_bytes, gen_grid(ctx, static_cast<uint64_t>(total)), dim3(256), 0, st, d_bytes,
                         static_cast<uint64_t>(total), gen::stream_base(seed, stream + 1));
      PRISM_CHECK_LAUNCH();
    }
  });
}

}  // extern "C"
