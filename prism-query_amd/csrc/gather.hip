// Gather ("take") and selection-vector kernels: assembling join outputs
// (compute::take, hash_join.rs:424-453,569-625) and filtered batches
// (FilterBuilder, filter_project.rs:126-135) in HBM.
#include <hipcub/hipcub.hpp>

#include "gather.hpp"

namespace prism {

constexpr int kGB = 256;

static dim3 ggrid(prism_ctx* ctx, uint64_t n) {
  uint64_t b = ceil_div(n ? n : 1, kGB);
  uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 16;
  return dim3(static_cast<unsigned>(b < cap ? b : cap));
}

__device__ __forceinline__ bool bitv(const uint8_t* bm, uint64_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }

// Every take kernel checks its valid indices against the source length; an
// index past it reads nothing, and sets *err (the host then fails the call).
template <class T>
__global__ void k_take_fixed(const T* __restrict__ src, int64_t src_off, uint64_t src_len,
                             const uint32_t* __restrict__ idx, const uint8_t* __restrict__ idx_valid, uint64_t n,
                             T* __restrict__ dst, uint32_t* __restrict__ err) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    bool ok = idx_valid == nullptr || bitv(idx_valid, i);
    const uint32_t r = ok ? idx[i] : 0;
    if (ok && r >= src_len) {
      *err = 1;
      ok = false;
    }
    dst[i] = ok ? src[src_off + r] : T{};
  }
}

// Bit gather, one 64-row word per wave: out bit i = index valid && src bit.
// src == nullptr means "every source row is set" (validity of a column without nulls).
__global__ void k_take_bits(const uint8_t* __restrict__ src, int64_t src_off, uint64_t src_len,
                            const uint32_t* __restrict__ idx, const uint8_t* __restrict__ idx_valid, uint64_t n,
                            unsigned long long* __restrict__ dst, uint32_t* __restrict__ err) {
  const uint64_t n_words = ceil_div(n, 64);
  const uint64_t waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) / 64;
  const uint64_t wave0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64;
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t w = wave0; w < n_words; w += waves) {
    uint64_t i = w * 64 + lane;
    bool bit = false;
    if (i < n) {
      bool ok = idx_valid == nullptr || bitv(idx_valid, i);
      const uint32_t r = ok ? idx[i] : 0;
      if (ok && r >= src_len) {
        *err = 1;
        ok = false;
      }
      bit = ok && (src == nullptr || bitv(src, src_off + r));
    }
    uint64_t word = __ballot(bit);
    if (lane == 0) dst[w] = word;
  }
}

// Offsets were checked non-decreasing when the column was uploaded, so each
// length is >= 0 and the copy below stays inside the scanned extents.
__global__ void k_take_str_len(const int32_t* __restrict__ off, int64_t src_off, uint64_t src_len,
                               const uint32_t* __restrict__ idx, const uint8_t* __restrict__ idx_valid, uint64_t n,
                               uint64_t* __restrict__ len, uint32_t* __restrict__ err) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    bool ok = idx_valid == nullptr || bitv(idx_valid, i);
    const uint32_t ri = ok ? idx[i] : 0;
    if (ok && ri >= src_len) {
      *err = 1;
      ok = false;
    }
    const uint64_t r = src_off + ri;
    len[i] = ok ? static_cast<uint64_t>(off[r + 1] - off[r]) : 0;
  }
}

__global__ void k_take_str_copy(const int32_t* __restrict__ off, int64_t src_off, uint64_t src_len,
                                const uint8_t* __restrict__ bytes, const uint32_t* __restrict__ idx,
                                const uint8_t* __restrict__ idx_valid, uint64_t n, const uint64_t* __restrict__ start,
                                int32_t* __restrict__ out_off, uint8_t* __restrict__ out_bytes) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = start[i];
    out_off[i] = static_cast<int32_t>(s);
    bool ok = idx_valid == nullptr || bitv(idx_valid, i);
    if (!ok || idx[i] >= src_len) continue;
    uint64_t r = src_off + idx[i];
    const int32_t b = off[r], e = off[r + 1];
    for (int32_t k = b; k < e; ++k) out_bytes[s + (k - b)] = bytes[k];
  }
}

__global__ void k_set_last_offset(int32_t* out_off, uint64_t n, const uint64_t* __restrict__ start,
                                  const uint64_t* __restrict__ len) {
  out_off[n] = n ? static_cast<int32_t>(start[n - 1] + len[n - 1]) : 0;
}

__global__ void k_mark(const uint32_t* __restrict__ idx, uint64_t n, uint8_t* __restrict__ flags) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    flags[idx[i]] = 1;
}

__global__ void k_flags_to_bits(const uint8_t* __restrict__ flags, uint64_t n, uint8_t want,
                                unsigned long long* __restrict__ dst) {
  const uint64_t n_words = ceil_div(n, 64);
  const uint64_t waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) / 64;
  const uint64_t wave0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64;
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t w = wave0; w < n_words; w += waves) {
    uint64_t i = w * 64 + lane;
    uint64_t word = __ballot(i < n && (flags[i] != 0) == (want != 0));
    if (lane == 0) dst[w] = word;
  }
}

// Selection vector from a bitmask: word counts -> exclusive scan -> ordered
// row indices (the stable order FilterBuilder produces).
__global__ void k_word_counts(const unsigned long long* __restrict__ bits, uint64_t n_words, uint32_t* __restrict__ cnt) {
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < n_words; w += (uint64_t)gridDim.x * blockDim.x)
    cnt[w] = __popcll(bits[w]);
}

__global__ void k_word_select(const unsigned long long* __restrict__ bits, uint64_t n_words,
                              const uint64_t* __restrict__ start, uint32_t* __restrict__ out) {
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < n_words; w += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long b = bits[w];
    uint64_t p = start[w];
    while (b) {
      int k = __builtin_ctzll(b);
      out[p++] = static_cast<uint32_t>(w * 64 + k);
      b &= b - 1;
    }
  }
}

__global__ void k_iota_base(uint32_t* out, uint64_t n, uint32_t base) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = base + static_cast<uint32_t>(i);
}

__global__ void k_fill_bytes(uint8_t* p, uint64_t n, uint8_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_concat_u32(uint32_t* dst, const uint32_t* __restrict__ src, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Host wrappers.
// ---------------------------------------------------------------------------
DevColumn take(prism_ctx* ctx, const DevColumn& src, const uint32_t* idx, const uint8_t* idx_valid, uint64_t n,
               uint32_t* err_word) {
  Ctx* c = as_ctx(ctx);
  hipStream_t st = ctx->stream;
  DevColumn d;
  d.type = src.type;
  d.name = src.name;
  d.nullable = src.nullable || idx_valid != nullptr;
  d.length = static_cast<int64_t>(n);
  d.dev_offset = 0;
  const dim3 g = ggrid(ctx, n), b(kGB);
  const uint64_t sl = src.length > 0 ? static_cast<uint64_t>(src.length) : 0;
  // With the caller's error word the kernels only set it; else an own word is
  // read back here (a synchronisation).
  DeviceBuffer err;
  uint32_t* e = err_word;
  if (!e) {
    err.alloc(c, 4);
    fill_bytes(ctx, err.as<uint8_t>(), 4, 0);
    e = err.as<uint32_t>();
  }
  auto check_err = [&] {
    if (err_word) return;
    uint32_t h = 0;
    PRISM_HIP(hipMemcpyAsync(&h, e, 4, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    if (h)
      fail(PRISM_ERR_ARROW, "Arrow error: take index out of bounds (source length " + std::to_string(sl) + ")");
  };
  switch (src.type) {
    case DType::Int64: case DType::Float64:
      d.data.alloc(c, n ? n * 8 : 8);
      if (n) hipLaunchKernelGGL(k_take_fixed<uint64_t>, g, b, 0, st, src.data.as<uint64_t>(), src.dev_offset, sl, idx,
                                idx_valid, n, d.data.as<uint64_t>(), e);
      break;
    case DType::Int32: case DType::Date32: case DType::UInt32:
      d.data.alloc(c, n ? n * 4 : 4);
      if (n) hipLaunchKernelGGL(k_take_fixed<uint32_t>, g, b, 0, st, src.data.as<uint32_t>(), src.dev_offset, sl, idx,
                                idx_valid, n, d.data.as<uint32_t>(), e);
      break;
    case DType::Decimal128:
      d.data.alloc(c, n ? n * 16 : 16);
      if (n) hipLaunchKernelGGL(k_take_fixed<uint4>, g, b, 0, st, src.data.as<uint4>(), src.dev_offset, sl, idx,
                                idx_valid, n, d.data.as<uint4>(), e);
      break;
    case DType::Bool:
      d.data.alloc(c, ceil_div(n ? n : 1, 64) * 8);
      if (n) hipLaunchKernelGGL(k_take_bits, g, b, 0, st, src.data.as<uint8_t>(), src.dev_offset, sl, idx, idx_valid,
                                n, d.data.as<unsigned long long>(), e);
      break;
    case DType::Utf8: {
      DeviceBuffer len(c, (n ? n : 1) * 8), start(c, (n ? n : 1) * 8);
      d.offsets.alloc(c, (n + 1) * 4);
      if (n) {
        hipLaunchKernelGGL(k_take_str_len, g, b, 0, st, src.offsets.as<int32_t>(), src.dev_offset, sl, idx, idx_valid,
                           n, len.as<uint64_t>(), e);
        PRISM_CHECK_LAUNCH();
        check_err();
        size_t tmp = 0;
        PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, len.as<uint64_t>(), start.as<uint64_t>(),
                                                   static_cast<int>(n), st));
        DeviceBuffer t(c, tmp);
        PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(t.ptr, tmp, len.as<uint64_t>(), start.as<uint64_t>(),
                                                   static_cast<int>(n), st));
        uint64_t last[2] = {0, 0};
        PRISM_HIP(hipMemcpyAsync(&last[0], start.as<uint64_t>() + n - 1, 8, hipMemcpyDeviceToHost, st));
        PRISM_HIP(hipMemcpyAsync(&last[1], len.as<uint64_t>() + n - 1, 8, hipMemcpyDeviceToHost, st));
        PRISM_HIP(hipStreamSynchronize(st));
        const uint64_t total = last[0] + last[1];
        if (total > 0x7fffffffULL) fail(PRISM_ERR_ARROW, "Arrow error: Utf8 output exceeds i32 offsets");
        d.data.alloc(c, total ? total : 1);
        hipLaunchKernelGGL(k_take_str_copy, g, b, 0, st, src.offsets.as<int32_t>(), src.dev_offset, sl,
                           src.data.as<uint8_t>(), idx, idx_valid, n, start.as<uint64_t>(), d.offsets.as<int32_t>(),
                           d.data.as<uint8_t>());
        hipLaunchKernelGGL(k_set_last_offset, dim3(1), dim3(1), 0, st, d.offsets.as<int32_t>(), n,
                           start.as<uint64_t>(), len.as<uint64_t>());
        PRISM_CHECK_LAUNCH();
        PRISM_HIP(hipStreamSynchronize(st));
      } else {
        d.data.alloc(c, 1);
        hipLaunchKernelGGL(k_set_last_offset, dim3(1), dim3(1), 0, st, d.offsets.as<int32_t>(), 0, nullptr, nullptr);
      }
      break;
    }
  }
  PRISM_CHECK_LAUNCH();
  if (src.validity.ptr || idx_valid) {
    d.validity.alloc(c, ceil_div(n ? n : 1, 64) * 8);
    if (n) hipLaunchKernelGGL(k_take_bits, g, b, 0, st, src.validity.as<uint8_t>(), src.dev_offset, sl, idx,
                              idx_valid, n, d.validity.as<unsigned long long>(), e);
    PRISM_CHECK_LAUNCH();
  }
  if (n) check_err();
  return d;
}

void mark_rows(prism_ctx* ctx, const uint32_t* idx, uint64_t n, uint8_t* flags, uint64_t n_rows) {
  hipLaunchKernelGGL(k_fill_bytes, ggrid(ctx, n_rows), dim3(kGB), 0, ctx->stream, flags, n_rows, 0);
  if (n) hipLaunchKernelGGL(k_mark, ggrid(ctx, n), dim3(kGB), 0, ctx->stream, idx, n, flags);
  PRISM_CHECK_LAUNCH();
}

void flags_to_bits(prism_ctx* ctx, const uint8_t* flags, uint64_t n, bool want, unsigned long long* bits) {
  if (n) hipLaunchKernelGGL(k_flags_to_bits, ggrid(ctx, n), dim3(kGB), 0, ctx->stream, flags, n, want ? 1 : 0, bits);
  PRISM_CHECK_LAUNCH();
}

uint64_t select_bits(prism_ctx* ctx, const unsigned long long* bits, uint64_t n_rows, DeviceBuffer& out) {
  Ctx* c = as_ctx(ctx);
  hipStream_t st = ctx->stream;
  const uint64_t nw = ceil_div(n_rows, 64);
  if (nw == 0) {
    out.reserve(c, 4);
    return 0;
  }
  DeviceBuffer cnt(c, nw * 4), cnt64(c, nw * 8), start(c, nw * 8);
  hipLaunchKernelGGL(k_word_counts, ggrid(ctx, nw), dim3(kGB), 0, st, bits, nw, cnt.as<uint32_t>());
  PRISM_CHECK_LAUNCH();
  // widen to u64 before the scan so that totals beyond 2^32 cannot wrap
  size_t tmp = 0;
  auto in_it = cnt.as<uint32_t>();
  PRISM_HIP(hipcub::DeviceScan::ExclusiveScan(nullptr, tmp, in_it, start.as<uint64_t>(), hipcub::Sum(), uint64_t(0),
                                              static_cast<int>(nw), st));
  DeviceBuffer t(c, tmp);
  PRISM_HIP(hipcub::DeviceScan::ExclusiveScan(t.ptr, tmp, in_it, start.as<uint64_t>(), hipcub::Sum(), uint64_t(0),
                                              static_cast<int>(nw), st));
  uint64_t last_start = 0;
  uint32_t last_cnt = 0;
  PRISM_HIP(hipMemcpyAsync(&last_start, start.as<uint64_t>() + nw - 1, 8, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipMemcpyAsync(&last_cnt, cnt.as<uint32_t>() + nw - 1, 4, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipStreamSynchronize(st));
  const uint64_t total = last_start + last_cnt;
  out.reserve(c, (total ? total : 1) * 4);
  hipLaunchKernelGGL(k_word_select, ggrid(ctx, nw), dim3(kGB), 0, st, bits, nw, start.as<uint64_t>(),
                     out.as<uint32_t>());
  PRISM_CHECK_LAUNCH();
  PRISM_HIP(hipStreamSynchronize(st));
  return total;
}

void iota(prism_ctx* ctx, uint32_t* out, uint64_t n, uint32_t base) {
  if (n) hipLaunchKernelGGL(k_iota_base, ggrid(ctx, n), dim3(kGB), 0, ctx->stream, out, n, base);
  PRISM_CHECK_LAUNCH();
}

void fill_bytes(prism_ctx* ctx, uint8_t* p, uint64_t n, uint8_t v) {
  if (n) hipLaunchKernelGGL(k_fill_bytes, ggrid(ctx, n), dim3(kGB), 0, ctx->stream, p, n, v);
  PRISM_CHECK_LAUNCH();
}

void copy_u32(prism_ctx* ctx, uint32_t* dst, const uint32_t* src, uint64_t n) {
  if (n) hipLaunchKernelGGL(k_concat_u32, ggrid(ctx, n), dim3(kGB), 0, ctx->stream, dst, src, n);
  PRISM_CHECK_LAUNCH();
}

}  // namespace prism
