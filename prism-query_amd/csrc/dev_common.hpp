// Device helpers shared by the prebuilt kernels and the kernels generated at
// run time (hipRTC): no host-only headers here.
#pragma once

#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "prism_hip.h"

namespace prism {

// FxHasher multiplier (hash_aggregate.rs:47,53) and hash_i64_fast
// (hash_join.rs:371-373).
constexpr uint64_t kFxK = 0x517cc1b727220a95ULL;

__host__ __device__ __forceinline__ uint64_t fx_step(uint64_t state, uint64_t word) {
  return state * kFxK + word;
}
__host__ __device__ __forceinline__ uint64_t hash_i64_fast(int64_t v) {
  return static_cast<uint64_t>(v) * kFxK;
}

// Monotone u64 encoding of an f64 (total order with -0 < +0; NaN is never fed).
__host__ __device__ __forceinline__ uint64_t f64_order(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ __forceinline__ double f64_unorder(uint64_t o) {
  uint64_t b = (o >> 63) ? (o & 0x7fffffffffffffffULL) : ~o;
  double x;
  __builtin_memcpy(&x, &b, 8);
  return x;
}

// Native 128-bit vector of two int64 (usable with the non-temporal builtins).
typedef long long v2i64 __attribute__((ext_vector_type(2)));
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));

// Pointers typed as global memory.  A load through a generic pointer (any
// pointer read from memory, such as a column descriptor's buffers) is a flat
// load, which also counts against the LDS counter, so every wait for an LDS
// result waits for it too; the same load through a global pointer does not.
#define PRISM_GLOBAL __attribute__((address_space(1)))
#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
template <int N> struct U32Vec;
template <> struct U32Vec<2> { typedef unsigned int type __attribute__((ext_vector_type(2))); };
template <> struct U32Vec<4> { typedef v4u32 type; };
template <class T>
__device__ __forceinline__ T gload(const void* p, uint64_t i) {
  if constexpr (__is_scalar(T)) {
    return ((const PRISM_GLOBAL T*)(p))[i];
  } else {  // a struct of 8 or 16 bytes: one vector load (no copy constructor for a global lvalue)
    static_assert(sizeof(T) == 8 || sizeof(T) == 16, "8- or 16-byte records");
    typedef typename U32Vec<sizeof(T) / 4>::type W;
    const W w = ((const PRISM_GLOBAL W*)(p))[i];
    T t;
    __builtin_memcpy(&t, &w, sizeof(T));
    return t;
  }
}
#endif

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }
// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Exact int128 accumulation with 64-bit atomics: the low word is added with
// its old value returned, so the carry of this very addition is known and goes
// into the high word together with the addend's high word.  Order independent.
__device__ __forceinline__ void atomic_add_i128(unsigned long long* lo, unsigned long long* hi,
                                                uint64_t add_lo, int64_t add_hi) {
  unsigned long long old = atomicAdd(lo, static_cast<unsigned long long>(add_lo));
  uint64_t carry = (old + add_lo) < old ? 1u : 0u;
  uint64_t h = static_cast<uint64_t>(add_hi) + carry;
  if (h) atomicAdd(hi, static_cast<unsigned long long>(h));
}

// Bounded spin: the look-back waits end after this many polls and flag an error.
constexpr uint32_t kSpinLimit = 1u << 26;

// Split accumulator of an int128 sum of int64 addends: low 32-bit halves summed
// unsigned, high 32-bit halves summed signed.  Neither overflows for < 2^31 addends.
struct Split128 {
  uint64_t lo32 = 0;
  int64_t hi32 = 0;
  __host__ __device__ __forceinline__ void add(int64_t v) {
    lo32 += static_cast<uint32_t>(static_cast<uint64_t>(v));
    hi32 += static_cast<int64_t>(v >> 32);  // arithmetic shift: floor(v / 2^32)
  }
  // value = hi32 * 2^32 + lo32 as (lo, hi) int128 words.
  __host__ __device__ __forceinline__ void to_i128(uint64_t& lo, int64_t& hi) const {
    uint64_t hl = static_cast<uint64_t>(hi32) << 32;
    int64_t hh = hi32 >> 32;
    lo = hl + lo32;
    hi = hh + ((lo < hl) ? 1 : 0);
  }
};

__host__ __device__ __forceinline__ uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace prism
