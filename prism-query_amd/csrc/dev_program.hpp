// Row-program definitions and device helpers shared by the prebuilt kernels
// and the kernels generated at run time (hipRTC).
#pragma once

#include "dev_common.hpp"

namespace prism {

constexpr int kMaxCols = 16;
constexpr int kMaxKeys = 8;
constexpr int kMaxValues = 16;
constexpr int kMaxLanes = 48;
constexpr int kMaxCode = 160;
constexpr int kMaxStrings = 16;
constexpr int kMaxStrBytes = 1024;
constexpr int kStackDepth = 8;
constexpr int kMaxPref = 8;
// Internal opcode (never in user code): push prefetched column slot `arg`.
constexpr int32_t kOpLoadPre = 64;

// Static result type of a program.
enum ValType : int32_t { VT_I = 0, VT_F = 1, VT_B = 2, VT_NULL = 3 };

struct ColsArg {
  prism_column c[kMaxCols];
};

// LIKE pattern shapes with a direct scan (no wildcard but leading / trailing
// '%', no '_', no escape); anything else runs the general matcher.
enum LikeKind : uint8_t { LK_GENERAL = 0, LK_CONTAINS = 1, LK_PREFIX = 2, LK_SUFFIX = 3, LK_EXACT = 4 };

struct StrTable {
  int32_t n;
  int32_t off[kMaxStrings + 1];
  uint8_t like_kind[kMaxStrings];
  uint8_t bytes[kMaxStrBytes];
};

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)


__device__ __forceinline__ bool bit_at(const uint8_t* bm, uint64_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }

__device__ __forceinline__ bool bit_at(const PRISM_GLOBAL uint8_t* bm, uint64_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }

__device__ __forceinline__ bool gbit(const void* bm, uint64_t i) {
  return bit_at((const PRISM_GLOBAL uint8_t*)(bm), i);
}

__device__ __forceinline__ bool col_valid(const prism_column& c, uint64_t row) {
  return c.validity == nullptr || gbit(c.validity, c.offset + row);
}

__device__ __forceinline__ int64_t col_load_i(const prism_column& c, uint64_t row) {
  const uint64_t i = c.offset + row;
  switch (c.type) {
    case PRISM_TYPE_INT64:
      return gload<int64_t>(c.data, i);
    case PRISM_TYPE_BOOL:
      return gbit(c.data, i) ? 1 : 0;
    default:  // INT32, DATE32
      return gload<int32_t>(c.data, i);
  }
}
__device__ __forceinline__ double col_load_f(const prism_column& c, uint64_t row) {
  return gload<double>(c.data, c.offset + row);
}
// Raw 64-bit word of a fixed-width column: int64 / f64 bits as stored, int32 /
// date32 sign-extended, bool as 0 / 1.
__device__ __forceinline__ uint64_t col_load_raw(const prism_column& c, uint64_t row) {
  const uint64_t i = c.offset + row;
  switch (c.type) {
    case PRISM_TYPE_INT64:
    case PRISM_TYPE_FLOAT64:
      return gload<uint64_t>(c.data, i);
    case PRISM_TYPE_BOOL:
      return gbit(c.data, i) ? 1 : 0;
    default:
      return static_cast<uint64_t>(static_cast<int64_t>(gload<int32_t>(c.data, i)));
  }
}
// FxHash word of a non-null fixed-width key from its raw word (the same words
// hash_key writes).
__device__ __forceinline__ uint64_t key_word(int32_t type, uint64_t raw) {
  switch (type) {
    case PRISM_TYPE_DATE32:
      return static_cast<uint64_t>(static_cast<uint32_t>(raw)) | (1ULL << 32);
    case PRISM_TYPE_BOOL:
      return 2 + raw;
    default:
      return raw;
  }
}

__device__ __forceinline__ double as_f(uint64_t b) { return __longlong_as_double(static_cast<long long>(b)); }
__device__ __forceinline__ uint64_t f_bits(double d) { return static_cast<uint64_t>(__double_as_longlong(d)); }

__device__ __forceinline__ bool cmp_i(int op, int64_t a, int64_t b) {
  switch (op) {
    case PRISM_CMP_EQ: return a == b;
    case PRISM_CMP_NE: return a != b;
    case PRISM_CMP_LT: return a < b;
    case PRISM_CMP_LE: return a <= b;
    case PRISM_CMP_GT: return a > b;
    default: return a >= b;
  }
}
// IEEE total order (f64::total_cmp): negative values have every bit but the
// sign flipped, then the words compare as signed integers.
__device__ __forceinline__ int64_t total_key(double d) {
  const int64_t b = static_cast<int64_t>(__double_as_longlong(d));
  return b ^ static_cast<int64_t>(static_cast<uint64_t>(b >> 63) >> 1);
}
__device__ __forceinline__ bool cmp_f(int op, double a, double b) { return cmp_i(op, total_key(a), total_key(b)); }

__device__ __forceinline__ int str_cmp(const uint8_t* a, int32_t na, const uint8_t* b, int32_t nb) {
  int32_t n = na < nb ? na : nb;
  for (int32_t i = 0; i < n; ++i) {
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  }
  return na == nb ? 0 : (na < nb ? -1 : 1);
}

__device__ __forceinline__ uint8_t fold_ascii(uint8_t c) { return (c >= 'A' && c <= 'Z') ? c + 32 : c; }

// Length of the UTF-8 sequence starting with byte c (1 for invalid lead bytes).
__device__ __forceinline__ int utf8_len(uint8_t c) {
  return c < 0x80 ? 1 : (c >> 5) == 6 ? 2 : (c >> 4) == 14 ? 3 : (c >> 3) == 30 ? 4 : 1;
}

// SQL LIKE: '%' any sequence, '_' one character, '\' escapes the next pattern
// character.  Greedy match with backtracking to the last '%' (linear for the
// common '%lit%' shapes).
// Bytes equal (ASCII letters folded when ci) between s and lit.
__device__ __forceinline__ bool bytes_eq(const uint8_t* s, const uint8_t* lit, int32_t m, bool ci) {
  for (int32_t j = 0; j < m; ++j) {
    const uint8_t a = s[j], b = lit[j];
    if (ci ? fold_ascii(a) != fold_ascii(b) : a != b) return false;
  }
  return true;
}

// Does s[0, n) contain lit[0, m)?  Four bytes per load: a SWAR test marks the
// bytes equal to lit[0] (folded for ILIKE), and only those starts are
// verified.  The aligned words may reach up to 3 bytes outside the string,
// never outside the 4-byte word holding its last byte.
__device__ __forceinline__ bool contains_lit(const uint8_t* s, int32_t n, const uint8_t* lit, int32_t m, bool ci) {
  if (m == 0) return true;
  if (n < m) return false;
  const uint8_t f0 = lit[0];
  const bool letter = ((f0 | 0x20) >= 'a' && (f0 | 0x20) <= 'z');
  const bool fold = ci && letter;
  const uint32_t first = (fold ? (f0 | 0x20u) : f0) * 0x01010101u;
  const uintptr_t a = reinterpret_cast<uintptr_t>(s);
  const uint32_t* w = reinterpret_cast<const uint32_t*>(a & ~static_cast<uintptr_t>(3));
  const int32_t lead = static_cast<int32_t>(a & 3);
  const int32_t last_start = n - m;  // last candidate position
  for (int32_t k = 0;; ++k) {
    const int32_t p0 = 4 * k - lead;  // string position of byte 0 of this word
    if (p0 > last_start) return false;
    uint32_t x = w[k];
    if (fold) x |= 0x20202020u;
    const uint32_t y = x ^ first;
    uint32_t hit = (y - 0x01010101u) & ~y & 0x80808080u;  // bytes of x equal to first (may over-report above a hit)
    while (hit) {
      const int32_t b = __builtin_ctz(hit) >> 3;
      hit &= hit - 1;
      const int32_t p = p0 + b;
      if (p < 0 || p > last_start) continue;
      if (bytes_eq(s + p + 1, lit + 1, m - 1, ci) && (ci ? fold_ascii(s[p]) == fold_ascii(f0) : s[p] == f0)) return true;
    }
  }
}

__device__ __forceinline__ bool like_fast(int kind, const uint8_t* s, int32_t n, const uint8_t* p, int32_t np, bool ci) {
  switch (kind) {
    case LK_CONTAINS: return np <= 2 || contains_lit(s, n, p + 1, np - 2, ci);  // "%" and "%%" match all
    case LK_PREFIX: return n >= np - 1 && bytes_eq(s, p, np - 1, ci);
    case LK_SUFFIX: return n >= np - 1 && bytes_eq(s + n - (np - 1), p + 1, np - 1, ci);
    default: return n == np && bytes_eq(s, p, np, ci);  // LK_EXACT
  }
}

__device__ inline bool like_match(const uint8_t* s, int32_t ns, const uint8_t* p, int32_t np, bool ci) {
  int32_t si = 0, pi = 0, star_p = -1, star_s = 0;
  while (si < ns) {
    if (pi < np) {
      uint8_t pc = p[pi];
      if (pc == '%') {
        star_p = ++pi;
        star_s = si;
        continue;
      }
      if (pc == '_') {
        si += utf8_len(s[si]);
        ++pi;
        continue;
      }
      int32_t pj = pi;
      if (pc == '\\' && pi + 1 < np) {
        pj = pi + 1;
        pc = p[pj];
      }
      uint8_t sc = s[si];
      if (ci ? fold_ascii(sc) == fold_ascii(pc) : sc == pc) {
        si += 1;
        pi = pj + 1;
        continue;
      }
    }
    if (star_p < 0) return false;
    star_s += utf8_len(s[star_s]);
    si = star_s;
    pi = star_p;
  }
  while (pi < np && p[pi] == '%') ++pi;
  return pi == np;
}

// str::hash: write(bytes) as zero-padded 8-byte LE words, then write_u8(0xff).
__device__ __forceinline__ uint64_t hash_str(const uint8_t* s, int32_t len, uint64_t state) {
  for (int32_t p = 0; p < len; p += 8) {
    uint64_t w = 0;
    const int32_t m = len - p < 8 ? len - p : 8;
    for (int32_t k = 0; k < m; ++k) w |= static_cast<uint64_t>(s[p + k]) << (8 * k);
    state = fx_step(state, w);
  }
  return fx_step(state, 0xff);
}

// Batched Utf8 hashing (same words as hash_str): the first up-to-8 bytes of a
// string come from the aligned 4-byte words it occupies, loaded together for
// all rows before any is used; only longer strings loop over the rest.
// Words past the last one holding a byte of the string repeat that one, so
// no load leaves the string's own words; an empty string reads `fallback`.
__device__ __forceinline__ void str_head_load(const uint8_t* s, int32_t len, const void* fallback, uint32_t& w0,
                                              uint32_t& w1, uint32_t& w2) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(s);
  const uint32_t* w = len > 0 ? reinterpret_cast<const uint32_t*>(a & ~static_cast<uintptr_t>(3))
                              : static_cast<const uint32_t*>(fallback);
  const int32_t last = len > 0 ? static_cast<int32_t>(((a + static_cast<uintptr_t>(len) - 1) >> 2) - (a >> 2)) : 0;
  w0 = w[0];
  w1 = w[last < 1 ? last : 1];
  w2 = w[last < 2 ? last : 2];
}
// str_head_load over a column in global memory (global, not flat, loads).
__device__ __forceinline__ void str_head_load(const PRISM_GLOBAL uint8_t* s, int32_t len, const PRISM_GLOBAL void* fallback,
                                              uint32_t& w0, uint32_t& w1, uint32_t& w2) {
  const uintptr_t a = (uintptr_t)(s);
  const PRISM_GLOBAL uint32_t* w = len > 0 ? (const PRISM_GLOBAL uint32_t*)(a & ~static_cast<uintptr_t>(3))
                                           : static_cast<const PRISM_GLOBAL uint32_t*>(fallback);
  const int32_t last = len > 0 ? static_cast<int32_t>(((a + static_cast<uintptr_t>(len) - 1) >> 2) - (a >> 2)) : 0;
  w0 = w[0];
  w1 = w[last < 1 ? last : 1];
  w2 = w[last < 2 ? last : 2];
}
// First min(len, 8) bytes of s as a zero-padded little-endian word.
__device__ __forceinline__ uint64_t str_head_word(const uint8_t* s, int32_t len, uint32_t w0, uint32_t w1,
                                                  uint32_t w2) {
  const uint32_t sh = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(s) & 3);
  const uint32_t lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
  const uint32_t hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
  const uint64_t v = (static_cast<uint64_t>(hi) << 32) | lo;
  return len >= 8 ? v : (len <= 0 ? 0 : v & ((1ULL << (8 * len)) - 1));
}
// hash_str from the head word: chunks after the first 8 bytes, then 0xff.
__device__ __forceinline__ uint64_t hash_str_head(const uint8_t* s, int32_t len, uint64_t head, uint64_t state) {
  if (len > 0) state = fx_step(state, head);
  if (len > 8) {
    for (int32_t p = 8; p < len; p += 8) {
      uint64_t w = 0;
      const int32_t m = len - p < 8 ? len - p : 8;
      for (int32_t k = 0; k < m; ++k) w |= static_cast<uint64_t>(s[p + k]) << (8 * k);
      state = fx_step(state, w);
    }
  }
  return fx_step(state, 0xff);
}

// FxHash word(s) of one key value (hash_aggregate.rs:24-75,492-504): null
// writes u8 0; Int32 sign-extends; Float64 hashes its bits; Utf8 writes its
// bytes as zero-padded 8-byte LE words then 0xff.  Date32 / Boolean use an
// injective word away from 0 (the reference hashes a Debug string there; only
// the partition into groups is observable).
__device__ __forceinline__ uint64_t hash_key(const prism_column& c, uint64_t row, uint64_t state) {
  if (!col_valid(c, row)) return fx_step(state, 0);  // 0u8.hash(): write_u8(0)
  const uint64_t i = c.offset + row;
  switch (c.type) {
    case PRISM_TYPE_INT64:
      return fx_step(state, static_cast<uint64_t>(gload<int64_t>(c.data, i)));
    case PRISM_TYPE_INT32:  // write_i32 -> write_u64(i as u64): sign extension
      return fx_step(state, static_cast<uint64_t>(static_cast<int64_t>(gload<int32_t>(c.data, i))));
    case PRISM_TYPE_FLOAT64:  // to_bits().hash()
      return fx_step(state, gload<uint64_t>(c.data, i));
    case PRISM_TYPE_DATE32:  // value word kept away from the null word 0
      return fx_step(state, static_cast<uint64_t>(static_cast<uint32_t>(gload<int32_t>(c.data, i))) | (1ULL << 32));
    case PRISM_TYPE_BOOL:
      return fx_step(state, 2 + (gbit(c.data, i) ? 1 : 0));
    default: {  // UTF8
      const int32_t b = gload<int32_t>(c.offsets, i), e = gload<int32_t>(c.offsets, i + 1);
      return hash_str(static_cast<const uint8_t*>(c.data) + b, e - b, state);
    }
  }
}

#endif  // __HIPCC__ || __HIPCC_RTC__

}  // namespace prism
