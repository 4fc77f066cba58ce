// Row programs: the fused predicate / projection stage evaluated inside the
// aggregation and filter kernels.  Predicate (filter_project.rs:22-51) and
// ScalarExpr (filter_project.rs:195-213) trees are compiled on the host into
// postfix code; every lane runs the same instruction at the same time, so the
// dispatch is scalar and the value stack lives in VGPRs (uniform stack index).
#pragma once

#include "common.hpp"

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

// Host-side checks: validates stack discipline and operand types, returns the
// result type.  Throws Error(InvalidArgument) on a malformed program.
ValType check_program(const prism_instr* code, uint32_t len, const int32_t* col_types, uint32_t n_cols,
                      uint32_t n_strings, int* max_depth);

// Build the device string table.
void build_str_table(const prism_string* s, uint32_t n, StrTable& out);

#if defined(__HIPCC__)

__device__ __forceinline__ bool bit_at(const uint8_t* bm, uint64_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }

__device__ __forceinline__ bool col_valid(const prism_column& c, uint64_t row) {
  return c.validity == nullptr || bit_at(c.validity, c.offset + row);
}

__device__ __forceinline__ int64_t col_load_i(const prism_column& c, uint64_t row) {
  const uint64_t i = c.offset + row;
  switch (c.type) {
    case PRISM_TYPE_INT64:
      return static_cast<const int64_t*>(c.data)[i];
    case PRISM_TYPE_BOOL:
      return bit_at(static_cast<const uint8_t*>(c.data), i) ? 1 : 0;
    default:  // INT32, DATE32
      return static_cast<const int32_t*>(c.data)[i];
  }
}
__device__ __forceinline__ double col_load_f(const prism_column& c, uint64_t row) {
  return static_cast<const double*>(c.data)[c.offset + row];
}
// Raw 64-bit word of a fixed-width column: int64 / f64 bits as stored, int32 /
// date32 sign-extended, bool as 0 / 1.
__device__ __forceinline__ uint64_t col_load_raw(const prism_column& c, uint64_t row) {
  const uint64_t i = c.offset + row;
  switch (c.type) {
    case PRISM_TYPE_INT64:
    case PRISM_TYPE_FLOAT64:
      return static_cast<const uint64_t*>(c.data)[i];
    case PRISM_TYPE_BOOL:
      return bit_at(static_cast<const uint8_t*>(c.data), i) ? 1 : 0;
    default:
      return static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int32_t*>(c.data)[i]));
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
      return fx_step(state, static_cast<uint64_t>(static_cast<const int64_t*>(c.data)[i]));
    case PRISM_TYPE_INT32:  // write_i32 -> write_u64(i as u64): sign extension
      return fx_step(state, static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int32_t*>(c.data)[i])));
    case PRISM_TYPE_FLOAT64:  // to_bits().hash()
      return fx_step(state, static_cast<const uint64_t*>(c.data)[i]);
    case PRISM_TYPE_DATE32:  // value word kept away from the null word 0
      return fx_step(state, static_cast<uint64_t>(static_cast<uint32_t>(static_cast<const int32_t*>(c.data)[i])) |
                                (1ULL << 32));
    case PRISM_TYPE_BOOL:
      return fx_step(state, 2 + (bit_at(static_cast<const uint8_t*>(c.data), i) ? 1 : 0));
    default: {  // UTF8: str::hash = write(bytes) in 8-byte LE words, then write_u8(0xff)
      const int32_t b = c.offsets[i], e = c.offsets[i + 1];
      const uint8_t* s = static_cast<const uint8_t*>(c.data) + b;
      const int32_t len = e - b;
      for (int32_t p = 0; p < len; p += 8) {
        uint64_t w = 0;
        const int32_t m = len - p < 8 ? len - p : 8;
        for (int32_t k = 0; k < m; ++k) w |= static_cast<uint64_t>(s[p + k]) << (8 * k);
        state = fx_step(state, w);
      }
      return fx_step(state, 0xff);
    }
  }
}

// Evaluate `len` instructions for R rows held by this lane.  Values are raw
// 64-bit words (int64 or f64 bits); `valid` bit r = row r non-null.
// `pre` holds NP prefetched fixed-width columns (kOpLoadPre reads them);
// bit j*R + r of pre_v = row r of slot j is non-null.
template <int R, int DEPTH, int NP>
__device__ __forceinline__ void run_program(const prism_instr* __restrict__ code, int len,
                                            const prism_column* __restrict__ cols,
                                            const StrTable* __restrict__ strs, const uint64_t (&row)[R],
                                            const bool (&in)[R], const uint64_t (&pre)[NP][R], uint32_t pre_v,
                                            uint64_t (&out)[R], uint32_t& out_valid, uint32_t* err) {
  static_assert(NP * R <= 32, "prefetch validity bits fit one word");
  static_assert(DEPTH * R <= 32, "validity bits fit one word");
  uint64_t st[DEPTH * R];
  // Validity bits of stack slot k live in bits [k*R, k*R + R) of one word.
  uint32_t vbits = 0;
  constexpr uint32_t kAll = (1u << R) - 1;
  auto vget = [&](int k) -> uint32_t { return (vbits >> (k * R)) & kAll; };
  auto vset = [&](int k, uint32_t v) { vbits = (vbits & ~(kAll << (k * R))) | ((v & kAll) << (k * R)); };
  int sp = 0;
  for (int pc = 0; pc < len; ++pc) {
    const int op = code[pc].op;
    const int arg = code[pc].arg;
    switch (op) {
      case kOpLoadPre: {
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < NP; ++j)
          if (arg == j) {
#pragma unroll
            for (int r = 0; r < R; ++r) st[sp * R + r] = pre[j][r];
            v = (pre_v >> (j * R)) & kAll;
          }
        vset(sp++, v);
        break;
      }
      case PRISM_OP_LOAD_I: {
        const prism_column& c = cols[arg];
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          st[sp * R + r] = in[r] ? static_cast<uint64_t>(col_load_i(c, row[r])) : 0;
          v |= (in[r] && col_valid(c, row[r])) ? (1u << r) : 0u;
        }
        vset(sp++, v);
        break;
      }
      case PRISM_OP_LOAD_F: {
        const prism_column& c = cols[arg];
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          st[sp * R + r] = in[r] ? f_bits(col_load_f(c, row[r])) : 0;
          v |= (in[r] && col_valid(c, row[r])) ? (1u << r) : 0u;
        }
        vset(sp++, v);
        break;
      }
      case PRISM_OP_CONST_I:
      case PRISM_OP_CONST_BOOL: {
        const uint64_t x = op == PRISM_OP_CONST_I ? static_cast<uint64_t>(code[pc].i64) : (code[pc].i64 != 0);
#pragma unroll
        for (int r = 0; r < R; ++r) st[sp * R + r] = x;
        vset(sp++, kAll);
        break;
      }
      case PRISM_OP_CONST_F: {
        const uint64_t x = f_bits(code[pc].f64);
#pragma unroll
        for (int r = 0; r < R; ++r) st[sp * R + r] = x;
        vset(sp++, kAll);
        break;
      }
      case PRISM_OP_CONST_NULL: {
#pragma unroll
        for (int r = 0; r < R; ++r) st[sp * R + r] = 0;
        vset(sp++, 0);
        break;
      }
      case PRISM_OP_I2F: {
#pragma unroll
        for (int r = 0; r < R; ++r) st[(sp - 1) * R + r] = f_bits(static_cast<double>(static_cast<int64_t>(st[(sp - 1) * R + r])));
        break;
      }
      case PRISM_OP_NEG_F: {
#pragma unroll
        for (int r = 0; r < R; ++r) st[(sp - 1) * R + r] = f_bits(-as_f(st[(sp - 1) * R + r]));
        break;
      }
      case PRISM_OP_NEG_I: {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int64_t a = static_cast<int64_t>(st[(sp - 1) * R + r]), o;
          if (__builtin_sub_overflow(static_cast<int64_t>(0), a, &o) && ((vget(sp - 1) >> r) & 1)) atomicOr(err, 2u);
          st[(sp - 1) * R + r] = static_cast<uint64_t>(o);
        }
        break;
      }
      case PRISM_OP_ADD_F:
      case PRISM_OP_SUB_F:
      case PRISM_OP_MUL_F:
      case PRISM_OP_DIV_F: {
        --sp;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          double a = as_f(st[(sp - 1) * R + r]), b = as_f(st[sp * R + r]), o;
          o = op == PRISM_OP_ADD_F ? a + b : op == PRISM_OP_SUB_F ? a - b : op == PRISM_OP_MUL_F ? a * b : a / b;
          st[(sp - 1) * R + r] = f_bits(o);
        }
        vset(sp - 1, vget(sp - 1) & vget(sp));
        break;
      }
      case PRISM_OP_ADD_I:
      case PRISM_OP_SUB_I:
      case PRISM_OP_MUL_I: {
        --sp;
        const uint32_t both = vget(sp - 1) & vget(sp);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int64_t a = static_cast<int64_t>(st[(sp - 1) * R + r]), b = static_cast<int64_t>(st[sp * R + r]), o;
          bool ov = op == PRISM_OP_ADD_I ? __builtin_add_overflow(a, b, &o)
                    : op == PRISM_OP_SUB_I ? __builtin_sub_overflow(a, b, &o)
                                           : __builtin_mul_overflow(a, b, &o);
          if (ov && ((both >> r) & 1)) atomicOr(err, 2u);
          st[(sp - 1) * R + r] = static_cast<uint64_t>(o);
        }
        vset(sp - 1, both);
        break;
      }
      case PRISM_OP_CMP_I: {
        --sp;
#pragma unroll
        for (int r = 0; r < R; ++r)
          st[(sp - 1) * R + r] = cmp_i(arg, static_cast<int64_t>(st[(sp - 1) * R + r]), static_cast<int64_t>(st[sp * R + r]));
        vset(sp - 1, vget(sp - 1) & vget(sp));
        break;
      }
      case PRISM_OP_CMP_F: {
        --sp;
#pragma unroll
        for (int r = 0; r < R; ++r) st[(sp - 1) * R + r] = cmp_f(arg, as_f(st[(sp - 1) * R + r]), as_f(st[sp * R + r]));
        vset(sp - 1, vget(sp - 1) & vget(sp));
        break;
      }
      case PRISM_OP_CMP_STR:
      case PRISM_OP_LIKE:
      case PRISM_OP_ILIKE: {
        const prism_column& c = cols[arg];
        const int sidx = static_cast<int>(code[pc].i64 & 0xff);
        const int cop = static_cast<int>(code[pc].i64 >> 8);
        const uint8_t* lit = strs->bytes + strs->off[sidx];
        const int32_t nlit = strs->off[sidx + 1] - strs->off[sidx];
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          bool res = false;
          if (in[r]) {
            const uint64_t i = c.offset + row[r];
            const int32_t b = c.offsets[i], e = c.offsets[i + 1];
            const uint8_t* s = static_cast<const uint8_t*>(c.data) + b;
            if (op == PRISM_OP_CMP_STR) {
              int k = str_cmp(s, e - b, lit, nlit);
              res = cmp_i(cop, k, 0);
            } else {
              const int lk = strs->like_kind[sidx];
              res = lk == LK_GENERAL ? like_match(s, e - b, lit, nlit, op == PRISM_OP_ILIKE)
                                     : like_fast(lk, s, e - b, lit, nlit, op == PRISM_OP_ILIKE);
            }
            v |= col_valid(c, row[r]) ? (1u << r) : 0u;
          }
          st[sp * R + r] = res;
        }
        vset(sp++, v);
        break;
      }
      case PRISM_OP_IS_NULL:
      case PRISM_OP_IS_NOT_NULL: {
        const prism_column& c = cols[arg];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          bool nul = in[r] && !col_valid(c, row[r]);
          st[sp * R + r] = op == PRISM_OP_IS_NULL ? nul : !nul;
        }
        vset(sp++, kAll);
        break;
      }
      case PRISM_OP_AND:
      case PRISM_OP_OR: {
        --sp;
#pragma unroll
        for (int r = 0; r < R; ++r)
          st[(sp - 1) * R + r] = op == PRISM_OP_AND ? (st[(sp - 1) * R + r] & st[sp * R + r]) : (st[(sp - 1) * R + r] | st[sp * R + r]);
        vset(sp - 1, vget(sp - 1) & vget(sp));
        break;
      }
      case PRISM_OP_NOT: {
#pragma unroll
        for (int r = 0; r < R; ++r) st[(sp - 1) * R + r] = st[(sp - 1) * R + r] ^ 1;
        break;
      }
      case PRISM_OP_ZIP: {
        // stack: else, cond, then  ->  cond ? then : else (a null cond picks else)
        sp -= 2;
        const uint32_t cv = vget(sp);
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const bool take = ((cv >> r) & 1) && (st[sp * R + r] & 1);
          st[(sp - 1) * R + r] = take ? st[(sp + 1) * R + r] : st[(sp - 1) * R + r];
          v |= ((take ? (vget(sp + 1) >> r) : (vget(sp - 1) >> r)) & 1) ? (1u << r) : 0u;
        }
        vset(sp - 1, v);
        break;
      }
      case PRISM_OP_SELECT: {
        // stack: cond, then, else  ->  zip(cond, then, else): a null cond picks else
        sp -= 2;
        const uint32_t cv = vget(sp - 1);
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const bool take = ((cv >> r) & 1) && (st[(sp - 1) * R + r] & 1);
          st[(sp - 1) * R + r] = take ? st[sp * R + r] : st[(sp + 1) * R + r];
          v |= ((take ? (vget(sp) >> r) : (vget(sp + 1) >> r)) & 1) ? (1u << r) : 0u;
        }
        vset(sp - 1, v);
        break;
      }
      default:
        break;
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) out[r] = st[r];
  out_valid = vget(0);
}

#endif  // __HIPCC__

}  // namespace prism
