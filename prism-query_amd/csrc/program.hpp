// Row programs: the fused predicate / projection stage evaluated inside the
// aggregation and filter kernels.  Predicate (filter_project.rs:22-51) and
// ScalarExpr (filter_project.rs:195-213) trees are compiled on the host into
// postfix code; every lane runs the same instruction at the same time, so the
// dispatch is scalar and the value stack lives in VGPRs (uniform stack index).
#pragma once

#include "common.hpp"
#include "dev_program.hpp"

namespace prism {

// Host-side checks: validates stack discipline and operand types, returns the
// result type.  Throws Error(InvalidArgument) on a malformed program.
ValType check_program(const prism_instr* code, uint32_t len, const int32_t* col_types, uint32_t n_cols,
                      uint32_t n_strings, int* max_depth);

// Build the device string table.
void build_str_table(const prism_string* s, uint32_t n, StrTable& out);

#if defined(__HIPCC__)

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
            const int32_t b = gload<int32_t>(c.offsets, i), e = gload<int32_t>(c.offsets, i + 1);
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
