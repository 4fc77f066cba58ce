// Filter / project over Arrow RecordBatches on gfx950
// (native/prism-executor/src/filter_project.rs:20-404, string_ops.rs:16-25).
//
// Predicate and ScalarExpr trees are compiled on the host into the row-program
// bytecode of program.hpp and evaluated by one kernel per tree: every wave
// takes 64 consecutive rows per step, so a ballot turns a boolean result into
// one 64-bit word of an Arrow bitmap.  Filtering compacts the selected rows
// with a selection vector (word popcounts, exclusive scan, in-word ranks) and
// gathers only the columns that are kept.
//
// Semantics kept from the reference:
//   comparisons    - only the column/scalar type pairs compare_column accepts
//                    (filter_project.rs:346-390); floats in IEEE total order
//                    as arrow-ord compares them (NaN == NaN, -0.0 < +0.0);
//                    null rows give null.
//   AND / OR / NOT - arrow `and` / `or` / `not`: null if any input is null.
//   LIKE / ILIKE   - '%' and '_' (one UTF-8 character), '\' escapes;
//                    ILIKE folds ASCII letters only.
//   filter         - a null mask entry drops the row.
//   arithmetic     - both operands cast to Float64 (filter_project.rs:236-247).
//   IfThen         - result type = first THEN branch's type; later clauses
//                    are applied first, so the earliest true clause wins
//                    (filter_project.rs:256-292).
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "arrow_io.hpp"
#include "gather.hpp"
#include "like_scan.hpp"
#include "program.hpp"

namespace prism {
namespace fp {

struct DevProgram {
  int32_t len;
  int32_t pad;
  prism_instr code[kMaxCode];
  StrTable strs;
};

enum OutKind : int32_t { OUT_BITS = 0, OUT_W8 = 1, OUT_W4 = 2 };

// A program that is one LIKE '%literal%' of a Utf8 column (col >= 0, from
// like_scan_prepare): a bitmap of it runs as the byte scan.
struct LikeScan {
  int col = -1;
  LikeArg arg;
};

constexpr int kEB = 256;  // threads per workgroup

// One wave = 64 consecutive rows per word; R words per lane per step.
template <int R, int DEPTH>
__global__ __launch_bounds__(kEB) void k_eval(const DevProgram* __restrict__ P, const prism_column* __restrict__ cols,
                                              uint64_t n, int out_kind, void* __restrict__ out,
                                              unsigned long long* __restrict__ out_valid,
                                              unsigned long long* __restrict__ null_count, uint32_t* err) {
  const int lane = threadIdx.x & 63;
  const uint64_t n_words = ceil_div(n, 64);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(kEB) + threadIdx.x) >> 6;
  const uint64_t n_waves = (static_cast<uint64_t>(gridDim.x) * kEB) >> 6;
  const uint64_t pre[1][R] = {};
  unsigned long long nulls = 0, selected = 0;
  for (uint64_t w0 = wave * R; w0 < n_words; w0 += n_waves * R) {
    uint64_t row[R];
    bool in[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      row[r] = (w0 + r) * 64 + lane;
      in[r] = row[r] < n;
    }
    uint64_t val[R];
    uint32_t valid;
    run_program<R, DEPTH, 1>(P->code, P->len, cols, &P->strs, row, in, pre, 0u, val, valid, err);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (w0 + r >= n_words) break;
      const bool v = in[r] && ((valid >> r) & 1);
      const unsigned long long vw = __ballot(v);
      const unsigned long long inw = __ballot(in[r]);
      nulls += __popcll(inw & ~vw);
      if (out_kind == OUT_BITS) {
        const unsigned long long bw = __ballot(v && (val[r] & 1));
        selected += __popcll(bw);
        if (lane == 0) static_cast<unsigned long long*>(out)[w0 + r] = bw;
      } else if (in[r]) {
        if (out_kind == OUT_W8) static_cast<uint64_t*>(out)[row[r]] = v ? val[r] : 0;
        else static_cast<uint32_t*>(out)[row[r]] = v ? static_cast<uint32_t>(val[r]) : 0;
      }
      if (lane == 0 && out_valid) out_valid[w0 + r] = vw;
    }
  }
  // one atomic per wave (wave-uniform value)
  if (lane == 0 && nulls) atomicAdd(null_count, nulls);
  if (lane == 0 && selected) atomicAdd(null_count + 2, selected);
}

// A program that is one LIKE '%literal%' of a Utf8 column: the byte scan of
// like_scan.hpp over sub-tiles of kLsRows rows.
__global__ __launch_bounds__(kLsThreads) void k_like_scan(const LikeArg* __restrict__ A,
                                                          const prism_column* __restrict__ cols, int col, uint64_t n,
                                                          unsigned long long* __restrict__ out,
                                                          unsigned long long* __restrict__ out_valid,
                                                          unsigned long long* __restrict__ counters) {
  __shared__ uint32_t s_bits[kLsBitWords];
  const prism_column c = cols[col];
  unsigned long long nulls = 0, selected = 0;
  for (uint64_t r0 = blockIdx.x * static_cast<uint64_t>(kLsRows); r0 < n;
       r0 += static_cast<uint64_t>(gridDim.x) * kLsRows)
    like_scan_tile(c, r0, r0 + kLsRows < n ? r0 + kLsRows : n, *A, s_bits, out, out_valid, nulls, selected);
  if ((threadIdx.x & 63) == 0) {
    if (nulls) atomicAdd(counters, nulls);
    if (selected) atomicAdd(counters + 2, selected);
  }
}

// Mask words from an Arrow BooleanArray (values AND validity), bit offset < 8.
__global__ void k_mask_words(const uint8_t* __restrict__ vals, const uint8_t* __restrict__ valid, int64_t off,
                             uint64_t n, unsigned long long* __restrict__ words) {
  const int lane = threadIdx.x & 63;
  const uint64_t n_words = ceil_div(n, 64);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  const uint64_t n_waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t w = wave; w < n_words; w += n_waves) {
    const uint64_t i = w * 64 + lane;
    bool b = false;
    if (i < n) {
      const uint64_t j = off + i;
      b = ((vals[j >> 3] >> (j & 7)) & 1) && (valid == nullptr || ((valid[j >> 3] >> (j & 7)) & 1));
    }
    const unsigned long long m = __ballot(b);
    if (lane == 0) words[w] = m;
  }
}

// Multi-source Utf8 select for IfThen over string branches: idx[r] picks a
// source (column or literal); a null index or a null source value is null.
struct StrSrc {
  int32_t is_lit;
  int32_t lit_off, lit_len;
  int32_t pad;
  prism_column col;
};
constexpr int kMaxSrc = 16;
struct StrSrcs {
  int32_t n;
  int32_t pad;
  StrSrc s[kMaxSrc];
  uint8_t lit[kMaxStrBytes];
};

__device__ __forceinline__ bool src_pick(const StrSrcs* S, const uint32_t* idx, const unsigned long long* ivalid,
                                         uint64_t r, const uint8_t*& p, int32_t& len) {
  if (!((ivalid[r >> 6] >> (r & 63)) & 1)) return false;
  const StrSrc& s = S->s[idx[r]];
  if (s.is_lit) {
    p = S->lit + s.lit_off;
    len = s.lit_len;
    return true;
  }
  if (!col_valid(s.col, r)) return false;
  const uint64_t i = s.col.offset + r;
  const int32_t b = gload<int32_t>(s.col.offsets, i), e = gload<int32_t>(s.col.offsets, i + 1);
  p = static_cast<const uint8_t*>(s.col.data) + b;
  len = e - b;
  return true;
}

__global__ void k_str_select_len(const StrSrcs* __restrict__ S, const uint32_t* __restrict__ idx,
                                 const unsigned long long* __restrict__ ivalid, uint64_t n, uint64_t* __restrict__ len,
                                 unsigned long long* __restrict__ out_valid, unsigned long long* __restrict__ null_count) {
  const int lane = threadIdx.x & 63;
  const uint64_t n_words = ceil_div(n, 64);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  const uint64_t n_waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  unsigned long long nulls = 0;
  for (uint64_t w = wave; w < n_words; w += n_waves) {
    const uint64_t r = w * 64 + lane;
    bool ok = false;
    int32_t l = 0;
    if (r < n) {
      const uint8_t* p;
      ok = src_pick(S, idx, ivalid, r, p, l);
      len[r] = ok ? static_cast<uint64_t>(l) : 0;
    }
    const unsigned long long vw = __ballot(ok), inw = __ballot(r < n);
    nulls += __popcll(inw & ~vw);
    if (lane == 0) out_valid[w] = vw;
  }
  if (lane == 0 && nulls) atomicAdd(null_count, nulls);
}

__global__ void k_str_select_copy(const StrSrcs* __restrict__ S, const uint32_t* __restrict__ idx,
                                  const unsigned long long* __restrict__ ivalid, uint64_t n,
                                  const uint64_t* __restrict__ start, const uint64_t* __restrict__ len,
                                  int32_t* __restrict__ out_off, uint8_t* __restrict__ out_bytes) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t s = start[r];
    out_off[r] = static_cast<int32_t>(s);
    if (r == n - 1) out_off[n] = static_cast<int32_t>(s + len[r]);
    const uint8_t* p;
    int32_t l;
    if (!src_pick(S, idx, ivalid, r, p, l)) continue;
    for (int32_t k = 0; k < l; ++k) out_bytes[s + k] = p[k];
  }
}

__global__ void k_zero_u64(unsigned long long* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

static dim3 egrid(prism_ctx* ctx, uint64_t n_words, int words_per_wave) {
  const uint64_t waves = ceil_div(n_words ? n_words : 1, static_cast<uint64_t>(words_per_wave));
  uint64_t b = ceil_div(waves, static_cast<uint64_t>(kEB / 64));
  const uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 8;
  return dim3(static_cast<unsigned>(b < cap ? b : cap));
}

// ---------------------------------------------------------------------------
// Host: tree -> row program.
// ---------------------------------------------------------------------------
static std::string type_debug(DType t) {
  switch (t) {
    case DType::Int32: return "Int32";
    case DType::Int64: return "Int64";
    case DType::Float64: return "Float64";
    case DType::Utf8: return "Utf8";
    case DType::Bool: return "Boolean";
    case DType::Date32: return "Date32";
    default: return dtype_name(t);
  }
}

static std::string scalar_debug(const prism_scalar& v) {
  char buf[64];
  switch (v.type) {
    case PRISM_SCALAR_INT32: return "Int32(" + std::to_string(v.i64) + ")";
    case PRISM_SCALAR_INT64: return "Int64(" + std::to_string(v.i64) + ")";
    case PRISM_SCALAR_FLOAT64: snprintf(buf, sizeof(buf), "Float64(%g)", v.f64); return buf;
    case PRISM_SCALAR_UTF8: return "Utf8(\"" + std::string(v.str ? v.str : "", v.str ? v.str_len : 0) + "\")";
    case PRISM_SCALAR_BOOLEAN: return std::string("Boolean(") + (v.i64 ? "true" : "false") + ")";
    case PRISM_SCALAR_DATE32: return "Date32(" + std::to_string(v.i64) + ")";
    default: return "?";
  }
}

static DType scalar_dtype(const prism_scalar& v) {
  switch (v.type) {
    case PRISM_SCALAR_INT32: return DType::Int32;
    case PRISM_SCALAR_INT64: return DType::Int64;
    case PRISM_SCALAR_FLOAT64: return DType::Float64;
    case PRISM_SCALAR_UTF8: return DType::Utf8;
    case PRISM_SCALAR_BOOLEAN: return DType::Bool;
    case PRISM_SCALAR_DATE32: return DType::Date32;
    default: fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown scalar type " + std::to_string(v.type));
  }
}

static bool is_int_like(DType t) { return t == DType::Int32 || t == DType::Int64 || t == DType::Date32; }

struct Compiler {
  const std::vector<HostColumn>& cols;
  const prism_pred_node* preds;
  uint32_t n_preds;
  const prism_expr_node* exprs;
  uint32_t n_exprs;
  std::vector<prism_instr> code;
  std::vector<std::string> strings;

  Compiler(const std::vector<HostColumn>& c, const prism_pred_node* p, uint32_t np, const prism_expr_node* e,
           uint32_t ne)
      : cols(c), preds(p), n_preds(np), exprs(e), n_exprs(ne) {}

  void emit(int32_t op, int32_t arg = 0, int64_t i64 = 0, double f64 = 0.0) {
    prism_instr x;
    x.op = op;
    x.arg = arg;
    x.i64 = i64;
    x.f64 = f64;
    code.push_back(x);
  }
  int64_t add_string(const char* p, uint32_t len) {
    if (strings.size() >= static_cast<size_t>(kMaxStrings))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: more than 16 string literals in one expression");
    strings.emplace_back(p ? p : "", p ? len : 0);
    return static_cast<int64_t>(strings.size() - 1);
  }
  const HostColumn& column(uint32_t c) {
    if (c >= cols.size())
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: column index " + std::to_string(c) + " out of range");
    return cols[c];
  }
  const prism_pred_node& pnode(int32_t i) {
    if (i < 0 || static_cast<uint32_t>(i) >= n_preds)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: predicate node index out of range");
    return preds[i];
  }
  const prism_expr_node& enode(int32_t i) {
    if (i < 0 || static_cast<uint32_t>(i) >= n_exprs)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: expression node index out of range");
    return exprs[i];
  }

  // Predicate -> boolean value on the stack (evaluate_predicate, filter_project.rs:65-123).
  void pred(int32_t i, int level) {
    if (level > 64) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: predicate nested too deeply");
    const prism_pred_node& p = pnode(i);
    switch (p.kind) {
      case PRISM_PRED_LITERAL:
        emit(PRISM_OP_CONST_BOOL, 0, p.value.i64 != 0);
        return;
      case PRISM_PRED_EQ: case PRISM_PRED_NE: case PRISM_PRED_LT:
      case PRISM_PRED_LE: case PRISM_PRED_GT: case PRISM_PRED_GE: {
        static const int32_t cmp[] = {0, PRISM_CMP_EQ, PRISM_CMP_NE, PRISM_CMP_LT, PRISM_CMP_LE, PRISM_CMP_GT, PRISM_CMP_GE};
        const int32_t op = cmp[p.kind];
        const HostColumn& c = column(p.column);
        const int st = p.value.type;
        const int32_t col = static_cast<int32_t>(p.column);
        if ((c.type == DType::Int32 && st == PRISM_SCALAR_INT32) || (c.type == DType::Int64 && st == PRISM_SCALAR_INT64) ||
            (c.type == DType::Date32 && (st == PRISM_SCALAR_DATE32 || st == PRISM_SCALAR_INT32))) {
          emit(PRISM_OP_LOAD_I, col);
          emit(PRISM_OP_CONST_I, 0, p.value.i64);
          emit(PRISM_OP_CMP_I, op);
        } else if (c.type == DType::Float64 && st == PRISM_SCALAR_FLOAT64) {
          emit(PRISM_OP_LOAD_F, col);
          emit(PRISM_OP_CONST_F, 0, 0, p.value.f64);
          emit(PRISM_OP_CMP_F, op);
        } else if (c.type == DType::Utf8 && st == PRISM_SCALAR_UTF8) {
          const int64_t s = add_string(p.value.str, p.value.str_len);
          emit(PRISM_OP_CMP_STR, col, s | (static_cast<int64_t>(op) << 8));
        } else {
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: Cannot compare column type " + type_debug(c.type) +
                                               " with scalar " + scalar_debug(p.value));
        }
        return;
      }
      case PRISM_PRED_IS_NULL:
      case PRISM_PRED_IS_NOT_NULL:
        column(p.column);
        emit(p.kind == PRISM_PRED_IS_NULL ? PRISM_OP_IS_NULL : PRISM_OP_IS_NOT_NULL, static_cast<int32_t>(p.column));
        return;
      case PRISM_PRED_AND:
      case PRISM_PRED_OR:
        pred(p.left, level + 1);
        pred(p.right, level + 1);
        emit(p.kind == PRISM_PRED_AND ? PRISM_OP_AND : PRISM_OP_OR);
        return;
      case PRISM_PRED_NOT:
        pred(p.left, level + 1);
        emit(PRISM_OP_NOT);
        return;
      case PRISM_PRED_LIKE:
      case PRISM_PRED_ILIKE: {
        const HostColumn& c = column(p.column);
        const char* what = p.kind == PRISM_PRED_LIKE ? "LIKE" : "ILIKE";
        if (c.type != DType::Utf8)
          fail(PRISM_ERR_INVALID_ARGUMENT,
               std::string("Invalid argument: ") + what + " requires Utf8 column, got " + type_debug(c.type));
        const int64_t s = add_string(p.pattern, p.pattern_len);
        emit(p.kind == PRISM_PRED_LIKE ? PRISM_OP_LIKE : PRISM_OP_ILIKE, static_cast<int32_t>(p.column), s);
        return;
      }
      default:
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown predicate kind " + std::to_string(p.kind));
    }
  }

  // Arrow type an expression evaluates to.
  DType type_of(int32_t i, int level = 0) {
    if (level > 64) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: expression nested too deeply");
    const prism_expr_node& e = enode(i);
    switch (e.kind) {
      case PRISM_EXPR_COLUMN: return column(e.column).type;
      case PRISM_EXPR_LITERAL: return scalar_dtype(e.value);
      case PRISM_EXPR_BINARY:
      case PRISM_EXPR_NEGATE: return DType::Float64;
      case PRISM_EXPR_IF_THEN:
        if (e.n_clauses == 0) {
          if (e.else_expr < 0)
            fail(PRISM_ERR_INVALID_ARGUMENT,
                 "Invalid argument: IfThen expression requires at least one clause or an else branch");
          return type_of(e.else_expr, level + 1);
        }
        return type_of(e.clause_exprs[0], level + 1);
      default:
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown expression kind " + std::to_string(e.kind));
    }
  }

  // Cast of the value on the stack from `from` to `to` (arrow cast, the cases
  // the device evaluates).
  void cast(DType from, DType to) {
    if (from == to) return;
    auto unsupported = [&] {
      fail(PRISM_ERR_ARROW, "Arrow error: Cast error: Casting from " + type_debug(from) + " to " + type_debug(to) +
                                " not supported by the device evaluator");
    };
    switch (to) {
      case DType::Float64:
        if (from == DType::Int32 || from == DType::Int64 || from == DType::Bool) emit(PRISM_OP_I2F);
        else unsupported();
        return;
      case DType::Int64:
        if (from == DType::Int32 || from == DType::Bool || from == DType::Date32) return;  // int64 words already
        unsupported();
        return;
      case DType::Int32:
        if (from == DType::Bool || from == DType::Date32) return;
        unsupported();
        return;
      case DType::Date32:
        if (from == DType::Int32) return;
        unsupported();
        return;
      case DType::Bool:
        if (from == DType::Int32 || from == DType::Int64) {
          emit(PRISM_OP_CONST_I, 0, 0);
          emit(PRISM_OP_CMP_I, PRISM_CMP_NE);
          return;
        }
        unsupported();
        return;
      default:
        unsupported();
    }
  }

  // Expression value of type `to` on the stack.
  void value(int32_t i, DType to, int level) {
    if (level > 64) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: expression nested too deeply");
    const prism_expr_node& e = enode(i);
    switch (e.kind) {
      case PRISM_EXPR_COLUMN: {
        const HostColumn& c = column(e.column);
        if (c.type == DType::Utf8 || to == DType::Utf8)
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: Utf8 value inside a numeric expression");
        if (c.type == DType::Float64) emit(PRISM_OP_LOAD_F, static_cast<int32_t>(e.column));
        else emit(PRISM_OP_LOAD_I, static_cast<int32_t>(e.column));
        cast(c.type, to);
        return;
      }
      case PRISM_EXPR_LITERAL: {
        const DType t = scalar_dtype(e.value);
        if (t == DType::Utf8 || to == DType::Utf8)
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: Utf8 literal inside a numeric expression");
        if (t == DType::Float64) emit(PRISM_OP_CONST_F, 0, 0, e.value.f64);
        else if (t == DType::Bool) emit(PRISM_OP_CONST_BOOL, 0, e.value.i64 != 0);
        else emit(PRISM_OP_CONST_I, 0, e.value.i64);
        cast(t, to);
        return;
      }
      case PRISM_EXPR_BINARY: {
        static const int32_t ops[] = {PRISM_OP_ADD_F, PRISM_OP_SUB_F, PRISM_OP_MUL_F, PRISM_OP_DIV_F};
        if (e.op < 0 || e.op > 3) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown arithmetic op");
        value(e.left, DType::Float64, level + 1);
        value(e.right, DType::Float64, level + 1);
        emit(ops[e.op]);
        cast(DType::Float64, to);
        return;
      }
      case PRISM_EXPR_NEGATE:
        value(e.left, DType::Float64, level + 1);
        emit(PRISM_OP_NEG_F);
        cast(DType::Float64, to);
        return;
      case PRISM_EXPR_IF_THEN: {
        const DType t = type_of(i, level);
        if (t == DType::Utf8) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: Utf8 IfThen inside an expression");
        if_then(e, t, level);
        cast(t, to);
        return;
      }
      default:
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown expression kind " + std::to_string(e.kind));
    }
  }

  // IfThen of non-string type t: result = else; clauses applied from the last
  // to the first with ZIP (else, cond, then).
  void if_then(const prism_expr_node& e, DType t, int level) {
    if (e.n_clauses == 0) {
      value(e.else_expr, t, level + 1);
      return;
    }
    if (!e.clause_preds || !e.clause_exprs) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL IfThen clauses");
    if (e.else_expr >= 0) value(e.else_expr, t, level + 1);
    else emit(PRISM_OP_CONST_NULL);
    for (int64_t k = static_cast<int64_t>(e.n_clauses) - 1; k >= 0; --k) {
      pred(e.clause_preds[k], level + 1);
      value(e.clause_exprs[k], t, level + 1);
      emit(PRISM_OP_ZIP);
    }
  }

  // Check the program and upload it.
  ValType finish(prism_ctx* ctx, DeviceBuffer& dev, int* depth, LikeScan* like = nullptr) {
    std::vector<int32_t> types(cols.size());
    for (size_t i = 0; i < cols.size(); ++i)
      types[i] = cols[i].type == DType::UInt32 || cols[i].type == DType::Decimal128 ? 0 : static_cast<int32_t>(cols[i].type);
    if (code.size() > static_cast<size_t>(kMaxCode))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: expression needs more than 160 instructions");
    const ValType vt = check_program(code.data(), static_cast<uint32_t>(code.size()), types.data(),
                                     static_cast<uint32_t>(types.size()), static_cast<uint32_t>(strings.size()), depth);
    auto P = std::make_unique<DevProgram>();
    memset(P.get(), 0, sizeof(DevProgram));
    P->len = static_cast<int32_t>(code.size());
    memcpy(P->code, code.data(), code.size() * sizeof(prism_instr));
    std::vector<prism_string> ps(strings.size());
    for (size_t i = 0; i < strings.size(); ++i) {
      ps[i].data = strings[i].data();
      ps[i].len = static_cast<uint32_t>(strings[i].size());
    }
    build_str_table(ps.data(), static_cast<uint32_t>(ps.size()), P->strs);
    if (like) like->col = like_scan_prepare(code.data(), static_cast<uint32_t>(code.size()), P->strs, like->arg);
    dev.alloc(as_ctx(ctx), sizeof(DevProgram));
    PRISM_HIP(hipMemcpyAsync(dev.ptr, P.get(), sizeof(DevProgram), hipMemcpyHostToDevice, ctx->stream));
    PRISM_HIP(hipStreamSynchronize(ctx->stream));  // P is a host temporary
    return vt;
  }
};

// Device copies of the columns a program reads, plus their descriptors.
struct DevBatch {
  std::vector<DevColumn> cols;
  std::vector<bool> present;
  DeviceBuffer desc;
};

static void collect_cols(const std::vector<prism_instr>& code, std::vector<bool>& used) {
  for (const prism_instr& x : code)
    switch (x.op) {
      case PRISM_OP_LOAD_I: case PRISM_OP_LOAD_F: case PRISM_OP_CMP_STR: case PRISM_OP_LIKE:
      case PRISM_OP_ILIKE: case PRISM_OP_IS_NULL: case PRISM_OP_IS_NOT_NULL:
        used[x.arg] = true;
        break;
      default:
        break;
    }
}

static void upload_cols(prism_ctx* ctx, const HostBatch& hb, const std::vector<bool>& used, DevBatch& db) {
  db.cols.resize(hb.cols.size());
  db.present.assign(hb.cols.size(), false);
  std::vector<prism_column> d(hb.cols.size() ? hb.cols.size() : 1);
  memset(d.data(), 0, d.size() * sizeof(prism_column));
  for (size_t i = 0; i < hb.cols.size(); ++i) {
    if (!used[i]) continue;
    db.cols[i] = upload(ctx, hb.cols[i]);
    db.present[i] = true;
    d[i] = db.cols[i].desc();
  }
  db.desc.alloc(as_ctx(ctx), d.size() * sizeof(prism_column));
  PRISM_HIP(hipMemcpyAsync(db.desc.ptr, d.data(), d.size() * sizeof(prism_column), hipMemcpyHostToDevice, ctx->stream));
  PRISM_HIP(hipStreamSynchronize(ctx->stream));
}

struct EvalOut {
  DeviceBuffer values, valid, counters;  // counters: null count, error flags, selected rows (bit output)
  uint64_t nulls = 0;
  uint64_t selected = 0;
};

static void run_eval(prism_ctx* ctx, const DevProgram* P, int depth, const prism_column* cols, uint64_t n,
                     OutKind kind, EvalOut& o, void* out_values = nullptr, const LikeScan* like = nullptr) {
  Ctx* c = as_ctx(ctx);
  hipStream_t st = ctx->stream;
  const uint64_t nw = ceil_div(n, 64);
  if (!out_values) o.values.alloc(c, kind == OUT_BITS ? (nw ? nw : 1) * 8 : (n ? n : 1) * (kind == OUT_W8 ? 8 : 4));
  void* vals = out_values ? out_values : o.values.ptr;
  o.valid.alloc(c, (nw ? nw : 1) * 8);
  o.counters.alloc(c, 32);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, o.counters.as<unsigned long long>(), 3);
  DeviceBuffer la;
  if (n && like && like->col >= 0 && kind == OUT_BITS) {
    la.alloc(c, sizeof(LikeArg));
    PRISM_HIP(hipMemcpyAsync(la.ptr, &like->arg, sizeof(LikeArg), hipMemcpyHostToDevice, st));
    const uint64_t tiles = ceil_div(n, static_cast<uint64_t>(kLsRows));
    const uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 8;
    KernelTimer kt(ctx, "filter_like_scan");
    hipLaunchKernelGGL(k_like_scan, dim3(static_cast<unsigned>(tiles < cap ? tiles : cap)), dim3(kLsThreads), 0, st,
                       la.as<LikeArg>(), cols, like->col, n, static_cast<unsigned long long*>(vals),
                       o.valid.as<unsigned long long>(), o.counters.as<unsigned long long>());
  } else if (n) {
    KernelTimer kt(ctx, "filter_eval");
    if (depth <= 4)
      hipLaunchKernelGGL((k_eval<4, 4>), egrid(ctx, nw, 4), dim3(kEB), 0, st, P, cols, n, static_cast<int>(kind),
                         vals, o.valid.as<unsigned long long>(), o.counters.as<unsigned long long>(),
                         reinterpret_cast<uint32_t*>(o.counters.as<char>() + 8));
    else
      hipLaunchKernelGGL((k_eval<2, kStackDepth>), egrid(ctx, nw, 2), dim3(kEB), 0, st, P, cols, n,
                         static_cast<int>(kind), vals, o.valid.as<unsigned long long>(),
                         o.counters.as<unsigned long long>(), reinterpret_cast<uint32_t*>(o.counters.as<char>() + 8));
  }
  PRISM_CHECK_LAUNCH();
  unsigned long long h[3];
  PRISM_HIP(hipMemcpyAsync(h, o.counters.ptr, 24, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipStreamSynchronize(st));
  o.nulls = h[0];
  o.selected = h[2];
}

// Mask words (value AND valid) of a predicate over the batch.
static void predicate_words(prism_ctx* ctx, const HostBatch& hb, const prism_pred_node* nodes, uint32_t n_nodes,
                            EvalOut& o, DevBatch& db) {
  if (!nodes || n_nodes == 0) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: empty predicate");
  Compiler C(hb.cols, nodes, n_nodes, nullptr, 0);
  C.pred(0, 0);
  std::vector<bool> used(hb.cols.size(), false);
  collect_cols(C.code, used);
  upload_cols(ctx, hb, used, db);
  DeviceBuffer prog;
  int depth = 1;
  LikeScan like;
  C.finish(ctx, prog, &depth, &like);
  run_eval(ctx, prog.as<DevProgram>(), depth, db.desc.as<prism_column>(), static_cast<uint64_t>(hb.n_rows), OUT_BITS, o,
           nullptr, &like);
}

static OutColumn bits_column(prism_ctx* ctx, const EvalOut& o, uint64_t n, const std::string& name, DType type) {
  OutColumn c;
  c.type = type;
  c.name = name;
  c.nullable = true;
  c.length = static_cast<int64_t>(n);
  c.null_count = static_cast<int64_t>(o.nulls);
  const size_t nb = (n + 7) / 8;
  hipStream_t st = ctx->stream;
  if (type == DType::Bool) {
    c.data.resize(nb);
    if (nb) PRISM_HIP(hipMemcpyAsync(c.data.data(), o.values.ptr, nb, hipMemcpyDeviceToHost, st));
  } else {
    const size_t w = static_cast<size_t>(dtype_width(type));
    c.data.resize(n * w);
    if (n) PRISM_HIP(hipMemcpyAsync(c.data.data(), o.values.ptr, n * w, hipMemcpyDeviceToHost, st));
  }
  if (o.nulls && nb) {
    c.validity.resize(nb);
    PRISM_HIP(hipMemcpyAsync(c.validity.data(), o.valid.ptr, nb, hipMemcpyDeviceToHost, st));
  }
  PRISM_HIP(hipStreamSynchronize(st));
  return c;
}

// Host copy of an input column with offset 0 (projection passes columns
// through unchanged; there is nothing to compute).
static OutColumn host_copy(const HostColumn& h, const std::string& name) {
  OutColumn o;
  o.type = h.type;
  o.name = name;
  o.nullable = h.nullable;
  o.length = h.length;
  const int64_t n = h.length, off = h.offset;
  auto copy_bits = [&](const uint8_t* src, std::vector<uint8_t>& dst) {
    dst.assign(static_cast<size_t>((n + 7) / 8), 0);
    for (int64_t i = 0; i < n; ++i)
      if ((src[(off + i) >> 3] >> ((off + i) & 7)) & 1) dst[i >> 3] |= static_cast<uint8_t>(1u << (i & 7));
  };
  if (h.type == DType::Utf8) {
    o.offsets.resize(static_cast<size_t>(n + 1));
    const int32_t base = h.offsets ? h.offsets[off] : 0;
    for (int64_t i = 0; i <= n; ++i) o.offsets[i] = h.offsets ? h.offsets[off + i] - base : 0;
    o.data.assign(static_cast<const uint8_t*>(h.data) + base, static_cast<const uint8_t*>(h.data) + base + o.offsets[n]);
  } else if (h.type == DType::Bool) {
    copy_bits(static_cast<const uint8_t*>(h.data), o.data);
  } else {
    const size_t w = static_cast<size_t>(dtype_width(h.type));
    const uint8_t* p = static_cast<const uint8_t*>(h.data) + off * w;
    o.data.assign(p, p + n * w);
  }
  if (h.validity && h.null_count != 0) {
    copy_bits(h.validity, o.validity);
    int64_t valid = 0;
    for (int64_t i = 0; i < n; ++i) valid += (o.validity[i >> 3] >> (i & 7)) & 1;
    o.null_count = n - valid;
    if (o.null_count == 0) o.validity.clear();
  }
  return o;
}

// Broadcast of a literal (ScalarExpr::Literal, filter_project.rs:217-231).
static OutColumn literal_column(const prism_scalar& v, uint64_t n, const std::string& name) {
  OutColumn o;
  o.type = scalar_dtype(v);
  o.name = name;
  o.nullable = true;
  o.length = static_cast<int64_t>(n);
  switch (o.type) {
    case DType::Utf8: {
      const uint32_t l = v.str ? v.str_len : 0;
      o.offsets.resize(n + 1);
      for (uint64_t i = 0; i <= n; ++i) o.offsets[i] = static_cast<int32_t>(i * l);
      o.data.resize(n * l);
      for (uint64_t i = 0; i < n; ++i) memcpy(o.data.data() + i * l, v.str, l);
      break;
    }
    case DType::Bool:
      o.data.assign((n + 7) / 8, v.i64 ? 0xff : 0);
      break;
    case DType::Float64: {
      o.data.resize(n * 8);
      for (uint64_t i = 0; i < n; ++i) memcpy(o.data.data() + i * 8, &v.f64, 8);
      break;
    }
    case DType::Int64: {
      o.data.resize(n * 8);
      for (uint64_t i = 0; i < n; ++i) memcpy(o.data.data() + i * 8, &v.i64, 8);
      break;
    }
    default: {  // Int32, Date32
      const int32_t x = static_cast<int32_t>(v.i64);
      o.data.resize(n * 4);
      for (uint64_t i = 0; i < n; ++i) memcpy(o.data.data() + i * 4, &x, 4);
    }
  }
  return o;
}

// Evaluate one expression root into an output column.
static OutColumn eval_expr(prism_ctx* ctx, const HostBatch& hb, const prism_expr_node* exprs, uint32_t n_exprs,
                           const prism_pred_node* preds, uint32_t n_preds, int32_t root, const std::string& name) {
  Compiler C(hb.cols, preds, n_preds, exprs, n_exprs);
  const prism_expr_node& e = C.enode(root);
  const uint64_t n = static_cast<uint64_t>(hb.n_rows);
  if (e.kind == PRISM_EXPR_COLUMN) return host_copy(C.column(e.column), name);
  if (e.kind == PRISM_EXPR_LITERAL) return literal_column(e.value, n, name);
  if (e.kind == PRISM_EXPR_IF_THEN && e.n_clauses == 0) {
    if (e.else_expr < 0)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: IfThen expression requires at least one clause or an else branch");
    return eval_expr(ctx, hb, exprs, n_exprs, preds, n_preds, e.else_expr, name);
  }
  const DType t = C.type_of(root);
  if (t == DType::Utf8) {
    // IfThen over string branches: a program picks the source per row.
    if (e.n_clauses > static_cast<uint32_t>(kMaxSrc - 1))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: at most 15 IfThen clauses over Utf8");
    if (!e.clause_preds || !e.clause_exprs) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL IfThen clauses");
    auto S = std::make_unique<StrSrcs>();
    memset(S.get(), 0, sizeof(StrSrcs));
    std::vector<int32_t> src_expr;
    for (uint32_t k = 0; k < e.n_clauses; ++k) src_expr.push_back(e.clause_exprs[k]);
    if (e.else_expr >= 0) src_expr.push_back(e.else_expr);
    if (e.else_expr >= 0) C.emit(PRISM_OP_CONST_I, 0, static_cast<int64_t>(e.n_clauses));
    else C.emit(PRISM_OP_CONST_NULL);
    for (int64_t k = static_cast<int64_t>(e.n_clauses) - 1; k >= 0; --k) {
      C.pred(e.clause_preds[k], 1);
      C.emit(PRISM_OP_CONST_I, 0, k);
      C.emit(PRISM_OP_ZIP);
    }
    std::vector<bool> used(hb.cols.size(), false);
    collect_cols(C.code, used);
    int lit_bytes = 0;
    for (size_t k = 0; k < src_expr.size(); ++k) {
      const prism_expr_node& b = C.enode(src_expr[k]);
      if (b.kind == PRISM_EXPR_COLUMN && C.column(b.column).type == DType::Utf8) {
        used[b.column] = true;
        S->s[k].is_lit = 0;
      } else if (b.kind == PRISM_EXPR_LITERAL && b.value.type == PRISM_SCALAR_UTF8) {
        const int l = static_cast<int>(b.value.str ? b.value.str_len : 0);
        if (lit_bytes + l > kMaxStrBytes) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: IfThen literals too long");
        S->s[k].is_lit = 1;
        S->s[k].lit_off = lit_bytes;
        S->s[k].lit_len = l;
        if (l) memcpy(S->lit + lit_bytes, b.value.str, l);
        lit_bytes += l;
      } else {
        fail(PRISM_ERR_ARROW, "Arrow error: Cast error: IfThen branch of type " + type_debug(C.type_of(src_expr[k])) +
                                  " cannot be cast to Utf8 by the device evaluator");
      }
    }
    DevBatch db;
    upload_cols(ctx, hb, used, db);
    for (size_t k = 0; k < src_expr.size(); ++k)
      if (!S->s[k].is_lit) S->s[k].col = db.cols[C.enode(src_expr[k]).column].desc();
    S->n = static_cast<int32_t>(src_expr.size());
    DeviceBuffer prog;
    int depth = 1;
    C.finish(ctx, prog, &depth);
    EvalOut idx;
    run_eval(ctx, prog.as<DevProgram>(), depth, db.desc.as<prism_column>(), n, OUT_W4, idx);
    Ctx* c = as_ctx(ctx);
    hipStream_t st = ctx->stream;
    DeviceBuffer dS(c, sizeof(StrSrcs)), len(c, (n ? n : 1) * 8), start(c, (n ? n : 1) * 8), valid(c, (ceil_div(n, 64) + 1) * 8),
        cnt(c, 16);
    PRISM_HIP(hipMemcpyAsync(dS.ptr, S.get(), sizeof(StrSrcs), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, cnt.as<unsigned long long>(), 2);
    DevColumn out;
    out.type = DType::Utf8;
    out.name = name;
    out.length = static_cast<int64_t>(n);
    out.offsets.alloc(c, (n + 1) * 4);
    uint64_t total = 0;
    if (n) {
      const dim3 g = egrid(ctx, ceil_div(n, 64), 1);
      hipLaunchKernelGGL(k_str_select_len, g, dim3(kEB), 0, st, dS.as<StrSrcs>(), idx.values.as<uint32_t>(),
                         idx.valid.as<unsigned long long>(), n, len.as<uint64_t>(), valid.as<unsigned long long>(),
                         cnt.as<unsigned long long>());
      size_t tmp = 0;
      PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, len.as<uint64_t>(), start.as<uint64_t>(), static_cast<int>(n), st));
      DeviceBuffer t(c, tmp);
      PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(t.ptr, tmp, len.as<uint64_t>(), start.as<uint64_t>(), static_cast<int>(n), st));
      uint64_t last[2];
      PRISM_HIP(hipMemcpyAsync(&last[0], start.as<uint64_t>() + n - 1, 8, hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipMemcpyAsync(&last[1], len.as<uint64_t>() + n - 1, 8, hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      total = last[0] + last[1];
      if (total > 0x7fffffffULL) fail(PRISM_ERR_EXECUTION, "Execution error: Utf8 result exceeds i32 offsets");
      out.data.alloc(c, total ? total : 1);
      hipLaunchKernelGGL(k_str_select_copy, egrid(ctx, ceil_div(n, 64), 1), dim3(kEB), 0, st, dS.as<StrSrcs>(),
                         idx.values.as<uint32_t>(), idx.valid.as<unsigned long long>(), n, start.as<uint64_t>(),
                         len.as<uint64_t>(), out.offsets.as<int32_t>(), out.data.as<uint8_t>());
      PRISM_CHECK_LAUNCH();
      out.validity = std::move(valid);
    } else {
      out.data.alloc(c, 1);
      int32_t z = 0;
      PRISM_HIP(hipMemcpyAsync(out.offsets.ptr, &z, 4, hipMemcpyHostToDevice, st));
      PRISM_HIP(hipStreamSynchronize(st));
    }
    OutColumn o = download(ctx, out);
    o.nullable = true;
    return o;
  }
  C.value(root, t, 0);
  std::vector<bool> used(hb.cols.size(), false);
  collect_cols(C.code, used);
  DevBatch db;
  upload_cols(ctx, hb, used, db);
  DeviceBuffer prog;
  int depth = 1;
  C.finish(ctx, prog, &depth);
  EvalOut o;
  const OutKind kind = t == DType::Bool ? OUT_BITS : (t == DType::Float64 || t == DType::Int64) ? OUT_W8 : OUT_W4;
  run_eval(ctx, prog.as<DevProgram>(), depth, db.desc.as<prism_column>(), n, kind, o);
  return bits_column(ctx, o, n, name, t);
}

// Rows of `hb` selected by mask words, for the listed columns.
static void compact(prism_ctx* ctx, const HostBatch& hb, const unsigned long long* words,
                    const std::vector<uint32_t>& keep, ArrowArray* out, ArrowSchema* out_schema) {
  DeviceBuffer sel;
  const uint64_t m = select_bits(ctx, words, static_cast<uint64_t>(hb.n_rows), sel);
  std::vector<OutColumn> cols;
  for (uint32_t ci : keep) {
    const HostColumn& h = hb.cols[ci];
    if (m == static_cast<uint64_t>(hb.n_rows)) {  // everything kept
      cols.push_back(host_copy(h, h.name));
      continue;
    }
    DevColumn d = upload(ctx, h);
    DevColumn t = take(ctx, d, sel.as<uint32_t>(), nullptr, m);
    t.nullable = h.nullable;
    OutColumn o = download(ctx, t);
    o.nullable = h.nullable;
    cols.push_back(std::move(o));
  }
  export_batch(std::move(cols), static_cast<int64_t>(m), out, out_schema);
}

}  // namespace fp
}  // namespace prism

using namespace prism;
using namespace prism::fp;

extern "C" {

prism_status prism_evaluate_predicate(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                      const prism_pred_node* nodes, uint32_t n_nodes, struct ArrowArray* out_mask,
                                      struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !out_mask || !out_schema) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/output");
    Lane lane(ctx);
    HostBatch hb = import_batch(batch, schema);
    EvalOut o;
    DevBatch db;
    predicate_words(ctx, hb, nodes, n_nodes, o, db);
    export_array(bits_column(ctx, o, static_cast<uint64_t>(hb.n_rows), "mask", DType::Bool), out_mask, out_schema);
  });
}

prism_status prism_filter_batch(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                const struct ArrowArray* mask, const struct ArrowSchema* mask_schema,
                                struct ArrowArray* out, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !mask || !mask_schema || !out || !out_schema)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/mask/output");
    Lane lane(ctx);
    HostBatch hb = import_batch(batch, schema);
    // The mask arrives as a one-column batch view of the BooleanArray.
    ArrowSchema wrap_schema;
    memset(&wrap_schema, 0, sizeof(wrap_schema));
    wrap_schema.format = "+s";
    wrap_schema.name = "";
    wrap_schema.n_children = 1;
    ArrowSchema* sch_child = const_cast<ArrowSchema*>(mask_schema);
    wrap_schema.children = &sch_child;
    ArrowArray wrap;
    memset(&wrap, 0, sizeof(wrap));
    wrap.length = mask->length;
    wrap.n_buffers = 1;
    const void* no_validity = nullptr;
    wrap.buffers = &no_validity;
    wrap.n_children = 1;
    ArrowArray* arr_child = const_cast<ArrowArray*>(mask);
    wrap.children = &arr_child;
    HostBatch mb = import_batch(&wrap, &wrap_schema);
    const HostColumn& mc = mb.cols[0];
    if (mc.type != DType::Bool) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: filter mask is not Boolean");
    if (mc.length != hb.n_rows)
      fail(PRISM_ERR_ARROW, "Arrow error: Invalid argument error: Filter predicate of length " +
                                std::to_string(mc.length) + " is larger than target array of length " +
                                std::to_string(hb.n_rows));
    DevColumn dm = upload(ctx, mc);
    Ctx* c = as_ctx(ctx);
    const uint64_t n = static_cast<uint64_t>(hb.n_rows);
    DeviceBuffer words(c, (ceil_div(n, 64) + 1) * 8);
    if (n)
      hipLaunchKernelGGL(k_mask_words, egrid(ctx, ceil_div(n, 64), 1), dim3(kEB), 0, ctx->stream, dm.data.as<uint8_t>(),
                         dm.validity.ptr ? dm.validity.as<uint8_t>() : nullptr, dm.dev_offset, n,
                         words.as<unsigned long long>());
    PRISM_CHECK_LAUNCH();
    std::vector<uint32_t> keep(hb.cols.size());
    for (size_t i = 0; i < keep.size(); ++i) keep[i] = static_cast<uint32_t>(i);
    compact(ctx, hb, words.as<unsigned long long>(), keep, out, out_schema);
  });
}

prism_status prism_project_batch(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                 const uint32_t* indices, uint32_t n_indices, struct ArrowArray* out,
                                 struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !out || !out_schema || (n_indices && !indices))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/indices/output");
    HostBatch hb = import_batch(batch, schema);
    std::vector<OutColumn> cols;
    for (uint32_t k = 0; k < n_indices; ++k) {
      if (indices[k] >= hb.cols.size())
        fail(PRISM_ERR_ARROW, "Arrow error: Schema error: project index " + std::to_string(indices[k]) +
                                  " out of bounds, max field " + std::to_string(hb.cols.size()));
      cols.push_back(host_copy(hb.cols[indices[k]], hb.cols[indices[k]].name));
    }
    export_batch(std::move(cols), hb.n_rows, out, out_schema);
  });
}

prism_status prism_filter_and_project(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                      const prism_pred_node* nodes, uint32_t n_nodes, const uint32_t* indices,
                                      uint32_t n_indices, struct ArrowArray* out, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !out || !out_schema || (n_indices && !indices))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/indices/output");
    Lane lane(ctx);
    HostBatch hb = import_batch(batch, schema);
    std::vector<uint32_t> keep;
    if (n_indices == 0) {
      for (size_t i = 0; i < hb.cols.size(); ++i) keep.push_back(static_cast<uint32_t>(i));
    } else {
      for (uint32_t k = 0; k < n_indices; ++k) {
        if (indices[k] >= hb.cols.size())
          fail(PRISM_ERR_ARROW, "Arrow error: Schema error: project index " + std::to_string(indices[k]) +
                                    " out of bounds, max field " + std::to_string(hb.cols.size()));
        keep.push_back(indices[k]);
      }
    }
    if (!nodes || n_nodes == 0) {
      std::vector<OutColumn> cols;
      for (uint32_t ci : keep) cols.push_back(host_copy(hb.cols[ci], hb.cols[ci].name));
      export_batch(std::move(cols), hb.n_rows, out, out_schema);
      return;
    }
    EvalOut o;
    DevBatch db;
    predicate_words(ctx, hb, nodes, n_nodes, o, db);
    // mask words already fold validity in: value bits are set only for valid rows
    compact(ctx, hb, o.values.as<unsigned long long>(), keep, out, out_schema);
  });
}

prism_status prism_evaluate_scalar_expr(prism_ctx* ctx, const struct ArrowArray* batch,
                                        const struct ArrowSchema* schema, const prism_expr_node* exprs,
                                        uint32_t n_exprs, const prism_pred_node* preds, uint32_t n_preds,
                                        int32_t root, struct ArrowArray* out, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !out || !out_schema) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/output");
    Lane lane(ctx);
    HostBatch hb = import_batch(batch, schema);
    export_array(eval_expr(ctx, hb, exprs, n_exprs, preds, n_preds, root, "expr"), out, out_schema);
  });
}

prism_status prism_project_batch_with_exprs(prism_ctx* ctx, const struct ArrowArray* batch,
                                            const struct ArrowSchema* schema, const uint32_t* indices,
                                            uint32_t n_indices, const prism_expr_node* exprs, uint32_t n_exprs,
                                            const int32_t* roots, uint32_t n_roots, const prism_pred_node* preds,
                                            uint32_t n_preds, struct ArrowArray* out, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !out || !out_schema || (n_indices && !indices) || (n_roots && !roots))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/indices/roots/output");
    Lane lane(ctx);
    HostBatch hb = import_batch(batch, schema);
    std::vector<OutColumn> cols;
    for (uint32_t k = 0; k < n_indices; ++k) {
      if (indices[k] >= hb.cols.size())
        fail(PRISM_ERR_ARROW, "Arrow error: Schema error: project index " + std::to_string(indices[k]) +
                                  " out of bounds, max field " + std::to_string(hb.cols.size()));
      cols.push_back(host_copy(hb.cols[indices[k]], hb.cols[indices[k]].name));
    }
    for (uint32_t k = 0; k < n_roots; ++k) {
      OutColumn c = eval_expr(ctx, hb, exprs, n_exprs, preds, n_preds, roots[k], "expr_" + std::to_string(k));
      c.nullable = true;
      cols.push_back(std::move(c));
    }
    export_batch(std::move(cols), hb.n_rows, out, out_schema);
  });
}

prism_status prism_filter_mask(prism_ctx* ctx, const prism_instr* code, uint32_t n_code, const prism_string* strings,
                               uint32_t n_strings, const int32_t* column_types, const prism_column* columns,
                               uint32_t n_columns, uint64_t n_rows, unsigned long long* d_mask, uint64_t* selected) {
  return guarded([&] {
    if (!ctx || !code || !selected || (n_columns && (!column_types || !columns)))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/program/columns/output");
    if (n_columns > static_cast<uint32_t>(kMaxCols)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: at most 16 columns");
    use_device(ctx);
    int depth = 1;
    const ValType vt = check_program(code, n_code, column_types, n_columns, n_strings, &depth);
    if (vt != VT_B && vt != VT_NULL) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: filter program is not boolean");
    auto P = std::make_unique<DevProgram>();
    memset(P.get(), 0, sizeof(DevProgram));
    P->len = static_cast<int32_t>(n_code);
    memcpy(P->code, code, n_code * sizeof(prism_instr));
    build_str_table(strings, n_strings, P->strs);
    LikeScan like;
    like.col = like_scan_prepare(code, n_code, P->strs, like.arg);
    Ctx* c = as_ctx(ctx);
    DeviceBuffer prog(c, sizeof(DevProgram)), desc(c, (n_columns ? n_columns : 1) * sizeof(prism_column));
    PRISM_HIP(hipMemcpyAsync(prog.ptr, P.get(), sizeof(DevProgram), hipMemcpyHostToDevice, ctx->stream));
    if (n_columns)
      PRISM_HIP(hipMemcpyAsync(desc.ptr, columns, n_columns * sizeof(prism_column), hipMemcpyHostToDevice, ctx->stream));
    EvalOut o;
    run_eval(ctx, prog.as<DevProgram>(), depth, desc.as<prism_column>(), n_rows, OUT_BITS, o, d_mask, &like);
    *selected = o.selected;
  });
}

prism_status prism_string_like(prism_ctx* ctx, const struct ArrowArray* array, const struct ArrowSchema* schema,
                               const char* pattern, uint32_t pattern_len, int32_t case_insensitive,
                               struct ArrowArray* out, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !array || !schema || !out || !out_schema)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/array/output");
    Lane lane(ctx);
    ArrowSchema wrap_schema;
    memset(&wrap_schema, 0, sizeof(wrap_schema));
    wrap_schema.format = "+s";
    wrap_schema.name = "";
    wrap_schema.n_children = 1;
    ArrowSchema* sch_child = const_cast<ArrowSchema*>(schema);
    wrap_schema.children = &sch_child;
    ArrowArray wrap;
    memset(&wrap, 0, sizeof(wrap));
    wrap.length = array->length;
    wrap.n_buffers = 1;
    const void* no_validity = nullptr;
    wrap.buffers = &no_validity;
    wrap.n_children = 1;
    ArrowArray* arr_child = const_cast<ArrowArray*>(array);
    wrap.children = &arr_child;
    HostBatch hb = import_batch(&wrap, &wrap_schema);
    prism_pred_node node;
    memset(&node, 0, sizeof(node));
    node.kind = case_insensitive ? PRISM_PRED_ILIKE : PRISM_PRED_LIKE;
    node.column = 0;
    node.pattern = pattern;
    node.pattern_len = pattern ? pattern_len : 0;
    EvalOut o;
    DevBatch db;
    predicate_words(ctx, hb, &node, 1, o, db);
    export_array(bits_column(ctx, o, static_cast<uint64_t>(hb.n_rows), "like", DType::Bool), out, out_schema);
  });
}

}  // extern "C"
