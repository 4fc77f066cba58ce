// Host-side validation of row programs (see program.hpp).
#include "program.hpp"

namespace prism {

static const char* op_name(int op) {
  switch (op) {
    case PRISM_OP_LOAD_I: return "LOAD_I";
    case PRISM_OP_LOAD_F: return "LOAD_F";
    case PRISM_OP_CONST_I: return "CONST_I";
    case PRISM_OP_CONST_F: return "CONST_F";
    case PRISM_OP_I2F: return "I2F";
    case PRISM_OP_ADD_F: return "ADD_F";
    case PRISM_OP_SUB_F: return "SUB_F";
    case PRISM_OP_MUL_F: return "MUL_F";
    case PRISM_OP_DIV_F: return "DIV_F";
    case PRISM_OP_NEG_F: return "NEG_F";
    case PRISM_OP_ADD_I: return "ADD_I";
    case PRISM_OP_SUB_I: return "SUB_I";
    case PRISM_OP_MUL_I: return "MUL_I";
    case PRISM_OP_NEG_I: return "NEG_I";
    case PRISM_OP_CMP_I: return "CMP_I";
    case PRISM_OP_CMP_F: return "CMP_F";
    case PRISM_OP_CMP_STR: return "CMP_STR";
    case PRISM_OP_LIKE: return "LIKE";
    case PRISM_OP_ILIKE: return "ILIKE";
    case PRISM_OP_IS_NULL: return "IS_NULL";
    case PRISM_OP_IS_NOT_NULL: return "IS_NOT_NULL";
    case PRISM_OP_AND: return "AND";
    case PRISM_OP_OR: return "OR";
    case PRISM_OP_NOT: return "NOT";
    case PRISM_OP_CONST_BOOL: return "CONST_BOOL";
    case PRISM_OP_SELECT: return "SELECT";
    case PRISM_OP_CONST_NULL: return "CONST_NULL";
    case PRISM_OP_ZIP: return "ZIP";
    default: return "?";
  }
}

ValType check_program(const prism_instr* code, uint32_t len, const int32_t* col_types, uint32_t n_cols,
                      uint32_t n_strings, int* max_depth) {
  auto bad = [&](uint32_t pc, const std::string& why) {
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: row program instruction " + std::to_string(pc) + " (" +
                                         op_name(code[pc].op) + "): " + why);
  };
  if (len == 0 || !code) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: empty row program");
  if (len > static_cast<uint32_t>(kMaxCode)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: row program too long");
  ValType st[kStackDepth];
  int sp = 0, deepest = 0;
  auto push = [&](uint32_t pc, ValType t) {
    if (sp >= kStackDepth) bad(pc, "stack deeper than " + std::to_string(kStackDepth));
    st[sp++] = t;
    deepest = sp > deepest ? sp : deepest;
  };
  auto need = [&](uint32_t pc, int k) {
    if (sp < k) bad(pc, "stack underflow");
  };
  auto col = [&](uint32_t pc) -> int32_t {
    int a = code[pc].arg;
    if (a < 0 || static_cast<uint32_t>(a) >= n_cols) bad(pc, "column index out of range");
    return col_types[a];
  };
  for (uint32_t pc = 0; pc < len; ++pc) {
    const prism_instr& in = code[pc];
    switch (in.op) {
      case PRISM_OP_LOAD_I: {
        int32_t t = col(pc);
        if (t != PRISM_TYPE_INT32 && t != PRISM_TYPE_INT64 && t != PRISM_TYPE_DATE32 && t != PRISM_TYPE_BOOL)
          bad(pc, "column is not an integer/date/bool column");
        push(pc, t == PRISM_TYPE_BOOL ? VT_B : VT_I);
        break;
      }
      case PRISM_OP_LOAD_F:
        if (col(pc) != PRISM_TYPE_FLOAT64) bad(pc, "column is not Float64");
        push(pc, VT_F);
        break;
      case PRISM_OP_CONST_I: push(pc, VT_I); break;
      case PRISM_OP_CONST_F: push(pc, VT_F); break;
      case PRISM_OP_CONST_BOOL: push(pc, VT_B); break;
      case PRISM_OP_CONST_NULL: push(pc, VT_NULL); break;
      case PRISM_OP_I2F:
        need(pc, 1);
        if (st[sp - 1] != VT_I && st[sp - 1] != VT_B && st[sp - 1] != VT_NULL) bad(pc, "operand is not integer");
        st[sp - 1] = VT_F;
        break;
      case PRISM_OP_NEG_F:
        need(pc, 1);
        if (st[sp - 1] != VT_F) bad(pc, "operand is not f64");
        break;
      case PRISM_OP_NEG_I:
        need(pc, 1);
        if (st[sp - 1] != VT_I) bad(pc, "operand is not int64");
        break;
      case PRISM_OP_ADD_F: case PRISM_OP_SUB_F: case PRISM_OP_MUL_F: case PRISM_OP_DIV_F:
        need(pc, 2);
        if (st[sp - 1] != VT_F || st[sp - 2] != VT_F) bad(pc, "operands are not f64");
        --sp;
        break;
      case PRISM_OP_ADD_I: case PRISM_OP_SUB_I: case PRISM_OP_MUL_I:
        need(pc, 2);
        if (st[sp - 1] != VT_I || st[sp - 2] != VT_I) bad(pc, "operands are not int64");
        --sp;
        break;
      case PRISM_OP_CMP_I:
        need(pc, 2);
        if ((st[sp - 1] != VT_I && st[sp - 1] != VT_B) || (st[sp - 2] != VT_I && st[sp - 2] != VT_B))
          bad(pc, "operands are not integer");
        if (in.arg < 0 || in.arg > 5) bad(pc, "bad comparison");
        --sp;
        st[sp - 1] = VT_B;
        break;
      case PRISM_OP_CMP_F:
        need(pc, 2);
        if (st[sp - 1] != VT_F || st[sp - 2] != VT_F) bad(pc, "operands are not f64");
        if (in.arg < 0 || in.arg > 5) bad(pc, "bad comparison");
        --sp;
        st[sp - 1] = VT_B;
        break;
      case PRISM_OP_CMP_STR: case PRISM_OP_LIKE: case PRISM_OP_ILIKE: {
        if (col(pc) != PRISM_TYPE_UTF8) bad(pc, "column is not Utf8");
        int64_t s = in.i64 & 0xff;
        if (s < 0 || s >= static_cast<int64_t>(n_strings)) bad(pc, "string index out of range");
        if (in.op == PRISM_OP_CMP_STR && ((in.i64 >> 8) < 0 || (in.i64 >> 8) > 5)) bad(pc, "bad comparison");
        push(pc, VT_B);
        break;
      }
      case PRISM_OP_IS_NULL: case PRISM_OP_IS_NOT_NULL:
        col(pc);
        push(pc, VT_B);
        break;
      case PRISM_OP_AND: case PRISM_OP_OR:
        need(pc, 2);
        if ((st[sp - 1] != VT_B && st[sp - 1] != VT_NULL) || (st[sp - 2] != VT_B && st[sp - 2] != VT_NULL))
          bad(pc, "operands are not boolean");
        --sp;
        st[sp - 1] = VT_B;
        break;
      case PRISM_OP_NOT:
        need(pc, 1);
        if (st[sp - 1] != VT_B && st[sp - 1] != VT_NULL) bad(pc, "operand is not boolean");
        st[sp - 1] = VT_B;
        break;
      case PRISM_OP_ZIP: {
        need(pc, 3);
        ValType e = st[sp - 3], c = st[sp - 2], a = st[sp - 1];
        if (c != VT_B && c != VT_NULL) bad(pc, "condition is not boolean");
        if (!(a == e || a == VT_NULL || e == VT_NULL)) bad(pc, "branches differ in type");
        sp -= 2;
        st[sp - 1] = a == VT_NULL ? e : a;
        break;
      }
      case PRISM_OP_SELECT: {
        need(pc, 3);
        ValType c = st[sp - 3], a = st[sp - 2], b = st[sp - 1];
        if (c != VT_B && c != VT_NULL) bad(pc, "condition is not boolean");
        ValType r = a == VT_NULL ? b : a;
        if (!(a == b || a == VT_NULL || b == VT_NULL)) bad(pc, "branches differ in type");
        sp -= 2;
        st[sp - 1] = r;
        break;
      }
      default:
        bad(pc, "unknown opcode " + std::to_string(in.op));
    }
  }
  if (sp != 1) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: row program must leave one value");
  if (max_depth) *max_depth = deepest;
  return st[0];
}

// Shape of a LIKE pattern: '%' only first and/or last, no '_' and no '\'.
static uint8_t like_kind(const uint8_t* p, int32_t n) {
  for (int32_t i = 0; i < n; ++i)
    if (p[i] == '_' || p[i] == '\\' || (p[i] == '%' && i != 0 && i != n - 1)) return LK_GENERAL;
  const bool lead = n >= 1 && p[0] == '%', trail = n >= 2 && p[n - 1] == '%';
  if (n == 1 && lead) return LK_CONTAINS;  // "%": any string
  if (lead && trail) return LK_CONTAINS;
  if (trail) return LK_PREFIX;
  if (lead) return LK_SUFFIX;
  return LK_EXACT;
}

void build_str_table(const prism_string* s, uint32_t n, StrTable& out) {
  if (n > static_cast<uint32_t>(kMaxStrings)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: too many string literals");
  memset(&out, 0, sizeof(out));
  out.n = static_cast<int32_t>(n);
  int32_t pos = 0;
  for (uint32_t i = 0; i < n; ++i) {
    out.off[i] = pos;
    if (s[i].len && !s[i].data) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL string literal");
    if (pos + static_cast<int64_t>(s[i].len) > kMaxStrBytes)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: string literals exceed 1 KiB");
    memcpy(out.bytes + pos, s[i].data, s[i].len);
    out.like_kind[i] = like_kind(reinterpret_cast<const uint8_t*>(s[i].data), static_cast<int32_t>(s[i].len));
    pos += static_cast<int32_t>(s[i].len);
  }
  out.off[n] = pos;
}

}  // namespace prism
