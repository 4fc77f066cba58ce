// Group-by update kernels generated per plan and compiled at run time.
//
// The generated kernel does what k_groupby does for the same plan (same LDS
// layout as its shared form, the same global table and merge), with the row
// programs written out as straight-line code over typed registers instead of
// interpreted, constants folded in, and loads of null-free columns without
// validity reads.  Two accumulation forms:
//   wave   - few groups (hash_aggregate.rs:215-269 groups by a 64-bit key
//            hash; a wave's rows rarely hold more than a handful): per round
//            the wave takes the group of its first pending row, every lane
//            sums its rows of that group, DPP reductions give the wave totals
//            and one lane adds them to the LDS (or HBM) slot.  After
//            kWaveRounds rounds the remaining rows take the per-row path.
//   shared - per row: LDS slot found by hashing, LDS atomics; HBM when the
//            LDS table is full.  The faster form unless every row falls in one
//            group (global aggregates), where all lanes would hit one LDS word.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "gb_jit.hpp"
#include "jit.hpp"

namespace prism {
namespace gb {
namespace {

constexpr int kWaveRounds = 24;

std::string hex64(uint64_t v) {
  char b[40];
  snprintf(b, sizeof(b), "0x%016llxULL", static_cast<unsigned long long>(v));
  return b;
}

const char* cmp_op(int c) {
  switch (c) {
    case PRISM_CMP_EQ: return "==";
    case PRISM_CMP_NE: return "!=";
    case PRISM_CMP_LT: return "<";
    case PRISM_CMP_LE: return "<=";
    case PRISM_CMP_GT: return ">";
    default: return ">=";
  }
}

struct Val {
  std::string v;   // expression of the value
  std::string ok;  // expression of its validity
  char t;          // 'I' int64 word, 'F' f64
};

class Gen {
 public:
  Gen(const prism_groupby* g, int form, int rows, uint32_t nullable)
      : g_(g), D_(g->plan), form_(form), R_(rows), nullable_(nullable) {}

  std::string run() {
    scan_columns();
    o_ << "#include \"dev_groupby.hpp\"\n"
          "using namespace prism;\nusing namespace prism::gb;\n"
       << "#define R_ " << R_ << "\n#define STRIDE " << D_.stride << "\n"
       << kJitWavesHint << "\n"
       // rows it0, it0 + step, ... below n, each batch 256 * R_ rows of this workgroup
       << "__device__ __forceinline__ void gb_body(const DevPlan* __restrict__ P, const prism_column* __restrict__ cols,\n"
          "    uint64_t it0, uint64_t step, uint64_t n, uint64_t row_base, GTable G, unsigned long long* lds,\n"
          "    int lds_slots, int lds_shift, uint32_t* flags) {\n"
          "  const int tid = threadIdx.x;\n"
          "  const int lane = tid & 63;\n"
          "  (void)lane;\n";
    declare_columns();
    o_ << "  for (uint64_t it = it0; it < n; it += step) {\n"
          "    uint64_t row[R_];\n    bool in[R_];\n"
          "#pragma unroll\n"
          "    for (int r = 0; r < R_; ++r) { row[r] = it + r * 256 + tid; in[r] = row[r] < n; }\n"
          // rows past the end load row n - 1 (no branches around the loads)
          "    uint64_t rr[R_];\n"
          "#pragma unroll\n"
          "    for (int r = 0; r < R_; ++r) rr[r] = in[r] ? row[r] : n - 1;\n";
    load_columns();
    if (D_.filter_len) filter();
    keys();
    values();
    if (form_ == kJitWave) wave_accumulate();
    else row_accumulate("in[r]");
    o_ << "  }\n}\n";
    // one launch over rows [first, n): grid-stride batches
    o_ << "extern \"C\" __global__ __launch_bounds__(256) PRISM_WAVES_HINT void " << kJitGroupbyKernel
       << "(const DevPlan* __restrict__ P, const prism_column* __restrict__ cols, uint64_t first, uint64_t n,\n"
          "    uint64_t row_base, GTable G, int lds_slots, int lds_shift, uint32_t* flags) {\n"
          "  extern __shared__ unsigned long long lds[];\n"
          "  gb_init<false>(P, lds, lds_slots);\n"
          "  gb_body(P, cols, first + static_cast<uint64_t>(blockIdx.x) * 256 * R_, static_cast<uint64_t>(gridDim.x) * 256 * R_,\n"
          "          n, row_base, G, lds, lds_slots, lds_shift, flags);\n"
          "  gb_flush<false>(P, G, lds, lds_slots, flags);\n}\n";
    // persistent tile-queue form: each workgroup takes (request, row range)
    // tiles of requests of this plan until the queue is empty
    o_ << "extern \"C\" __global__ __launch_bounds__(256) PRISM_WAVES_HINT void " << kJitGroupbyTiles
       << "(TileQueue* q, const RowTile* __restrict__ tiles, const TileReq* __restrict__ reqs) {\n"
          "  extern __shared__ unsigned long long lds[];\n"
          "  __shared__ uint32_t s_t;\n"
          "  const uint32_t n_tiles = q->n_tiles;\n"
          "  bool first = true;\n"
          "  while (true) {\n"
          "    const uint32_t t = next_tile(q, &s_t, first);\n"
          "    if (t >= n_tiles) break;\n"
          "    const RowTile tl = tiles[q->offset + t];\n"
          "    const TileReq& d = tile_req(reqs, tl.req);\n"
          "    gb_init<false>(d.P, lds, d.lds_slots);\n"
          "    gb_body(d.P, d.cols, tl.start, 256 * R_, tl.end, 0, d.G, lds, d.lds_slots, d.lds_shift, d.flags);\n"
          "    gb_flush<false>(d.P, d.G, lds, d.lds_slots, d.flags);\n"
          "    __syncthreads();\n"
          "  }\n}\n";
    return o_.str();
  }

 private:
  const prism_groupby* g_;
  const DevPlan& D_;
  int form_, R_;
  uint32_t nullable_;
  std::ostringstream o_;
  bool fixed_[kMaxCols] = {};  // fixed-width columns loaded once per row batch
  bool used_[kMaxCols] = {};   // columns whose descriptors the kernel reads
  int uid_ = 0;

  int type(int c) const { return g_->col_types[c]; }
  bool nullable(int c) const { return (nullable_ >> c) & 1; }
  bool is_fixed(int c) const { return type(c) != PRISM_TYPE_UTF8; }
  std::string fresh() { return "t" + std::to_string(uid_++); }

  void scan_columns() {
    auto use = [&](int c) {
      used_[c] = true;
      if (is_fixed(c)) fixed_[c] = true;
    };
    auto scan = [&](int start, int n) {
      for (int pc = start; pc < start + n; ++pc) {
        const prism_instr& ins = D_.code[pc];
        switch (ins.op) {
          case kOpLoadPre: use(D_.pref_col[ins.arg]); break;
          case PRISM_OP_LOAD_I:
          case PRISM_OP_LOAD_F:
          case PRISM_OP_CMP_STR:
          case PRISM_OP_LIKE:
          case PRISM_OP_ILIKE:
            use(ins.arg);
            break;
          case PRISM_OP_IS_NULL:
          case PRISM_OP_IS_NOT_NULL:
            used_[ins.arg] = true;
            break;
          default:
            break;
        }
      }
    };
    if (D_.filter_len) scan(D_.filter_start, D_.filter_len);
    for (int v = 0; v < D_.n_values; ++v) scan(D_.value_start[v], D_.value_len[v]);
    for (int k = 0; k < D_.n_keys; ++k) use(D_.key_cols[k]);
  }

  void declare_columns() {
    for (int c = 0; c < kMaxCols; ++c) {
      if (!used_[c]) continue;
      const std::string C = std::to_string(c);
      switch (type(c)) {
        case PRISM_TYPE_INT64:
          o_ << "  const PRISM_GLOBAL int64_t* __restrict__ d" << C << " = (const PRISM_GLOBAL int64_t*)(static_cast<const int64_t*>(cols["
             << C << "].data) + cols[" << C << "].offset);\n";
          break;
        case PRISM_TYPE_FLOAT64:
          o_ << "  const PRISM_GLOBAL double* __restrict__ d" << C << " = (const PRISM_GLOBAL double*)(static_cast<const double*>(cols["
             << C << "].data) + cols[" << C << "].offset);\n";
          break;
        case PRISM_TYPE_INT32:
        case PRISM_TYPE_DATE32:
          o_ << "  const PRISM_GLOBAL int32_t* __restrict__ d" << C << " = (const PRISM_GLOBAL int32_t*)(static_cast<const int32_t*>(cols["
             << C << "].data) + cols[" << C << "].offset);\n";
          break;
        case PRISM_TYPE_BOOL:
          o_ << "  const PRISM_GLOBAL uint8_t* __restrict__ d" << C << " = (const PRISM_GLOBAL uint8_t*)(cols[" << C << "].data);\n";
          break;
        default:  // UTF8
          o_ << "  const PRISM_GLOBAL int32_t* __restrict__ of" << C << " = (const PRISM_GLOBAL int32_t*)(cols[" << C
             << "].offsets + cols[" << C << "].offset);\n"
             << "  const PRISM_GLOBAL uint8_t* __restrict__ d" << C << " = (const PRISM_GLOBAL uint8_t*)(cols[" << C << "].data);\n";
      }
      if (type(c) == PRISM_TYPE_BOOL || nullable(c)) o_ << "  const uint64_t o" << C << " = cols[" << C << "].offset;\n";
      if (nullable(c))
        o_ << "  const PRISM_GLOBAL uint8_t* __restrict__ n" << C << " = (const PRISM_GLOBAL uint8_t*)(cols[" << C << "].validity);\n";
    }
  }

  // Validity of column c at row expression `rw` (in range is the caller's part).
  std::string valid(int c, const std::string& rw) const {
    if (!nullable(c)) return "true";
    const std::string C = std::to_string(c);
    return "bit_at(n" + C + ", o" + C + " + " + rw + ")";
  }

  void load_columns() {
    for (int c = 0; c < kMaxCols; ++c) {
      if (!fixed_[c]) continue;
      const std::string C = std::to_string(c);
      const bool f = type(c) == PRISM_TYPE_FLOAT64;
      std::string ld;
      switch (type(c)) {
        // columns are streamed once: non-temporal loads leave L2 to the group table's lines
        case PRISM_TYPE_INT64:
        case PRISM_TYPE_FLOAT64:
          ld = "__builtin_nontemporal_load(d" + C + " + rr[r])";
          break;
        case PRISM_TYPE_BOOL:
          ld = "static_cast<int64_t>(bit_at(d" + C + ", o" + C + " + rr[r]))";
          break;
        default:
          ld = "static_cast<int64_t>(__builtin_nontemporal_load(d" + C + " + rr[r]))";
      }
      o_ << "    " << (f ? "double" : "int64_t") << " x" << C << "[R_];\n    bool x" << C << "v[R_];\n"
         << "#pragma unroll\n    for (int r = 0; r < R_; ++r) { x" << C << "[r] = " << ld << "; x" << C
         << "v[r] = in[r] && " << valid(c, "rr[r]") << "; }\n";
    }
  }

  static std::string asI(const Val& a) {
    return a.t == 'I' ? a.v : "static_cast<int64_t>(f_bits(" + a.v + "))";
  }
  static std::string asF(const Val& a) {
    return a.t == 'F' ? a.v : "as_f(static_cast<uint64_t>(" + a.v + "))";
  }
  static std::string raw(const Val& a) {
    return a.t == 'F' ? "f_bits(" + a.v + ")" : "static_cast<uint64_t>(" + a.v + ")";
  }

  // Straight-line code of one program for row r; returns the result.  The
  // statements go to o_ (inside a loop over r).
  Val program(int start, int len) {
    std::vector<Val> st;
    auto pop = [&]() {
      Val v = st.back();
      st.pop_back();
      return v;
    };
    auto push_tmp = [&](const char* ty, const std::string& expr, const std::string& ok, char t) {
      const std::string n = fresh();
      o_ << "      const " << ty << " " << n << " = " << expr << ";\n";
      o_ << "      const bool " << n << "v = " << ok << ";\n";
      st.push_back({n, n + "v", t});
    };
    for (int pc = start; pc < start + len; ++pc) {
      const prism_instr& ins = D_.code[pc];
      switch (ins.op) {
        case kOpLoadPre:
        case PRISM_OP_LOAD_I:
        case PRISM_OP_LOAD_F: {
          const int c = ins.op == kOpLoadPre ? D_.pref_col[ins.arg] : ins.arg;
          const std::string C = std::to_string(c);
          st.push_back({"x" + C + "[r]", "x" + C + "v[r]", type(c) == PRISM_TYPE_FLOAT64 ? 'F' : 'I'});
          break;
        }
        case PRISM_OP_CONST_I:
          st.push_back({"static_cast<int64_t>(" + hex64(static_cast<uint64_t>(ins.i64)) + ")", "true", 'I'});
          break;
        case PRISM_OP_CONST_BOOL:
          st.push_back({ins.i64 != 0 ? "static_cast<int64_t>(1)" : "static_cast<int64_t>(0)", "true", 'I'});
          break;
        case PRISM_OP_CONST_F: {
          uint64_t b;
          memcpy(&b, &ins.f64, 8);
          st.push_back({"as_f(" + hex64(b) + ")", "true", 'F'});
          break;
        }
        case PRISM_OP_CONST_NULL:
          st.push_back({"static_cast<int64_t>(0)", "false", 'I'});
          break;
        case PRISM_OP_I2F: {
          Val a = pop();
          push_tmp("double", "static_cast<double>(" + asI(a) + ")", a.ok, 'F');
          break;
        }
        case PRISM_OP_NEG_F: {
          Val a = pop();
          push_tmp("double", "-(" + asF(a) + ")", a.ok, 'F');
          break;
        }
        case PRISM_OP_NEG_I: {
          Val a = pop();
          const std::string n = fresh();
          o_ << "      int64_t " << n << ";\n      const bool " << n << "v = " << a.ok << ";\n"
             << "      if (__builtin_sub_overflow(static_cast<int64_t>(0), " << asI(a) << ", &" << n << ") && " << n
             << "v) atomicOr(flags, 2u);\n";
          st.push_back({n, n + "v", 'I'});
          break;
        }
        case PRISM_OP_ADD_F:
        case PRISM_OP_SUB_F:
        case PRISM_OP_MUL_F:
        case PRISM_OP_DIV_F: {
          Val b = pop(), a = pop();
          const char* op = ins.op == PRISM_OP_ADD_F ? "+" : ins.op == PRISM_OP_SUB_F ? "-" : ins.op == PRISM_OP_MUL_F ? "*" : "/";
          push_tmp("double", asF(a) + " " + op + " " + asF(b), "(" + a.ok + ") && (" + b.ok + ")", 'F');
          break;
        }
        case PRISM_OP_ADD_I:
        case PRISM_OP_SUB_I:
        case PRISM_OP_MUL_I: {
          Val b = pop(), a = pop();
          const char* fn = ins.op == PRISM_OP_ADD_I ? "__builtin_add_overflow"
                           : ins.op == PRISM_OP_SUB_I ? "__builtin_sub_overflow"
                                                      : "__builtin_mul_overflow";
          const std::string n = fresh();
          o_ << "      int64_t " << n << ";\n      const bool " << n << "v = (" << a.ok << ") && (" << b.ok << ");\n"
             << "      if (" << fn << "(" << asI(a) << ", " << asI(b) << ", &" << n << ") && " << n
             << "v) atomicOr(flags, 2u);\n";
          st.push_back({n, n + "v", 'I'});
          break;
        }
        case PRISM_OP_CMP_I: {
          Val b = pop(), a = pop();
          push_tmp("int64_t", "(" + asI(a) + " " + cmp_op(ins.arg) + " " + asI(b) + ") ? 1 : 0",
                   "(" + a.ok + ") && (" + b.ok + ")", 'I');
          break;
        }
        case PRISM_OP_CMP_F: {
          Val b = pop(), a = pop();
          push_tmp("int64_t",
                   "(total_key(" + asF(a) + ") " + cmp_op(ins.arg) + " total_key(" + asF(b) + ")) ? 1 : 0",
                   "(" + a.ok + ") && (" + b.ok + ")", 'I');
          break;
        }
        case PRISM_OP_CMP_STR:
        case PRISM_OP_LIKE:
        case PRISM_OP_ILIKE: {
          const int c = ins.arg;
          const std::string C = std::to_string(c);
          const int sidx = static_cast<int>(ins.i64 & 0xff);
          const int cop = static_cast<int>(ins.i64 >> 8);
          const int off = D_.strs.off[sidx];
          const int nlit = D_.strs.off[sidx + 1] - off;
          const std::string lit = "P->strs.bytes + " + std::to_string(off);
          std::string e;
          if (ins.op == PRISM_OP_CMP_STR) {
            e = "(str_cmp(s, se - sb, " + lit + ", " + std::to_string(nlit) + ") " + cmp_op(cop) + " 0) ? 1 : 0";
          } else {
            const int lk = D_.strs.like_kind[sidx];
            const char* ci = ins.op == PRISM_OP_ILIKE ? "true" : "false";
            e = lk == LK_GENERAL
                    ? "like_match(s, se - sb, " + lit + ", " + std::to_string(nlit) + ", " + ci + ") ? 1 : 0"
                    : "like_fast(" + std::to_string(lk) + ", s, se - sb, " + lit + ", " + std::to_string(nlit) + ", " +
                          ci + ") ? 1 : 0";
          }
          const std::string n = fresh();
          o_ << "      int64_t " << n << " = 0;\n      bool " << n << "v = false;\n"
             << "      if (in[r]) {\n        const int32_t sb = of" << C << "[row[r]], se = of" << C << "[row[r] + 1];\n"
             << "        const uint8_t* s = (const uint8_t*)(d" << C << " + sb);\n        " << n << " = " << e << ";\n        " << n
             << "v = " << valid(c, "row[r]") << ";\n      }\n";
          st.push_back({n, n + "v", 'I'});
          break;
        }
        case PRISM_OP_IS_NULL:
        case PRISM_OP_IS_NOT_NULL: {
          const std::string nul = "(in[r] && !(" + valid(ins.arg, "row[r]") + "))";
          push_tmp("int64_t", ins.op == PRISM_OP_IS_NULL ? nul + " ? 1 : 0" : nul + " ? 0 : 1", "true", 'I');
          break;
        }
        case PRISM_OP_AND:
        case PRISM_OP_OR: {
          Val b = pop(), a = pop();
          push_tmp("int64_t", asI(a) + (ins.op == PRISM_OP_AND ? " & " : " | ") + asI(b),
                   "(" + a.ok + ") && (" + b.ok + ")", 'I');
          break;
        }
        case PRISM_OP_NOT: {
          Val a = pop();
          push_tmp("int64_t", asI(a) + " ^ 1", a.ok, 'I');
          break;
        }
        case PRISM_OP_ZIP:
        case PRISM_OP_SELECT: {
          Val cond, then_v, else_v;
          if (ins.op == PRISM_OP_ZIP) {  // else, cond, then
            then_v = pop();
            cond = pop();
            else_v = pop();
          } else {  // cond, then, else
            else_v = pop();
            then_v = pop();
            cond = pop();
          }
          const std::string take = fresh();
          o_ << "      const bool " << take << " = (" << cond.ok << ") && ((" << asI(cond) << " & 1) != 0);\n";
          if (then_v.t == else_v.t)
            push_tmp(then_v.t == 'F' ? "double" : "int64_t", take + " ? " + then_v.v + " : " + else_v.v,
                     take + " ? (" + then_v.ok + ") : (" + else_v.ok + ")", then_v.t);
          else
            push_tmp("int64_t",
                     "static_cast<int64_t>(" + take + " ? " + raw(then_v) + " : " + raw(else_v) + ")",
                     take + " ? (" + then_v.ok + ") : (" + else_v.ok + ")", 'I');
          break;
        }
        default:
          throw Error(PRISM_ERR_INTERNAL, "Internal error: no code for row-program opcode " + std::to_string(ins.op));
      }
    }
    if (st.size() != 1) throw Error(PRISM_ERR_INTERNAL, "Internal error: row program leaves a bad stack");
    return st[0];
  }

  void filter() {
    o_ << "#pragma unroll\n    for (int r = 0; r < R_; ++r) {\n";
    Val f = program(D_.filter_start, D_.filter_len);
    o_ << "      in[r] = in[r] && (" << f.ok << ") && ((" << asI(f) << " & 1) != 0);\n    }\n";
    // Rows the filter drops take no further part (no overflow checks on them).
    for (int c = 0; c < kMaxCols; ++c)
      if (fixed_[c])
        o_ << "#pragma unroll\n    for (int r = 0; r < R_; ++r) x" << c << "v[r] = x" << c << "v[r] && in[r];\n";
  }

  void keys() {
    // Utf8 keys: offsets of every row, then the head words of every row, all
    // in flight before the hashing uses them.
    for (int k = 0; k < D_.n_keys; ++k) {
      const int c = D_.key_cols[k];
      if (is_fixed(c)) continue;
      const std::string C = std::to_string(c);
      o_ << "    int32_t sb" << C << "[R_], sl" << C << "[R_];\n#pragma unroll\n"
         << "    for (int r = 0; r < R_; ++r) { sb" << C << "[r] = of" << C << "[rr[r]]; sl" << C << "[r] = of" << C
         << "[rr[r] + 1] - sb" << C << "[r]; }\n";
    }
    for (int k = 0; k < D_.n_keys; ++k) {
      const int c = D_.key_cols[k];
      if (is_fixed(c)) continue;
      const std::string C = std::to_string(c);
      o_ << "    uint32_t hw" << C << "[R_][3];\n#pragma unroll\n"
         << "    for (int r = 0; r < R_; ++r) str_head_load(d" << C << " + sb" << C << "[r], sl" << C << "[r], of" << C
         << " + rr[r], hw" << C << "[r][0], hw" << C << "[r][1], hw" << C << "[r][2]);\n";
    }
    o_ << "    uint64_t h[R_];\n#pragma unroll\n    for (int r = 0; r < R_; ++r) {\n      uint64_t hh = 0;\n";
    for (int k = 0; k < D_.n_keys; ++k) {
      const int c = D_.key_cols[k];
      const std::string C = std::to_string(c);
      if (is_fixed(c)) {
        std::string w;
        switch (type(c)) {
          case PRISM_TYPE_DATE32:
            w = "(static_cast<uint64_t>(static_cast<uint32_t>(x" + C + "[r])) | (1ULL << 32))";
            break;
          case PRISM_TYPE_BOOL:
            w = "(2 + static_cast<uint64_t>(x" + C + "[r]))";
            break;
          case PRISM_TYPE_FLOAT64:
            w = "f_bits(x" + C + "[r])";
            break;
          default:
            w = "static_cast<uint64_t>(x" + C + "[r])";
        }
        o_ << "      hh = fx_step(hh, x" << C << "v[r] ? " << w << " : 0);\n";
      } else {
        const std::string s = "(const uint8_t*)(d" + C + " + sb" + C + "[r])";
        const std::string hashed = "hash_str_head(" + s + ", sl" + C + "[r], str_head_word(" + s + ", sl" + C +
                                   "[r], hw" + C + "[r][0], hw" + C + "[r][1], hw" + C + "[r][2]), hh)";
        if (nullable(c))
          o_ << "      hh = " << valid(c, "rr[r]") << " ? " << hashed << " : fx_step(hh, 0);\n";
        else
          o_ << "      hh = " << hashed << ";\n";
      }
    }
    o_ << "      h[r] = hh;\n    }\n";
  }

  void values() {
    for (int v = 0; v < D_.n_values; ++v) {
      o_ << "    uint64_t w" << v << "[R_];\n    bool w" << v << "v[R_];\n#pragma unroll\n    for (int r = 0; r < R_; ++r) {\n";
      Val x = program(D_.value_start[v], D_.value_len[v]);
      o_ << "      w" << v << "[r] = " << raw(x) << ";\n      w" << v << "v[r] = " << x.ok << ";\n    }\n";
    }
  }

  std::string lane_cond(const Lane& ln, const std::string& row_cond) const {
    const bool want_valid = D_.skip_nulls || ln.op == L_CNT_VALID;
    return want_valid ? "(" + row_cond + ") && w" + std::to_string(ln.value) + "v[r]" : row_cond;
  }

  static std::string lane_lit(const Lane& ln) {
    return "Lane{" + std::to_string(ln.op) + ", " + std::to_string(ln.value) + ", " + std::to_string(ln.word) + ", " +
           std::to_string(ln.is_int) + "}";
  }

  // Per-row accumulation of the rows where `cond` holds (r is the loop index).
  void row_accumulate(const std::string& cond) {
    o_ << "#pragma unroll\n    for (int r = 0; r < R_; ++r) {\n      if (!(" << cond << ")) continue;\n"
       << "      const unsigned long long fr = row_base + row[r];\n"
       << "      const int ls = (lds_slots && h[r] != kEmptyKey) ? lds_find(lds, STRIDE, lds_slots, lds_shift, h[r]) : -1;\n"
       << "      if (ls >= 0) {\n        unsigned long long* L = &lds[ls * STRIDE];\n"
       << "        atomicAdd(&L[kW_COUNT], 1ULL);\n"
       << "        if (fr < __hip_atomic_load(&L[kW_FIRST], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMin(&L[kW_FIRST], fr);\n";
    for (int l = 0; l < D_.n_lanes; ++l) {
      const Lane& ln = D_.lanes[l];
      const bool want_valid = D_.skip_nulls || ln.op == L_CNT_VALID;
      o_ << "        " << (want_valid ? "if (w" + std::to_string(ln.value) + "v[r]) " : "") << "lane_lds(L, "
         << lane_lit(ln) << ", w" << ln.value << "[r]);\n";
    }
    o_ << "      } else {\n        const int64_t gs = g_find(G, h[r], flags);\n        if (gs < 0) continue;\n"
       << "        unsigned long long* gp = &G.w[static_cast<uint64_t>(gs) * STRIDE];\n"
       << "        atomicAdd(&gp[kW_COUNT], 1ULL);\n"
       << "        if (fr < __hip_atomic_load(&gp[kW_FIRST], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&gp[kW_FIRST], fr);\n";
    for (int l = 0; l < D_.n_lanes; ++l) {
      const Lane& ln = D_.lanes[l];
      const bool want_valid = D_.skip_nulls || ln.op == L_CNT_VALID;
      o_ << "        " << (want_valid ? "if (w" + std::to_string(ln.value) + "v[r]) " : "") << "lane_global(gp, "
         << lane_lit(ln) << ", w" << ln.value << "[r]);\n";
    }
    o_ << "      }\n    }\n";
  }

  void wave_accumulate() {
    o_ << "    {\n    uint32_t pend = 0;\n#pragma unroll\n    for (int r = 0; r < R_; ++r) pend |= in[r] ? (1u << r) : 0u;\n"
       << "    const uint64_t row0 = row_base + it + static_cast<uint64_t>(tid & ~63);\n"
       << "    for (int round = 0;; ++round) {\n"
       << "      const uint64_t act = __ballot(pend != 0);\n      if (act == 0) break;\n"
       << "      if (round == " << kWaveRounds << ") {\n";
    row_accumulate("(pend >> r) & 1");
    o_ << "        break;\n      }\n"
       << "      const int lead = __builtin_ctzll(act);\n      uint64_t mine = 0;\n"
       << "#pragma unroll\n      for (int r = R_ - 1; r >= 0; --r) mine = ((pend >> r) & 1) ? h[r] : mine;\n"
       << "      const uint64_t H = readlane_u64(mine, lead);\n      uint32_t take = 0;\n      uint64_t bal[R_];\n"
       << "#pragma unroll\n      for (int r = 0; r < R_; ++r) {\n"
       << "        const bool t = ((pend >> r) & 1) && h[r] == H;\n        take |= t ? (1u << r) : 0u;\n"
       << "        bal[r] = __ballot(t);\n      }\n      pend &= ~take;\n"
       << "      unsigned long long cnt = 0, fr = kNoRow;\n"
       << "#pragma unroll\n      for (int r = R_ - 1; r >= 0; --r) {\n"
       << "        cnt += __builtin_popcountll(bal[r]);\n"
       << "        if (bal[r]) fr = row0 + static_cast<uint64_t>(r) * 256 + __builtin_ctzll(bal[r]);\n      }\n";
    // Lane partials and wave totals.
    for (int l = 0; l < D_.n_lanes; ++l) {
      const Lane& ln = D_.lanes[l];
      const std::string L = std::to_string(l), V = std::to_string(ln.value);
      const std::string cond = lane_cond(ln, "(take >> r) & 1");
      switch (ln.op) {
        case L_SUM_I:
          o_ << "      uint64_t pl" << L << " = 0;\n      int64_t ph" << L << " = 0;\n"
             << "#pragma unroll\n      for (int r = 0; r < R_; ++r) {\n"
             << "        const int64_t x = (" << cond << ") ? static_cast<int64_t>(w" << V << "[r]) : 0;\n"
             << "        pl" << L << " += static_cast<uint32_t>(x);\n        ph" << L << " += x >> 32;\n      }\n"
             << "      const uint64_t sl" << L << " = wave_add_u64(pl" << L << ");\n"
             << "      const uint64_t sh" << L << " = wave_add_u64(static_cast<uint64_t>(ph" << L << "));\n";
          break;
        case L_SUM_F:
          o_ << "      uint64_t pl" << L << " = 0;\n      int64_t ph" << L << " = 0;\n      double pf" << L << " = 0;\n"
             << "#pragma unroll\n      for (int r = 0; r < R_; ++r) {\n        if (!(" << cond << ")) continue;\n"
             << "        const double d = as_f(w" << V << "[r]);\n"
             << "        if (integral_small(d)) { const int64_t x = static_cast<int64_t>(d); pl" << L
             << " += static_cast<uint32_t>(x); ph" << L << " += x >> 32; }\n        else pf" << L << " += d;\n      }\n"
             << "      const uint64_t sl" << L << " = wave_add_u64(pl" << L << ");\n"
             << "      const uint64_t sh" << L << " = wave_add_u64(static_cast<uint64_t>(ph" << L << "));\n"
             << "      const double sf" << L << " = wave_add_f64(pf" << L << ");\n";
          break;
        case L_MIN_F:
        case L_MAX_F: {
          const bool mn = ln.op == L_MIN_F;
          const std::string id = mn ? "~0ULL" : "0ULL";
          const std::string val = ln.is_int ? "static_cast<double>(static_cast<int64_t>(w" + V + "[r]))" : "as_f(w" + V + "[r])";
          o_ << "      uint64_t pm" << L << " = " << id << ";\n"
             << "#pragma unroll\n      for (int r = 0; r < R_; ++r) {\n        const double d = " << val << ";\n"
             << "        if ((" << cond << ") && !isnan(d)) pm" << L << " = " << (mn ? "umin64" : "umax64") << "(pm" << L
             << ", f64_order(d));\n      }\n"
             << "      const uint64_t sm" << L << " = " << (mn ? "wave_min_u64" : "wave_max_u64") << "(pm" << L << ");\n";
          break;
        }
        default:  // L_CNT_VALID
          o_ << "      unsigned long long sc" << L << " = 0;\n"
             << "#pragma unroll\n      for (int r = 0; r < R_; ++r) sc" << L << " += __builtin_popcountll(__ballot("
             << cond << "));\n";
      }
    }
    o_ << "      int slot = -1;\n      if (lds_slots && H != kEmptyKey) {\n"
       << "        int s = -1;\n        if (lane == lead) s = lds_find(lds, STRIDE, lds_slots, lds_shift, H);\n"
       << "        slot = __builtin_amdgcn_readlane(s, lead);\n      }\n"
       << "      if (lane == lead) {\n        if (slot >= 0) {\n          unsigned long long* L = &lds[slot * STRIDE];\n"
       << "          atomicAdd(&L[kW_COUNT], cnt);\n          atomicMin(&L[kW_FIRST], fr);\n";
    for (int l = 0; l < D_.n_lanes; ++l) {
      const Lane& ln = D_.lanes[l];
      const std::string L = std::to_string(l), W = std::to_string(ln.word);
      switch (ln.op) {
        case L_SUM_I:
        case L_SUM_F:
          o_ << "          if (sl" << L << ") atomicAdd(&L[" << W << "], sl" << L << ");\n"
             << "          if (sh" << L << ") atomicAdd(&L[" << W << " + 1], sh" << L << ");\n";
          if (ln.op == L_SUM_F)
            o_ << "          if (sf" << L << " != 0) atomicAdd(reinterpret_cast<double*>(&L[" << W << " + 2]), sf" << L
               << ");\n";
          break;
        case L_MIN_F:
          o_ << "          if (sm" << L << " != ~0ULL) atomicMin(&L[" << W << "], sm" << L << ");\n";
          break;
        case L_MAX_F:
          o_ << "          if (sm" << L << " != 0ULL) atomicMax(&L[" << W << "], sm" << L << ");\n";
          break;
        default:
          o_ << "          if (sc" << L << ") atomicAdd(&L[" << W << "], sc" << L << ");\n";
      }
    }
    o_ << "        } else {\n          const int64_t gs = g_find(G, H, flags);\n          if (gs >= 0) {\n"
       << "            unsigned long long* gp = &G.w[static_cast<uint64_t>(gs) * STRIDE];\n"
       << "            atomicAdd(&gp[kW_COUNT], cnt);\n            atomicMin(&gp[kW_FIRST], fr);\n";
    for (int l = 0; l < D_.n_lanes; ++l) {
      const Lane& ln = D_.lanes[l];
      const std::string L = std::to_string(l), W = std::to_string(ln.word);
      switch (ln.op) {
        case L_SUM_I:
        case L_SUM_F:
          o_ << "            { Split128 sp; sp.lo32 = sl" << L << "; sp.hi32 = static_cast<int64_t>(sh" << L
             << "); uint64_t lo; int64_t hi; sp.to_i128(lo, hi);\n"
             << "              if (lo | static_cast<uint64_t>(hi)) atomic_add_i128(&gp[" << W << "], &gp[" << W
             << " + 1], lo, hi); }\n";
          if (ln.op == L_SUM_F)
            o_ << "            if (sf" << L << " != 0) unsafeAtomicAdd(reinterpret_cast<double*>(&gp[" << W
               << " + 2]), sf" << L << ");\n";
          break;
        case L_MIN_F:
          o_ << "            if (sm" << L << " != ~0ULL) atomicMin(&gp[" << W << "], sm" << L << ");\n";
          break;
        case L_MAX_F:
          o_ << "            if (sm" << L << " != 0ULL) atomicMax(&gp[" << W << "], sm" << L << ");\n";
          break;
        default:
          o_ << "            if (sc" << L << ") atomicAdd(&gp[" << W << "], sc" << L << ");\n";
      }
    }
    o_ << "          }\n        }\n      }\n    }\n    }\n";
  }
};

}  // namespace

std::string jit_groupby_source(const prism_groupby* g, int form, int rows, uint32_t nullable) {
  return Gen(g, form, rows, nullable).run();
}

namespace {
std::mutex g_kmu;
std::unordered_map<std::string, JitKernels> g_kernels;
}  // namespace

JitKernels jit_kernels(const prism_groupby* g, int form, int rows, uint32_t nullable) {
  if (!jit::enabled()) return {};
  const int dev = g->ctx->device;
  // the plan image is zero-filled before it is built, so its bytes name it
  std::string key(reinterpret_cast<const char*>(&g->plan), sizeof(DevPlan));
  key.append(reinterpret_cast<const char*>(g->col_types.data()), g->col_types.size() * sizeof(int32_t));
  const int32_t shape[4] = {dev, form, rows, static_cast<int32_t>(nullable)};
  key.append(reinterpret_cast<const char*>(shape), sizeof(shape));
  std::lock_guard<std::mutex> lock(g_kmu);
  auto it = g_kernels.find(key);
  if (it != g_kernels.end()) return it->second;
  const std::string src = jit_groupby_source(g, form, rows, nullable);
  std::string used_src = src;
  hipFunction_t fn = jit::get(dev, src, kJitGroupbyKernel);
  int regs = 0;
  // Registers just above the budget of one more wave per SIMD: compile again
  // asking for that occupancy, keep it if nothing spills.
  if (fn && hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, fn) == hipSuccess && regs > 0) {
    constexpr int kVgprs = 512, kGranule = 8;
    const int used = (regs + kGranule - 1) / kGranule * kGranule;
    const int waves = kVgprs / used;
    if (waves < 8) {
      const int budget = kVgprs / (waves + 1) / kGranule * kGranule;
      const size_t at = src.find(kJitWavesHint);
      if ((used - budget) * 8 <= used && at != std::string::npos) {
        std::string hinted = src;
        hinted.insert(at + strlen(kJitWavesHint),
                      " __attribute__((amdgpu_waves_per_eu(" + std::to_string(waves + 1) + ")))");
        hipFunction_t f2 = jit::get(dev, hinted, kJitGroupbyKernel);
        int local = 1;
        if (f2 && hipFuncGetAttribute(&local, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f2) == hipSuccess && local == 0) {
          fn = f2;
          used_src = hinted;
        }
      }
    }
  }
  JitKernels k;
  if (fn) {
    k.update = fn;
    k.tiles = jit::get(dev, used_src, kJitGroupbyTiles);
  }
  g_kernels.emplace(std::move(key), k);
  return k;
}

}  // namespace gb
}  // namespace prism
