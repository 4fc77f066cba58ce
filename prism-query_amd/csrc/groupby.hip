#include <hipcub/hipcub.hpp>

#include <memory>
#include <vector>

#include "gb_jit.hpp"
#include "groupby.hpp"
#include "jit.hpp"

namespace prism {
namespace gb {

template <int kRows, int kDepth, int kPref, bool kPriv>
__global__ __launch_bounds__(kBlock) void k_groupby(const DevPlan* __restrict__ P, const prism_column* __restrict__ cols,
                                                    uint64_t first, uint64_t n, uint64_t row_base, GTable G,
                                                    int lds_slots, int lds_shift, uint32_t* flags) {
  extern __shared__ unsigned long long lds[];
  gb_init<kPriv>(P, lds, lds_slots);
  gb_rows<kRows, kDepth, kPref, kPriv>(P, cols, first + static_cast<uint64_t>(blockIdx.x) * kBlock * kRows,
                                       static_cast<uint64_t>(gridDim.x) * kBlock * kRows, n, row_base, G, lds,
                                       lds_slots, lds_shift, flags);
  gb_flush<kPriv>(P, G, lds, lds_slots, flags);
}

// Finish of a table of at most kFinishCap slots in one workgroup: groups in
// first-row order into `rows`, their count into *n_found.
__global__ __launch_bounds__(kFinishThreads) void k_finish_small(GTable G, unsigned long long* rows,
                                                                 unsigned long long* n_found) {
  __shared__ unsigned long long s_first[kFinishCap];
  __shared__ uint32_t s_slot[kFinishCap];
  __shared__ uint32_t s_n;
  const uint32_t n = order_groups(G, rows, s_first, s_slot, &s_n);
  if (threadIdx.x == 0) *n_found = n;
}

__global__ void k_fill_tmpl(unsigned long long* w, uint64_t n_words, const DevPlan* __restrict__ P) {
  const int stride = P->stride;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_words; i += (uint64_t)gridDim.x * blockDim.x)
    w[i] = P->tmpl[i % stride];
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

__global__ void k_clear_counters(unsigned long long* n_groups, uint32_t* flags, unsigned long long* n_found) {
  *n_groups = 0;
  *flags = 0;
  *n_found = 0;
}

// Re-insert every occupied slot of `old` into `G` (keys are unique there).
__global__ void k_rehash(const unsigned long long* __restrict__ old, uint64_t old_cap, GTable G, uint32_t* flags) {
  const int stride = G.stride;
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s <= old_cap; s += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long* src = &old[s * stride];
    if (src[kW_FIRST] == kNoRow) continue;
    int64_t d = (s == old_cap) ? static_cast<int64_t>(G.cap) : g_find(G, src[kW_KEY], flags);
    if (d < 0) continue;
    unsigned long long* dst = &G.w[static_cast<uint64_t>(d) * stride];
    for (int j = 1; j < stride; ++j) dst[j] = src[j];
  }
}

__global__ void k_collect(const unsigned long long* __restrict__ w, uint64_t cap, int stride,
                          unsigned long long* n_found, uint64_t* first_out, uint32_t* slot_out) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s <= cap; s += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t f = w[s * stride + kW_FIRST];
    if (f == kNoRow) continue;
    unsigned long long i = atomicAdd(n_found, 1ULL);
    first_out[i] = f;
    slot_out[i] = static_cast<uint32_t>(s);
  }
}

__global__ void k_gather_slots(const unsigned long long* __restrict__ w, int stride, const uint32_t* __restrict__ order,
                               uint64_t n, unsigned long long* out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n * stride; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t g = i / stride;
    int j = static_cast<int>(i % stride);
    out[i] = w[static_cast<uint64_t>(order[g]) * stride + j];
  }
}

}  // namespace gb
}  // namespace prism

using namespace prism;
using namespace prism::gb;


namespace prism {
namespace gb {

static dim3 fill_grid(prism_ctx* ctx, uint64_t n) {
  uint64_t b = ceil_div(n ? n : 1, 256);
  uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 8;
  return dim3(static_cast<unsigned>(b < cap ? b : cap));
}

static void alloc_table(prism_groupby* g, uint64_t cap) {
  g->table.alloc(g->ctx, (cap + 1) * g->slot_bytes());
  g->cap = cap;
  hipLaunchKernelGGL(k_fill_tmpl, fill_grid(g->ctx, (cap + 1) * g->plan.stride), dim3(256), 0, g->ctx->stream,
                     g->table.as<unsigned long long>(), (cap + 1) * g->plan.stride, g->d_plan.as<DevPlan>());
  PRISM_CHECK_LAUNCH();
}

// Group count and flags (pinned readback: the context lock is held and the
// stream drained before the block is reused).
static void read_counters(prism_groupby* g, uint64_t& groups, uint32_t& flags) {
  struct Counters {
    unsigned long long n;
    uint32_t f;
    uint32_t pad;
  };
  Counters* h = static_cast<Counters*>(readback_block(g->ctx));
  PRISM_HIP(hipMemcpyAsync(h, g->counters.ptr, 16, hipMemcpyDeviceToHost, g->ctx->stream));
  PRISM_HIP(hipStreamSynchronize(g->ctx->stream));
  groups = h->n;
  flags = h->f;
}

// Move the table into one of `new_cap` slots.
static void grow(prism_groupby* g, uint64_t new_cap) {
  DeviceBuffer old = std::move(g->table);
  const uint64_t old_cap = g->cap;
  alloc_table(g, new_cap);
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, g->ctx->stream, g->n_groups_ptr(), 0ULL);
  hipLaunchKernelGGL(k_rehash, fill_grid(g->ctx, old_cap + 1), dim3(256), 0, g->ctx->stream,
                     old.as<unsigned long long>(), old_cap, g->view(), g->flags_ptr());
  PRISM_CHECK_LAUNCH();
  PRISM_HIP(hipStreamSynchronize(g->ctx->stream));
}

template <int kRows, int kDepth, int kPref>
static void launch_form(prism_groupby* g, bool priv, dim3 grid, size_t lds_bytes, const prism_column* cols,
                        uint64_t first, uint64_t n_rows, uint64_t row_base, int lds_slots, int lds_shift) {
  Ctx* c = g->ctx;
  if (priv)
    hipLaunchKernelGGL((k_groupby<kRows, kDepth, kPref, true>), grid, dim3(kBlock), lds_bytes, c->stream,
                       g->d_plan.as<DevPlan>(), cols, first, n_rows, row_base, g->view(), lds_slots, lds_shift,
                       g->flags_ptr());
  else
    hipLaunchKernelGGL((k_groupby<kRows, kDepth, kPref, false>), grid, dim3(kBlock), lds_bytes, c->stream,
                       g->d_plan.as<DevPlan>(), cols, first, n_rows, row_base, g->view(), lds_slots, lds_shift,
                       g->flags_ptr());
}

// Group-count estimate behind the choice of form: what this state has seen
// (also before a reset), else the caller's expectation.
static uint64_t group_estimate(const prism_groupby* g) {
  const uint64_t seen = g->known_groups > g->seen_groups ? g->known_groups : g->seen_groups;
  return seen ? seen : g->expected;
}

// LDS table of the shared form for `known` groups (0 = unknown): slots, hash
// shift and bytes; no LDS table when the groups cannot fit.
size_t shared_slots(const prism_groupby* g, uint64_t known, int& lds_slots, int& lds_shift) {
  const size_t sb = g->slot_bytes();
  const uint64_t want = known ? known : 512;
  uint64_t s = next_pow2(want * 2 < 64 ? 64 : want * 2);
  uint64_t fit = 1;
  while ((fit * 2) * sb <= kLdsBudget) fit *= 2;
  if (s > fit) s = fit;
  lds_slots = 0;
  lds_shift = 0;
  if (want <= s && s >= 16) {
    lds_slots = static_cast<int>(s);
    lds_shift = 64 - log2_pow2(s);
  }
  return static_cast<size_t>(lds_slots) * sb;
}

// Rows [first, n_rows) of the update; bit c of `nullable` = column c has a
// validity bitmap in this update.
static void launch(prism_groupby* g, const prism_column* cols, uint64_t first, uint64_t n_rows, uint64_t row_base,
                   uint32_t nullable) {
  Ctx* c = g->ctx;
  const uint64_t known = group_estimate(g);
  // Private form when the group count is known to be small.
  const uint64_t pw = static_cast<uint64_t>(g->plan.stride - kW_COUNT);
  const uint64_t ps = next_pow2(known > 2 ? known : 2);
  const size_t priv_bytes = (2 * ps + ps * pw * kBlock) * 8;
  const bool priv = known > 0 && ps <= 16 && priv_bytes <= kPrivBudget;
  int lds_slots = 0, lds_shift = 0;
  size_t lds_bytes = 0;
  if (priv) {
    lds_slots = static_cast<int>(ps);
    lds_shift = 64 - log2_pow2(ps);
    lds_bytes = priv_bytes;
  } else {
    lds_bytes = shared_slots(g, known, lds_slots, lds_shift);
  }
  // Generated kernel: the shared-form LDS table; wave pre-aggregation when
  // every row falls in one group.
  {
    const int form = (g->plan.n_keys == 0 || known == 1) ? kJitWave : kJitShared;
    const int rows = g->plan.n_pref > 6 ? 2 : 4;
    hipFunction_t fn = jit_kernels(g, form, rows, nullable).update;
    if (fn) {
      int slots = 0, shift = 0;
      const size_t bytes = shared_slots(g, known, slots, shift);
      int per_cu = 0;
      PRISM_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, bytes));
      if (per_cu < 1) per_cu = 1;
      uint64_t blocks = ceil_div(n_rows - first, static_cast<uint64_t>(kBlock) * rows);
      const uint64_t max_blocks = static_cast<uint64_t>(c->num_cus) * per_cu;
      if (blocks > max_blocks) blocks = max_blocks;
      if (blocks == 0) blocks = 1;
      const DevPlan* P = g->d_plan.as<DevPlan>();
      GTable G = g->view();
      uint32_t* flags = g->flags_ptr();
      void* args[] = {&P, &cols, &first, &n_rows, &row_base, &G, &slots, &shift, &flags};
      KernelTimer kt(c, "groupby_update");
      PRISM_HIP(hipModuleLaunchKernel(fn, static_cast<unsigned>(blocks), 1, 1, kBlock, 1, 1,
                                      static_cast<unsigned>(bytes), c->stream, args, nullptr));
      return;
    }
  }
  const bool shallow = g->max_depth <= 4;
  const uint64_t rows_per_lane = shallow ? 4 : 2;
  uint64_t blocks = ceil_div(n_rows - first, static_cast<uint64_t>(kBlock) * rows_per_lane);
  const uint64_t per_cu = lds_bytes ? (160 * 1024 / lds_bytes < 8 ? 160 * 1024 / lds_bytes : 8) : 8;
  uint64_t max_blocks = static_cast<uint64_t>(c->num_cus) * (per_cu ? per_cu : 1);
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks == 0) blocks = 1;
  const dim3 grid(static_cast<unsigned>(blocks));
  KernelTimer kt(c, "groupby_update");
  if (shallow) {
    if (g->plan.n_pref <= 4)
      launch_form<4, 4, 4>(g, priv, grid, lds_bytes, cols, first, n_rows, row_base, lds_slots, lds_shift);
    else
      launch_form<4, 4, 8>(g, priv, grid, lds_bytes, cols, first, n_rows, row_base, lds_slots, lds_shift);
  } else {
    launch_form<2, kStackDepth, 8>(g, priv, grid, lds_bytes, cols, first, n_rows, row_base, lds_slots, lds_shift);
  }
  PRISM_CHECK_LAUNCH();
}

}  // namespace gb
}  // namespace prism

namespace prism {
namespace gb {

// Checks `plan` and fills the host side of `g` (device plan image, column
// types, aggregate lanes); touches no device.
static void build_plan(prism_groupby* g, const prism_groupby_plan* plan) {
  if (plan->n_columns > static_cast<uint32_t>(kMaxCols))
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: at most 16 input columns");
  if (plan->n_keys > static_cast<uint32_t>(kMaxKeys))
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: at most 8 group-by columns");
  if (plan->n_values > static_cast<uint32_t>(kMaxValues))
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: at most 16 value programs");
  if (plan->n_columns && !plan->column_types) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL column types");
  g->col_types.assign(plan->column_types, plan->column_types + plan->n_columns);
  DevPlan& D = g->plan;
  memset(&D, 0, sizeof(D));
  build_str_table(plan->strings, plan->n_strings, D.strs);

  D.n_keys = static_cast<int32_t>(plan->n_keys);
  for (uint32_t k = 0; k < plan->n_keys; ++k) {
    uint32_t col = plan->key_columns[k];
    if (col >= plan->n_columns) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: group-by column out of range");
    D.key_cols[k] = static_cast<int32_t>(col);
  }
  int32_t code_len = 0;
  auto append = [&](const prism_instr* code, uint32_t len) {
    if (code_len + static_cast<int64_t>(len) > kMaxCode)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: row programs exceed 160 instructions");
    memcpy(&D.code[code_len], code, len * sizeof(prism_instr));
    int32_t start = code_len;
    code_len += static_cast<int32_t>(len);
    return start;
  };
  if (plan->filter && plan->n_filter) {
    int depth = 1;
    ValType t = check_program(plan->filter, plan->n_filter, plan->column_types, plan->n_columns,
                              plan->n_strings, &depth);
    g->max_depth = depth > g->max_depth ? depth : g->max_depth;
    if (t != VT_B && t != VT_NULL) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: filter is not boolean");
    D.filter_start = append(plan->filter, plan->n_filter);
    D.filter_len = static_cast<int32_t>(plan->n_filter);
  }
  D.n_values = static_cast<int32_t>(plan->n_values);
  uint32_t off = 0;
  for (uint32_t v = 0; v < plan->n_values; ++v) {
    uint32_t len = plan->value_lengths[v];
    int depth = 1;
    ValType t = check_program(plan->values + off, len, plan->column_types, plan->n_columns, plan->n_strings,
                              &depth);
    g->max_depth = depth > g->max_depth ? depth : g->max_depth;
    g->value_types.push_back(t);
    D.value_start[v] = append(plan->values + off, len);
    D.value_len[v] = static_cast<int32_t>(len);
    off += len;
  }
  // Prefetch slots: fixed-width columns read by the programs or hashed as
  // keys; their loads become kOpLoadPre of the slot.
  auto fixed = [&](uint32_t col) { return plan->column_types[col] != PRISM_TYPE_UTF8; };
  auto pref_slot = [&](uint32_t col) -> int32_t {
    for (int32_t j = 0; j < D.n_pref; ++j)
      if (D.pref_col[j] == static_cast<int32_t>(col)) return j;
    if (D.n_pref >= kMaxPref) return -1;
    D.pref_col[D.n_pref] = static_cast<int32_t>(col);
    return D.n_pref++;
  };
  for (int32_t pc = 0; pc < code_len; ++pc) {
    prism_instr& ins = D.code[pc];
    if ((ins.op == PRISM_OP_LOAD_I || ins.op == PRISM_OP_LOAD_F) && fixed(static_cast<uint32_t>(ins.arg))) {
      const int32_t j = pref_slot(static_cast<uint32_t>(ins.arg));
      if (j >= 0) {
        ins.op = kOpLoadPre;
        ins.arg = j;
      }
    }
  }
  for (int32_t k = 0; k < D.n_keys; ++k)
    D.key_pref[k] = fixed(static_cast<uint32_t>(D.key_cols[k])) ? pref_slot(static_cast<uint32_t>(D.key_cols[k])) : -1;
  const bool global = plan->n_keys == 0;
  D.skip_nulls = global ? 1 : 0;
  int32_t word = kW_LANES;
  D.tmpl[kW_KEY] = kEmptyKey;
  D.tmpl[kW_FIRST] = kNoRow;
  D.tmpl[kW_COUNT] = 0;
  auto add_lane = [&](int32_t op, int32_t value, int32_t is_int) {
    for (int32_t i = 0; i < D.n_lanes; ++i)  // SUM and AVG of one value share a lane
      if (D.lanes[i].op == op && D.lanes[i].value == value) return i;
    if (D.n_lanes >= kMaxLanes) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: too many aggregates");
    Lane& L = D.lanes[D.n_lanes];
    L.op = op;
    L.value = value;
    L.word = word;
    L.is_int = is_int;
    int words = op == L_SUM_I ? 2 : op == L_SUM_F ? 3 : 1;
    for (int i = 0; i < words; ++i) D.tmpl[word + i] = 0;
    if (op == L_MIN_F) D.tmpl[word] = f64_order(INFINITY);
    if (op == L_MAX_F) D.tmpl[word] = f64_order(-INFINITY);
    word += words;
    return D.n_lanes++;
  };
  for (uint32_t a = 0; a < plan->n_aggs; ++a) {
    const prism_groupby_agg& A = plan->aggs[a];
    g->aggs.push_back(A);
    int32_t ls = -1, lm = -1, lc = -1;
    if (A.func == PRISM_AGG_COUNT_DISTINCT)
      fail(PRISM_ERR_UNSUPPORTED_AGGREGATION,
           "Aggregation error: unsupported function COUNT DISTINCT in the device aggregation");
    if (A.func < PRISM_AGG_SUM || A.func > PRISM_AGG_MAX)
      fail(PRISM_ERR_UNSUPPORTED_AGGREGATION, "Aggregation error: unsupported function " + std::to_string(A.func));
    if ((A.value < 0 || static_cast<uint32_t>(A.value) >= plan->n_values) && A.func != PRISM_AGG_COUNT)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: aggregate value index out of range");
    if (A.value >= 0 && static_cast<uint32_t>(A.value) >= plan->n_values)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: aggregate value index out of range");
    int32_t vt = (A.value >= 0 && static_cast<uint32_t>(A.value) < plan->n_values) ? g->value_types[A.value] : VT_I;
    if (vt == VT_NULL) vt = VT_I;
    bool is_int = vt != VT_F;
    switch (A.func) {
      case PRISM_AGG_SUM:
      case PRISM_AGG_AVG:
        ls = add_lane(is_int ? L_SUM_I : L_SUM_F, A.value, is_int);
        break;
      case PRISM_AGG_MIN:
        lm = add_lane(L_MIN_F, A.value, is_int);
        break;
      case PRISM_AGG_MAX:
        lm = add_lane(L_MAX_F, A.value, is_int);
        break;
      default:
        break;
    }
    if (global && A.func != PRISM_AGG_SUM && A.func != PRISM_AGG_MIN && A.func != PRISM_AGG_MAX && A.value >= 0)
      lc = add_lane(L_CNT_VALID, A.value, is_int);
    g->agg_lane_sum.push_back(ls);
    g->agg_lane_mm.push_back(lm);
    g->agg_lane_cnt.push_back(lc);
  }
  D.stride = word;
}

}  // namespace gb
}  // namespace prism

extern "C" {

prism_status prism_groupby_create(prism_ctx* ctx, const prism_groupby_plan* plan, uint64_t expected_groups,
                                  prism_groupby** out) {
  return guarded([&] {
    if (!ctx || !plan || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/plan/output");
    *out = nullptr;
    auto g = std::make_unique<prism_groupby>();
    build_plan(g.get(), plan);
    use_device(ctx);
    g->ctx = as_ctx(ctx);
    g->expected = expected_groups;
    const DevPlan& D = g->plan;
    // Lanes are visited per value program; keep the order stable.
    g->d_plan.alloc(g->ctx, sizeof(DevPlan));
    g->counters.alloc(g->ctx, 256);
    PRISM_HIP(hipMemcpyAsync(g->d_plan.ptr, &D, sizeof(DevPlan), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(1), 0, ctx->stream, g->n_groups_ptr(), g->flags_ptr(),
                       g->n_found_ptr());
    PRISM_CHECK_LAUNCH();
    uint64_t cap = next_pow2(expected_groups * 2 < 1024 ? 1024 : expected_groups * 2);
    alloc_table(g.get(), cap);
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
    *out = g.release();
  });
}

prism_status prism_groupby_jit_source(const prism_groupby_plan* plan, int32_t form, uint32_t nullable,
                                      int32_t compile, char* buf, uint64_t cap, uint64_t* len) {
  return guarded([&] {
    if (!plan || !len) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL plan/output");
    if (form != kJitWave && form != kJitShared) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown form");
    prism_groupby g;
    build_plan(&g, plan);
    const std::string src = jit_groupby_source(&g, form, g.plan.n_pref > 6 ? 2 : 4, nullable);
    *len = src.size();
    if (buf && cap) {
      const size_t n = src.size() < cap - 1 ? src.size() : cap - 1;
      memcpy(buf, src.data(), n);
      buf[n] = '\0';
    }
    if (compile) {
      std::string code, err;
      if (!jit::compile_code(src, "gfx950", code, err)) fail(PRISM_ERR_EXECUTION, "Execution error: " + err);
    }
  });
}

prism_status prism_groupby_update(prism_groupby* g, const prism_column* columns, uint64_t n_rows, uint64_t row_base) {
  return guarded([&] {
    if (!g) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL group-by state");
    if (n_rows == 0) return;
    if (!columns && !g->col_types.empty()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL columns");
    use_device(g->ctx);
    std::lock_guard<std::mutex> lock(g->ctx->mu);
    g->finished = false;
    ColsArg cols;
    memset(&cols, 0, sizeof(cols));
    uint32_t nullable = 0;
    for (size_t i = 0; i < g->col_types.size(); ++i) {
      if (columns[i].type != g->col_types[i])
        fail(PRISM_ERR_SCHEMA_MISMATCH, "Schema mismatch: column " + std::to_string(i) + " type differs from the plan");
      if (!columns[i].data && n_rows) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL column data");
      if (columns[i].type == PRISM_TYPE_UTF8 && !columns[i].offsets)
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: Utf8 column without offsets");
      cols.c[i] = columns[i];
      if (columns[i].validity) nullable |= 1u << i;
    }
    hipStream_t st = g->ctx->stream;
    g->d_cols.reserve(g->ctx, sizeof(ColsArg));
    // column descriptors go up from the pinned readback block (second KiB);
    // every update drains the stream before it returns
    static_assert(sizeof(ColsArg) <= 3072, "column descriptors fit the readback block");
    char* pin = static_cast<char*>(readback_block(g->ctx)) + 1024;
    memcpy(pin, &cols, sizeof(ColsArg));
    PRISM_HIP(hipMemcpyAsync(g->d_cols.ptr, pin, sizeof(ColsArg), hipMemcpyHostToDevice, st));
    const prism_column* dcols = g->d_cols.as<prism_column>();
    uint64_t done = 0;
    while (done < n_rows) {
      // With no group count to go by, a prefix of the rows measures it first.
      uint64_t end = n_rows;
      if (group_estimate(g) == 0 && n_rows - done > 2 * kSampleRows) end = done + kSampleRows;
      // A snapshot is needed only when this segment could overfill the table.
      const bool risky = (g->known_groups + (end - done)) * 4 > g->cap * 3;
      if (risky) {
        g->backup.reserve(g->ctx, g->table.bytes);
        PRISM_HIP(hipMemcpyAsync(g->backup.ptr, g->table.ptr, g->table.bytes, hipMemcpyDeviceToDevice, st));
      }
      launch(g, dcols, done, end, row_base, nullable);
      uint64_t groups;
      uint32_t flags;
      read_counters(g, groups, flags);
      if (flags & kFlagOverflow)
        fail(PRISM_ERR_EXECUTION, "Execution error: int64 overflow in exact integer arithmetic");
      if (flags & kFlagDropped) {
        if (!risky) fail(PRISM_ERR_INTERNAL, "Internal error: group table overflow without snapshot");
        PRISM_HIP(hipMemcpyAsync(g->table.ptr, g->backup.ptr, g->table.bytes, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, g->flags_ptr(), 0u);
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, g->n_groups_ptr(), static_cast<unsigned long long>(g->known_groups));
        PRISM_HIP(hipStreamSynchronize(st));
        uint64_t target = next_pow2((groups > g->known_groups ? groups : g->known_groups) * 4);
        grow(g, target > g->cap * 4 ? target : g->cap * 4);
        read_counters(g, g->known_groups, flags);
        continue;
      }
      g->known_groups = groups;
      if (flags & kFlagGrow) {
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, g->flags_ptr(), 0u);
        grow(g, g->cap * 4);
        read_counters(g, g->known_groups, flags);
      }
      if (g->known_groups > g->seen_groups) g->seen_groups = g->known_groups;
      done = end;
    }
  });
}

prism_status prism_groupby_finish(prism_groupby* g, uint64_t* n_groups) {
  return guarded([&] {
    if (!g || !n_groups) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL group-by state/output");
    use_device(g->ctx);
    std::lock_guard<std::mutex> lock(g->ctx->mu);
    hipStream_t st = g->ctx->stream;
    Ctx* c = g->ctx;
    const uint64_t slots = g->cap + 1;
    if (slots <= static_cast<uint64_t>(kFinishCap)) {  // one workgroup sorts in LDS; one read of the count
      g->result.reserve(c, slots * g->slot_bytes());
      hipLaunchKernelGGL(k_finish_small, dim3(1), dim3(kFinishThreads), 0, st, g->view(),
                         g->result.as<unsigned long long>(), g->n_found_ptr());
      PRISM_CHECK_LAUNCH();
      unsigned long long* h = static_cast<unsigned long long*>(readback_block(c));
      PRISM_HIP(hipMemcpyAsync(h, g->n_found_ptr(), 8, hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      g->n_result = *h;
      g->finished = true;
      *n_groups = *h;
      return;
    }
    DeviceBuffer first(c, slots * 8), first_sorted(c, slots * 8), idx(c, slots * 4), idx_sorted(c, slots * 4);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, g->n_found_ptr(), 0ULL);
    hipLaunchKernelGGL(k_collect, fill_grid(g->ctx, slots), dim3(256), 0, st, g->table.as<unsigned long long>(),
                       g->cap, g->plan.stride, g->n_found_ptr(), first.as<uint64_t>(), idx.as<uint32_t>());
    PRISM_CHECK_LAUNCH();
    unsigned long long found = 0;
    PRISM_HIP(hipMemcpyAsync(&found, g->n_found_ptr(), 8, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    if (found > 1) {
      size_t tmp_bytes = 0;
      PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, first.as<uint64_t>(), first_sorted.as<uint64_t>(),
                                                   idx.as<uint32_t>(), idx_sorted.as<uint32_t>(),
                                                   static_cast<int>(found), 0, 64, st));
      DeviceBuffer tmp(c, tmp_bytes);
      PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, first.as<uint64_t>(), first_sorted.as<uint64_t>(),
                                                   idx.as<uint32_t>(), idx_sorted.as<uint32_t>(),
                                                   static_cast<int>(found), 0, 64, st));
    } else if (found == 1) {
      PRISM_HIP(hipMemcpyAsync(idx_sorted.ptr, idx.ptr, 4, hipMemcpyDeviceToDevice, st));
    }
    g->result.reserve(c, (found ? found : 1) * g->slot_bytes());
    if (found)
      hipLaunchKernelGGL(k_gather_slots, fill_grid(g->ctx, found * g->plan.stride), dim3(256), 0, st,
                         g->table.as<unsigned long long>(), g->plan.stride, idx_sorted.as<uint32_t>(), found,
                         g->result.as<unsigned long long>());
    PRISM_CHECK_LAUNCH();
    PRISM_HIP(hipStreamSynchronize(st));
    g->n_result = found;
    g->finished = true;
    *n_groups = found;
  });
}

static double i128_to_f64(uint64_t lo, int64_t hi) {
  __int128 v = static_cast<__int128>((static_cast<unsigned __int128>(static_cast<uint64_t>(hi)) << 64) | lo);
  return static_cast<double>(v);  // correctly rounded (round to nearest even)
}

}  // extern "C"

namespace prism {
namespace gb {

void convert_rows(const prism_groupby* g, const unsigned long long* rows, uint64_t n, const prism_groupby_output* out) {
  const int stride = g->plan.stride;
  const DevPlan& D = g->plan;
  for (uint64_t i = 0; i < n; ++i) {
    const unsigned long long* s = &rows[i * stride];
    if (out->first_row) out->first_row[i] = s[kW_FIRST];
    if (out->key_hash) out->key_hash[i] = s[kW_KEY];
    if (out->row_count) out->row_count[i] = s[kW_COUNT];
    for (size_t a = 0; a < g->aggs.size(); ++a) {
      const int func = g->aggs[a].func;
      double sum_f = 0.0;
      uint64_t slo = 0;
      int64_t shi = 0;
      if (g->agg_lane_sum[a] >= 0) {
        const Lane& L = D.lanes[g->agg_lane_sum[a]];
        slo = s[L.word];
        shi = static_cast<int64_t>(s[L.word + 1]);
        sum_f = i128_to_f64(slo, shi);
        if (L.op == L_SUM_F) {
          double fpart;
          memcpy(&fpart, &s[L.word + 2], 8);
          sum_f = (slo == 0 && shi == 0) ? fpart : sum_f + fpart;
        }
      }
      const uint64_t count = g->agg_lane_cnt[a] >= 0 ? s[D.lanes[g->agg_lane_cnt[a]].word] : s[kW_COUNT];
      double f = 0.0;
      switch (func) {
        case PRISM_AGG_SUM: f = sum_f; break;
        case PRISM_AGG_COUNT: f = static_cast<double>(count); break;
        case PRISM_AGG_AVG: f = count ? sum_f / static_cast<double>(count) : 0.0; break;
        case PRISM_AGG_MIN:
        case PRISM_AGG_MAX: {
          const Lane& L = D.lanes[g->agg_lane_mm[a]];
          f = f64_unorder(s[L.word]);
          if (D.skip_nulls) {  // global_aggregate_batches: an empty fold gives 0.0 (hash_aggregate.rs:401-413)
            if (func == PRISM_AGG_MIN && f == INFINITY) f = 0.0;
            if (func == PRISM_AGG_MAX && f == -INFINITY) f = 0.0;
          } else if (s[kW_COUNT] == 0) {
            f = 0.0;
          }
          break;
        }
      }
      if (out->agg_f64 && out->agg_f64[a]) out->agg_f64[a][i] = f;
      if (out->agg_i64 && out->agg_i64[a]) out->agg_i64[a][i] = static_cast<int64_t>(count);
      if (out->agg_sum128 && out->agg_sum128[a]) {
        out->agg_sum128[a][2 * i] = slo;
        out->agg_sum128[a][2 * i + 1] = static_cast<uint64_t>(shi);
      }
    }
  }
}

}  // namespace gb
}  // namespace prism

extern "C" {

prism_status prism_groupby_read(prism_groupby* g, const prism_groupby_output* out) {
  return guarded([&] {
    if (!g || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL group-by state/output");
    if (!g->finished) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: call prism_groupby_finish first");
    use_device(g->ctx);
    const int stride = g->plan.stride;
    const uint64_t n = g->n_result;
    std::vector<unsigned long long> h(n * stride);
    if (n) {
      PRISM_HIP(hipMemcpyAsync(h.data(), g->result.ptr, n * stride * 8, hipMemcpyDeviceToHost, g->ctx->stream));
      PRISM_HIP(hipStreamSynchronize(g->ctx->stream));
    }
    convert_rows(g, h.data(), n, out);
  });
}

prism_status prism_groupby_reset(prism_groupby* g) {
  return guarded([&] {
    if (!g) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL group-by state");
    use_device(g->ctx);
    std::lock_guard<std::mutex> lock(g->ctx->mu);
    hipLaunchKernelGGL(k_fill_tmpl, fill_grid(g->ctx, (g->cap + 1) * g->plan.stride), dim3(256), 0, g->ctx->stream,
                       g->table.as<unsigned long long>(), (g->cap + 1) * g->plan.stride, g->d_plan.as<DevPlan>());
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(1), 0, g->ctx->stream, g->n_groups_ptr(), g->flags_ptr(),
                       g->n_found_ptr());
    PRISM_CHECK_LAUNCH();
    g->known_groups = 0;
    g->finished = false;
    g->n_result = 0;
  });
}

void prism_groupby_destroy(prism_groupby* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
}

}  // extern "C"
