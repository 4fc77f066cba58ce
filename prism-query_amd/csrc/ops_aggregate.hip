// Arrow-level hash aggregation: hash_aggregate / hash_aggregate_batches
// (native/prism-executor/src/hash_aggregate.rs:174-527) over the Arrow C Data
// Interface, running the device group-by (groupby.hip) and, for COUNT
// DISTINCT, a sort + reduce-by-key over (group id, value) pairs in HBM.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gather.hpp"
#include "program.hpp"

namespace prism {
namespace aagg {

constexpr int kB = 256;

static dim3 agrid(prism_ctx* ctx, uint64_t n) {
  uint64_t b = ceil_div(n ? n : 1, kB);
  uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 16;
  return dim3(static_cast<unsigned>(b < cap ? b : cap));
}

struct KeyArg {
  prism_column c[kMaxKeys];
  int n;
};

__global__ void k_fill_u64_zero(unsigned long long* p) { *p = 0; }

__device__ __forceinline__ uint64_t smix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// (group id, DistinctValue word, source row) for every non-null value
// (hash_aggregate.rs:457-481): Float64 folds -0.0 into 0.0 and every NaN into
// one NaN; a Utf8 value is represented by a 64-bit fingerprint of its bytes
// (masked by fp_mask), and runs of equal fingerprints are checked byte by
// byte afterwards.  Source row = batch << 40 | row.
__global__ void k_distinct_pairs(KeyArg K, prism_column v, uint64_t n, uint64_t batch, uint64_t fp_mask,
                                 unsigned long long* __restrict__ cursor, uint64_t* __restrict__ v_out,
                                 ulonglong2* __restrict__ gi_out) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    if (!col_valid(v, r)) continue;
    uint64_t h = 0;
    for (int k = 0; k < K.n; ++k) h = hash_key(K.c[k], r, h);
    const uint64_t i = v.offset + r;
    uint64_t w;
    switch (v.type) {
      case PRISM_TYPE_INT64: w = static_cast<uint64_t>(static_cast<const int64_t*>(v.data)[i]); break;
      case PRISM_TYPE_FLOAT64: {
        double d = static_cast<const double*>(v.data)[i];
        if (d == 0.0) d = 0.0;
        w = isnan(d) ? 0x7ff8000000000000ULL : static_cast<uint64_t>(__double_as_longlong(d));
        break;
      }
      case PRISM_TYPE_BOOL: w = bit_at(static_cast<const uint8_t*>(v.data), i); break;
      case PRISM_TYPE_UTF8: {
        const int32_t b = v.offsets[i], e = v.offsets[i + 1];
        const uint8_t* s = static_cast<const uint8_t*>(v.data) + b;
        uint64_t x = smix(static_cast<uint64_t>(e - b) + 0x51ED27ULL);
        for (int32_t p = 0; p < e - b; ++p) x = smix(x ^ (static_cast<uint64_t>(s[p]) << ((p & 7) * 8)) ^ p);
        w = x & fp_mask;
        break;
      }
      default: w = static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int32_t*>(v.data)[i]));
    }
    unsigned long long p = atomicAdd(cursor, 1ULL);
    v_out[p] = w;
    gi_out[p] = make_ulonglong2(h, (batch << 40) | r);
  }
}

// Between the two sort passes: group ids become the keys, (value, row) the payload.
__global__ void k_distinct_regroup(const uint64_t* __restrict__ v, const ulonglong2* __restrict__ gi, uint64_t n,
                                   uint64_t* __restrict__ g_out, ulonglong2* __restrict__ vi_out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const ulonglong2 x = gi[i];
    g_out[i] = x.x;
    vi_out[i] = make_ulonglong2(v[i], x.y);
  }
}

// Byte equality of two Utf8 values named by source row (batch << 40 | row);
// cols[b] = the device descriptor of batch b's value column.
__device__ __forceinline__ bool str_eq(const prism_column* __restrict__ cols, uint64_t a, uint64_t b) {
  const prism_column& A = cols[a >> 40];
  const prism_column& B = cols[b >> 40];
  const uint64_t ra = A.offset + (a & ((1ULL << 40) - 1)), rb = B.offset + (b & ((1ULL << 40) - 1));
  const int32_t a0 = A.offsets[ra], a1 = A.offsets[ra + 1], b0 = B.offsets[rb], b1 = B.offsets[rb + 1];
  if (a1 - a0 != b1 - b0) return false;
  const uint8_t* x = static_cast<const uint8_t*>(A.data) + a0;
  const uint8_t* y = static_cast<const uint8_t*>(B.data) + b0;
  for (int32_t k = 0; k < a1 - a0; ++k)
    if (x[k] != y[k]) return false;
  return true;
}

// flag[i] = 1 at the first element of each run of equal (group, word).  For
// Utf8 (check_bytes), neighbours inside a run whose bytes differ are a
// fingerprint collision: *collisions counts them and the host recounts the
// affected runs exactly.
__global__ void k_first_of_run(const uint64_t* __restrict__ g, const ulonglong2* __restrict__ vi, uint64_t n,
                               const prism_column* __restrict__ str_cols, uint32_t* __restrict__ flag,
                               unsigned long long* __restrict__ collisions) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const bool first = i == 0 || g[i] != g[i - 1] || vi[i].x != vi[i - 1].x;
    flag[i] = first ? 1u : 0u;
    if (!first && str_cols && !str_eq(str_cols, vi[i - 1].y, vi[i].y)) atomicAdd(collisions, 1ULL);
  }
}

static uint64_t distinct_fp_mask() {
  // test hook, read at each call: PRISM_DISTINCT_FP_BITS=b keeps b bits of the
  // Utf8 fingerprint, so that tests can force collisions through the exact recount
  const char* e = getenv("PRISM_DISTINCT_FP_BITS");
  const int b = e ? atoi(e) : 64;
  return b <= 0 ? 0ULL : b >= 64 ? ~0ULL : (1ULL << b) - 1;
}

static std::string host_str(const HostColumn& c, int64_t row) {
  const int64_t i = c.offset + row;
  const int32_t b = c.offsets[i], e = c.offsets[i + 1];
  return std::string(static_cast<const char*>(c.data) + b, static_cast<size_t>(e - b));
}

// Distinct counts per group id for one aggregate column across all batches:
// sort (group, word) pairs, count runs per group.
static std::unordered_map<uint64_t, int64_t> count_distinct(prism_ctx* ctx, const std::vector<HostBatch>& batches,
                                                            const std::vector<uint32_t>& group_by, uint32_t column) {
  Ctx* c = as_ctx(ctx);
  hipStream_t st = ctx->stream;
  uint64_t total = 0;
  for (const HostBatch& b : batches) total += static_cast<uint64_t>(b.n_rows);
  const bool utf8 = !batches.empty() && batches[0].cols[column].type == DType::Utf8;
  DeviceBuffer v(c, (total ? total : 1) * 8), gi(c, (total ? total : 1) * 16), cur(c, 8);
  hipLaunchKernelGGL(k_fill_u64_zero, dim3(1), dim3(1), 0, st, cur.as<unsigned long long>());
  std::vector<DevColumn> vcols(batches.size());  // Utf8: kept for the byte check
  std::vector<prism_column> str_desc(batches.size());
  memset(str_desc.data(), 0, str_desc.size() * sizeof(prism_column));
  const uint64_t mask = utf8 ? distinct_fp_mask() : ~0ULL;
  for (size_t bi = 0; bi < batches.size(); ++bi) {
    const HostBatch& b = batches[bi];
    if (b.n_rows == 0) continue;
    std::vector<DevColumn> keys;
    KeyArg K;
    memset(&K, 0, sizeof(K));
    K.n = static_cast<int>(group_by.size());
    for (size_t k = 0; k < group_by.size(); ++k) {
      keys.push_back(upload(ctx, b.cols[group_by[k]]));
      K.c[k] = keys.back().desc();
    }
    vcols[bi] = upload(ctx, b.cols[column]);
    if (utf8) str_desc[bi] = vcols[bi].desc();
    hipLaunchKernelGGL(k_distinct_pairs, agrid(ctx, b.n_rows), dim3(kB), 0, st, K, vcols[bi].desc(),
                       static_cast<uint64_t>(b.n_rows), static_cast<uint64_t>(bi), mask,
                       cur.as<unsigned long long>(), v.as<uint64_t>(), gi.as<ulonglong2>());
    PRISM_CHECK_LAUNCH();
    PRISM_HIP(hipStreamSynchronize(st));
    if (!utf8) vcols[bi] = DevColumn();
  }
  unsigned long long n = 0;
  PRISM_HIP(hipMemcpyAsync(&n, cur.ptr, 8, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipStreamSynchronize(st));
  std::unordered_map<uint64_t, int64_t> out;
  if (n == 0) return out;
  // LSD order: by word, then stably by group -> sorted by (group, word),
  // rows ascending inside a run.
  DeviceBuffer v2(c, n * 8), gi2(c, n * 16), g3(c, n * 8), g4(c, n * 8), vi3(c, n * 16), vi4(c, n * 16);
  DeviceBuffer flag(c, n * 4), ug(c, n * 8), uc(c, n * 8), nr(c, 8), coll(c, 8);
  size_t t1 = 0, t2 = 0, t3 = 0;
  const int ni = static_cast<int>(n);
  PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, v.as<uint64_t>(), v2.as<uint64_t>(), gi.as<ulonglong2>(),
                                               gi2.as<ulonglong2>(), ni, 0, 64, st));
  PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, g3.as<uint64_t>(), g4.as<uint64_t>(),
                                               vi3.as<ulonglong2>(), vi4.as<ulonglong2>(), ni, 0, 64, st));
  PRISM_HIP(hipcub::DeviceReduce::ReduceByKey(nullptr, t3, g4.as<uint64_t>(), ug.as<uint64_t>(), flag.as<uint32_t>(),
                                              uc.as<uint64_t>(), nr.as<uint64_t>(), hipcub::Sum(), ni, st));
  size_t tmax = t1 > t2 ? t1 : t2;
  tmax = tmax > t3 ? tmax : t3;
  DeviceBuffer tmp(c, tmax);
  PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, t1, v.as<uint64_t>(), v2.as<uint64_t>(), gi.as<ulonglong2>(),
                                               gi2.as<ulonglong2>(), ni, 0, 64, st));
  hipLaunchKernelGGL(k_distinct_regroup, agrid(ctx, n), dim3(kB), 0, st, v2.as<uint64_t>(), gi2.as<ulonglong2>(), n,
                     g3.as<uint64_t>(), vi3.as<ulonglong2>());
  PRISM_CHECK_LAUNCH();
  PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, t2, g3.as<uint64_t>(), g4.as<uint64_t>(),
                                               vi3.as<ulonglong2>(), vi4.as<ulonglong2>(), ni, 0, 64, st));
  hipLaunchKernelGGL(k_fill_u64_zero, dim3(1), dim3(1), 0, st, coll.as<unsigned long long>());
  DeviceBuffer d_str(c, str_desc.size() * sizeof(prism_column));
  if (utf8) {
    PRISM_HIP(hipMemcpyAsync(d_str.ptr, str_desc.data(), str_desc.size() * sizeof(prism_column),
                             hipMemcpyHostToDevice, st));
    PRISM_HIP(hipStreamSynchronize(st));
  }
  hipLaunchKernelGGL(k_first_of_run, agrid(ctx, n), dim3(kB), 0, st, g4.as<uint64_t>(), vi4.as<ulonglong2>(), n,
                     utf8 ? d_str.as<prism_column>() : nullptr, flag.as<uint32_t>(), coll.as<unsigned long long>());
  PRISM_CHECK_LAUNCH();
  PRISM_HIP(hipcub::DeviceReduce::ReduceByKey(tmp.ptr, t3, g4.as<uint64_t>(), ug.as<uint64_t>(), flag.as<uint32_t>(),
                                              uc.as<uint64_t>(), nr.as<uint64_t>(), hipcub::Sum(), ni, st));
  uint64_t runs = 0;
  unsigned long long collisions = 0;
  PRISM_HIP(hipMemcpyAsync(&runs, nr.ptr, 8, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipMemcpyAsync(&collisions, coll.ptr, 8, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipStreamSynchronize(st));
  std::vector<uint64_t> hg(runs), hc(runs);
  PRISM_HIP(hipMemcpyAsync(hg.data(), ug.ptr, runs * 8, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipMemcpyAsync(hc.data(), uc.ptr, runs * 8, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipStreamSynchronize(st));
  for (uint64_t i = 0; i < runs; ++i) out[hg[i]] = static_cast<int64_t>(hc[i]);
  if (collisions) {
    // Exact recount of the runs whose fingerprints collide: each counted
    // once above; add the further distinct strings it holds.
    std::vector<uint64_t> hgr(n);
    std::vector<ulonglong2> hvi(n);
    PRISM_HIP(hipMemcpyAsync(hgr.data(), g4.ptr, n * 8, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipMemcpyAsync(hvi.data(), vi4.ptr, n * 16, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    auto str_of = [&](uint64_t id) {
      const HostBatch& b = batches[id >> 40];
      return host_str(b.cols[column], static_cast<int64_t>(id & ((1ULL << 40) - 1)));
    };
    for (uint64_t i = 0; i < n;) {
      uint64_t j = i + 1;
      while (j < n && hgr[j] == hgr[i] && hvi[j].x == hvi[i].x) ++j;
      if (j - i > 1) {
        std::unordered_set<std::string> seen;
        for (uint64_t k = i; k < j; ++k) seen.insert(str_of(hvi[k].y));
        out[hgr[i]] += static_cast<int64_t>(seen.size()) - 1;
      }
      i = j;
    }
  }
  return out;
}

// Group key value of one row copied into an output column (slice + concat,
// hash_aggregate.rs:302-314).
static void append_value(OutColumn& o, const HostColumn& c, int64_t row, int64_t out_row) {
  const bool valid = host_valid(c, row);
  if (!valid) {
    if (o.validity.empty()) o.validity.assign(static_cast<size_t>((o.length + 7) / 8), 0xff);
    o.validity[out_row >> 3] &= static_cast<uint8_t>(~(1u << (out_row & 7)));
    o.null_count++;
  }
  const int64_t i = c.offset + row;
  switch (c.type) {
    case DType::Utf8: {
      const int32_t b = c.offsets[i], e = c.offsets[i + 1];
      const uint8_t* s = static_cast<const uint8_t*>(c.data) + b;
      o.data.insert(o.data.end(), s, s + (e - b));
      o.offsets.push_back(static_cast<int32_t>(o.data.size()));
      break;
    }
    case DType::Bool: {
      const bool bit = (static_cast<const uint8_t*>(c.data)[i >> 3] >> (i & 7)) & 1;
      if (bit) o.data[out_row >> 3] |= static_cast<uint8_t>(1u << (out_row & 7));
      break;
    }
    default: {
      const int w = dtype_width(c.type);
      memcpy(o.data.data() + out_row * w, static_cast<const uint8_t*>(c.data) + i * w, w);
    }
  }
}

static bool is_int_type(DType t) { return t == DType::Int32 || t == DType::Int64 || t == DType::Bool; }

static void aggregate(prism_ctx* ctx, const std::vector<HostBatch>& batches, const std::vector<HostColumn>& schema,
                      const prism_hash_agg_config* cfg, ArrowArray* out, ArrowSchema* out_schema) {
  if (!cfg) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL aggregation config");
  std::vector<uint32_t> group_by(cfg->group_by, cfg->group_by + cfg->n_group_by);
  std::vector<prism_agg_expr> aggs(cfg->aggregates, cfg->aggregates + cfg->n_aggregates);
  const bool exact = (cfg->flags & PRISM_AGG_FLAG_EXACT_DECIMAL) != 0;
  const size_t n_cols = schema.size();
  for (uint32_t k : group_by)
    if (k >= n_cols) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: group-by column out of range");
  for (const prism_agg_expr& a : aggs) {
    if (a.func < PRISM_AGG_SUM || a.func > PRISM_AGG_COUNT_DISTINCT)
      fail(PRISM_ERR_UNSUPPORTED_AGGREGATION, "Aggregation error: unsupported function " + std::to_string(a.func));
    const bool star = !group_by.empty() ? false : (a.func == PRISM_AGG_COUNT && a.column >= n_cols);
    if (a.column >= n_cols && !star && !(a.func == PRISM_AGG_COUNT && !group_by.empty() && a.column >= n_cols))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: aggregate column out of range");
    if (a.func != PRISM_AGG_COUNT && a.func != PRISM_AGG_COUNT_DISTINCT && a.column < n_cols) {
      DType t = schema[a.column].type;
      if (!(t == DType::Int32 || t == DType::Int64 || t == DType::Float64 || t == DType::Bool))
        fail(PRISM_ERR_ARROW, "Arrow error: cannot cast " + dtype_name(t) + " to Float64 for aggregation");
    }
  }
  bool any_rows = false;
  for (const HostBatch& b : batches) any_rows |= b.n_rows > 0;
  std::vector<OutColumn> cols;
  auto agg_type = [&](const prism_agg_expr& a) {
    if (a.func == PRISM_AGG_COUNT || a.func == PRISM_AGG_COUNT_DISTINCT) return DType::Int64;
    if (exact && a.func == PRISM_AGG_SUM && a.column < n_cols && is_int_type(schema[a.column].type))
      return DType::Decimal128;
    return DType::Float64;
  };
  if (!any_rows) {  // build_empty_result (hash_aggregate.rs:506-527)
    for (uint32_t k : group_by) {
      OutColumn o;
      o.type = schema[k].type;
      o.name = schema[k].name;
      o.nullable = schema[k].nullable;
      cols.push_back(std::move(o));
    }
    for (const prism_agg_expr& a : aggs) {
      OutColumn o;
      o.type = agg_type(a);
      o.name = a.output_name ? a.output_name : "";
      cols.push_back(std::move(o));
    }
    export_batch(std::move(cols), 0, out, out_schema);
    return;
  }
  const bool global = group_by.empty();
  // Plan over the columns the aggregation reads: keys first, then values.
  std::vector<uint32_t> used;
  std::map<uint32_t, uint32_t> slot;
  auto use = [&](uint32_t col) {
    auto it = slot.find(col);
    if (it != slot.end()) return it->second;
    uint32_t s = static_cast<uint32_t>(used.size());
    slot[col] = s;
    used.push_back(col);
    return s;
  };
  std::vector<uint32_t> key_slots;
  for (uint32_t k : group_by) key_slots.push_back(use(k));
  std::vector<prism_instr> code;
  std::vector<uint32_t> lens;
  std::map<uint32_t, int32_t> value_of_col;
  std::vector<prism_groupby_agg> gaggs;
  std::vector<int> distinct_idx;
  for (size_t i = 0; i < aggs.size(); ++i) {
    const prism_agg_expr& a = aggs[i];
    prism_groupby_agg g;
    g.func = a.func == PRISM_AGG_COUNT_DISTINCT ? PRISM_AGG_COUNT : a.func;
    g.value = -1;
    if (a.func == PRISM_AGG_COUNT_DISTINCT) distinct_idx.push_back(static_cast<int>(i));
    const bool needs_value = a.column < n_cols && (a.func != PRISM_AGG_COUNT || global) &&
                             a.func != PRISM_AGG_COUNT_DISTINCT;
    if (needs_value) {
      auto it = value_of_col.find(a.column);
      if (it == value_of_col.end()) {
        const uint32_t s = use(a.column);
        prism_instr ins;
        memset(&ins, 0, sizeof(ins));
        const DType t = schema[a.column].type;
        ins.op = t == DType::Float64 ? PRISM_OP_LOAD_F : PRISM_OP_LOAD_I;
        ins.arg = static_cast<int32_t>(s);
        code.push_back(ins);
        lens.push_back(1);
        it = value_of_col.emplace(a.column, static_cast<int32_t>(lens.size() - 1)).first;
      }
      g.value = it->second;
    }
    gaggs.push_back(g);
  }
  std::vector<int32_t> types;
  for (uint32_t col : used) types.push_back(static_cast<int32_t>(schema[col].type));
  prism_groupby_plan plan;
  memset(&plan, 0, sizeof(plan));
  plan.n_columns = static_cast<uint32_t>(used.size());
  plan.column_types = types.data();
  plan.key_columns = key_slots.data();
  plan.n_keys = static_cast<uint32_t>(key_slots.size());
  plan.values = code.data();
  plan.value_lengths = lens.data();
  plan.n_values = static_cast<uint32_t>(lens.size());
  plan.aggs = gaggs.data();
  plan.n_aggs = static_cast<uint32_t>(gaggs.size());
  prism_groupby* g = nullptr;
  prism_status s = prism_groupby_create(ctx, &plan, 0, &g);
  if (s != PRISM_OK) fail(s, prism_last_error());
  std::unique_ptr<prism_groupby, void (*)(prism_groupby*)> guard(g, prism_groupby_destroy);
  std::vector<int64_t> batch_start;
  int64_t base = 0;
  for (const HostBatch& b : batches) {
    batch_start.push_back(base);
    if (b.n_rows > 0) {
      std::vector<DevColumn> dc;
      std::vector<prism_column> desc;
      for (uint32_t col : used) {
        dc.push_back(upload(ctx, b.cols[col]));
        desc.push_back(dc.back().desc());
      }
      PRISM_HIP(hipStreamSynchronize(ctx->stream));
      s = prism_groupby_update(g, desc.data(), static_cast<uint64_t>(b.n_rows), static_cast<uint64_t>(base));
      if (s != PRISM_OK) fail(s, prism_last_error());
    }
    base += b.n_rows;
  }
  uint64_t n_groups = 0;
  s = prism_groupby_finish(g, &n_groups);
  if (s != PRISM_OK) fail(s, prism_last_error());
  const size_t na = aggs.size();
  std::vector<uint64_t> first(n_groups), khash(n_groups), rcount(n_groups);
  std::vector<std::vector<double>> f64(na, std::vector<double>(n_groups));
  std::vector<std::vector<int64_t>> i64(na, std::vector<int64_t>(n_groups));
  std::vector<std::vector<uint64_t>> s128(na, std::vector<uint64_t>(2 * n_groups));
  std::vector<double*> fp(na);
  std::vector<int64_t*> ip(na);
  std::vector<uint64_t*> sp(na);
  for (size_t a = 0; a < na; ++a) {
    fp[a] = f64[a].data();
    ip[a] = i64[a].data();
    sp[a] = s128[a].data();
  }
  prism_groupby_output go{first.data(), khash.data(), rcount.data(), fp.data(), ip.data(), sp.data()};
  s = prism_groupby_read(g, &go);
  if (s != PRISM_OK) fail(s, prism_last_error());
  // COUNT DISTINCT per group id
  std::vector<std::unordered_map<uint64_t, int64_t>> distinct(na);
  for (int a : distinct_idx) distinct[a] = count_distinct(ctx, batches, group_by, aggs[a].column);
  // Output: keys from each group's first row, then aggregates.
  for (uint32_t k : group_by) {
    const HostColumn& sc = schema[k];
    OutColumn o;
    o.type = sc.type;
    o.name = sc.name;
    o.nullable = sc.nullable;
    o.length = static_cast<int64_t>(n_groups);
    if (sc.type == DType::Utf8) o.offsets.push_back(0);
    else if (sc.type == DType::Bool) o.data.assign(static_cast<size_t>((n_groups + 7) / 8), 0);
    else o.data.resize(static_cast<size_t>(n_groups) * dtype_width(sc.type));
    for (uint64_t i = 0; i < n_groups; ++i) {
      size_t bi = std::upper_bound(batch_start.begin(), batch_start.end(), static_cast<int64_t>(first[i])) -
                  batch_start.begin() - 1;
      while (batches[bi].n_rows == 0) --bi;
      append_value(o, batches[bi].cols[k], static_cast<int64_t>(first[i]) - batch_start[bi], static_cast<int64_t>(i));
    }
    cols.push_back(std::move(o));
  }
  for (size_t a = 0; a < na; ++a) {
    OutColumn o;
    o.type = agg_type(aggs[a]);
    o.name = aggs[a].output_name ? aggs[a].output_name : "";
    o.length = static_cast<int64_t>(n_groups);
    if (o.type == DType::Int64) {
      o.data.resize(n_groups * 8);
      for (uint64_t i = 0; i < n_groups; ++i) {
        int64_t v = i64[a][i];
        if (aggs[a].func == PRISM_AGG_COUNT_DISTINCT) {
          auto it = distinct[a].find(khash[i]);
          v = it == distinct[a].end() ? 0 : it->second;
        }
        memcpy(o.data.data() + i * 8, &v, 8);
      }
    } else if (o.type == DType::Decimal128) {
      o.data.resize(n_groups * 16);
      memcpy(o.data.data(), s128[a].data(), n_groups * 16);
    } else {
      o.data.resize(n_groups * 8);
      memcpy(o.data.data(), f64[a].data(), n_groups * 8);
    }
    cols.push_back(std::move(o));
  }
  export_batch(std::move(cols), static_cast<int64_t>(n_groups), out, out_schema);
}

}  // namespace aagg
}  // namespace prism

using namespace prism;
using namespace prism::aagg;

extern "C" {

prism_status prism_hash_aggregate_batches(prism_ctx* ctx, const struct ArrowArray* const* batches, uint32_t n_batches,
                                          const struct ArrowSchema* schema, const prism_hash_agg_config* cfg,
                                          struct ArrowArray* out, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context");
    Lane lane(ctx);
    std::vector<HostBatch> bs;
    for (uint32_t i = 0; i < n_batches; ++i) bs.push_back(import_batch(batches[i], schema));
    std::vector<HostColumn> sch = import_schema(schema);
    aggregate(ctx, bs, sch, cfg, out, out_schema);
  });
}

prism_status prism_hash_aggregate(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                  const prism_hash_agg_config* cfg, struct ArrowArray* out,
                                  struct ArrowSchema* out_schema) {
  const struct ArrowArray* one[1] = {batch};
  return prism_hash_aggregate_batches(ctx, one, 1, schema, cfg, out, out_schema);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Distributed aggregate, reducer side (prism-bench handler.rs:270-344,
// 397-526): producers run the plan with AVG split into SUM + COUNT; the
// reducer merges every producer's partial batch with one device aggregation
// and rebuilds the original aggregates.
// ---------------------------------------------------------------------------
namespace {

struct Expanded {
  std::vector<prism_agg_expr> worker;
  std::vector<std::string> names;  // storage of worker[i].output_name
  std::vector<uint32_t> map;       // 2 per original aggregate; second UINT32_MAX unless AVG
};

// expand_worker_aggregates (handler.rs:397-428).
void expand(const prism_agg_expr* original, uint32_t n, Expanded& e) {
  e.worker.clear();
  e.names.clear();
  e.map.assign(2 * static_cast<size_t>(n), UINT32_MAX);
  auto push = [&](uint32_t column, int32_t func) {
    const uint32_t idx = static_cast<uint32_t>(e.worker.size());
    prism_agg_expr w;
    w.column = column;
    w.func = func;
    w.output_name = nullptr;
    e.worker.push_back(w);
    e.names.push_back("agg_" + std::to_string(idx));
    return idx;
  };
  for (uint32_t i = 0; i < n; ++i) {
    const prism_agg_expr& a = original[i];
    if (a.func < PRISM_AGG_SUM || a.func > PRISM_AGG_COUNT_DISTINCT)
      fail(PRISM_ERR_UNSUPPORTED_AGGREGATION, "Aggregation error: unsupported function " + std::to_string(a.func));
    if (a.func == PRISM_AGG_AVG) {
      e.map[2 * i] = push(a.column, PRISM_AGG_SUM);
      e.map[2 * i + 1] = push(a.column, PRISM_AGG_COUNT);
    } else {
      e.map[2 * i] = push(a.column, a.func);
    }
  }
  for (size_t i = 0; i < e.worker.size(); ++i) e.worker[i].output_name = e.names[i].c_str();
}

// numeric_value_as_f64 / numeric_value_as_i64 (handler.rs:528-566): nulls read as 0.
double value_f64(const HostColumn& c, int64_t row) {
  if (!host_valid(c, row)) return 0.0;
  const int64_t i = c.offset + row;
  switch (c.type) {
    case DType::Float64: return static_cast<const double*>(c.data)[i];
    case DType::Int64: return static_cast<double>(static_cast<const int64_t*>(c.data)[i]);
    case DType::Int32: return static_cast<double>(static_cast<const int32_t*>(c.data)[i]);
    default:
      fail(PRISM_ERR_INVALID_ARGUMENT,
           "Invalid argument: unsupported numeric type for distributed aggregate merge: " + dtype_name(c.type));
  }
}

// Rust's `f64 as i64`: truncation toward zero, saturating, NaN -> 0.
int64_t f64_as_i64(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 9223372036854775807.0) return INT64_MAX;
  if (v <= -9223372036854775808.0) return INT64_MIN;
  return static_cast<int64_t>(v);
}

int64_t value_i64(const HostColumn& c, int64_t row) {
  if (!host_valid(c, row)) return 0;
  const int64_t i = c.offset + row;
  switch (c.type) {
    case DType::Int64: return static_cast<const int64_t*>(c.data)[i];
    case DType::Int32: return static_cast<const int32_t*>(c.data)[i];
    case DType::Float64: return f64_as_i64(static_cast<const double*>(c.data)[i]);
    default:
      fail(PRISM_ERR_INVALID_ARGUMENT,
           "Invalid argument: unsupported integer type for distributed aggregate merge: " + dtype_name(c.type));
  }
}

// A host column copied into an output column (group keys pass through).
OutColumn copy_out(const HostColumn& c, int64_t n) {
  OutColumn o;
  o.type = c.type;
  o.name = c.name;
  o.nullable = c.nullable;
  o.length = n;
  if (c.validity && c.null_count != 0) {
    o.validity.assign(static_cast<size_t>((n + 7) / 8), 0);
    for (int64_t r = 0; r < n; ++r)
      if (host_valid(c, r)) o.validity[r >> 3] |= static_cast<uint8_t>(1u << (r & 7));
      else ++o.null_count;
  }
  if (c.type == DType::Utf8) {
    o.offsets.resize(static_cast<size_t>(n) + 1);
    const int32_t base = n ? c.offsets[c.offset] : 0;
    for (int64_t r = 0; r <= n; ++r) o.offsets[r] = c.offsets[c.offset + r] - base;
    const uint8_t* b = static_cast<const uint8_t*>(c.data);
    o.data.assign(b + base, b + base + o.offsets[n]);
  } else if (c.type == DType::Bool) {
    o.data.assign(static_cast<size_t>((n + 7) / 8), 0);
    const uint8_t* b = static_cast<const uint8_t*>(c.data);
    for (int64_t r = 0; r < n; ++r) {
      const int64_t i = c.offset + r;
      if ((b[i >> 3] >> (i & 7)) & 1) o.data[r >> 3] |= static_cast<uint8_t>(1u << (r & 7));
    }
  } else {
    const int w = dtype_width(c.type);
    const uint8_t* b = static_cast<const uint8_t*>(c.data) + c.offset * w;
    o.data.assign(b, b + n * w);
  }
  return o;
}

OutColumn f64_out(const char* name, std::vector<double>&& v) {
  OutColumn o;
  o.type = DType::Float64;
  o.name = name ? name : "";
  o.length = static_cast<int64_t>(v.size());
  o.data.resize(v.size() * 8);
  memcpy(o.data.data(), v.data(), v.size() * 8);
  return o;
}

OutColumn i64_out(const char* name, std::vector<int64_t>&& v) {
  OutColumn o;
  o.type = DType::Int64;
  o.name = name ? name : "";
  o.length = static_cast<int64_t>(v.size());
  o.data.resize(v.size() * 8);
  memcpy(o.data.data(), v.data(), v.size() * 8);
  return o;
}

// Releases an exported Arrow array / schema pair when it goes out of scope.
struct ArrowHold {
  ArrowArray a{};
  ArrowSchema s{};
  ~ArrowHold() {
    if (a.release) a.release(&a);
    if (s.release) s.release(&s);
  }
};

}  // namespace

extern "C" {

prism_status prism_expand_worker_aggregates(const prism_agg_expr* original, uint32_t n, prism_agg_expr* out_worker,
                                            uint32_t* n_worker, uint32_t* out_map) {
  return guarded([&] {
    if ((n && !original) || !out_worker || !n_worker || !out_map)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL aggregate list or output");
    static thread_local Expanded e;  // names stay valid until the next call on this thread
    expand(original, n, e);
    for (size_t i = 0; i < e.worker.size(); ++i) out_worker[i] = e.worker[i];
    *n_worker = static_cast<uint32_t>(e.worker.size());
    for (size_t i = 0; i < e.map.size(); ++i) out_map[i] = e.map[i];
  });
}

prism_status prism_merge_partial_aggregates(prism_ctx* ctx, const struct ArrowArray* const* partials,
                                            uint32_t n_partials, const struct ArrowSchema* schema,
                                            uint32_t group_by_count, const prism_agg_expr* original,
                                            uint32_t n_original, struct ArrowArray* out,
                                            struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !schema || !out || !out_schema) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL argument");
    if ((n_partials && !partials) || (n_original && !original))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL partial batches or aggregate list");
    Lane lane(ctx);
    Expanded e;
    expand(original, n_original, e);
    std::vector<HostColumn> sch = import_schema(schema);
    if (sch.size() < group_by_count + e.worker.size())
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: partial batches hold " + std::to_string(sch.size()) +
                                           " columns, expected the group columns and " +
                                           std::to_string(e.worker.size()) + " worker aggregates");
    // build_final_merge_config (handler.rs:430-457)
    std::vector<uint32_t> gb(group_by_count);
    for (uint32_t i = 0; i < group_by_count; ++i) gb[i] = i;
    std::vector<prism_agg_expr> merge(e.worker.size());
    for (size_t i = 0; i < e.worker.size(); ++i) {
      const int32_t f = e.worker[i].func;
      merge[i].column = group_by_count + static_cast<uint32_t>(i);
      merge[i].func = (f == PRISM_AGG_MIN || f == PRISM_AGG_MAX) ? f : PRISM_AGG_SUM;
      merge[i].output_name = e.worker[i].output_name;
    }
    prism_hash_agg_config cfg{gb.data(), group_by_count, merge.data(), static_cast<uint32_t>(merge.size()), 0};
    std::vector<HostBatch> bs;
    for (uint32_t i = 0; i < n_partials; ++i) bs.push_back(import_batch(partials[i], schema));
    ArrowHold merged;
    aggregate(ctx, bs, sch, &cfg, &merged.a, &merged.s);
    // reconstruct_original_aggregate_batch (handler.rs:459-526)
    const HostBatch m = import_batch(&merged.a, &merged.s);
    const int64_t n = m.n_rows;
    std::vector<OutColumn> cols;
    for (uint32_t k = 0; k < group_by_count; ++k) cols.push_back(copy_out(m.cols[k], n));
    for (uint32_t i = 0; i < n_original; ++i) {
      const prism_agg_expr& a = original[i];
      const HostColumn& src = m.cols[group_by_count + e.map[2 * i]];
      if (a.func == PRISM_AGG_COUNT || a.func == PRISM_AGG_COUNT_DISTINCT) {
        std::vector<int64_t> v(static_cast<size_t>(n));
        for (int64_t r = 0; r < n; ++r) v[r] = value_i64(src, r);
        cols.push_back(i64_out(a.output_name, std::move(v)));
      } else if (a.func == PRISM_AGG_AVG) {
        const HostColumn& cnt = m.cols[group_by_count + e.map[2 * i + 1]];
        std::vector<double> v(static_cast<size_t>(n));
        for (int64_t r = 0; r < n; ++r) {
          const double s = value_f64(src, r), c = value_f64(cnt, r);
          v[r] = c > 0.0 ? s / c : 0.0;
        }
        cols.push_back(f64_out(a.output_name, std::move(v)));
      } else {
        std::vector<double> v(static_cast<size_t>(n));
        for (int64_t r = 0; r < n; ++r) v[r] = value_f64(src, r);
        cols.push_back(f64_out(a.output_name, std::move(v)));
      }
    }
    export_batch(std::move(cols), n, out, out_schema);
  });
}

}  // extern "C"
