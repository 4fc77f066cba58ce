// Arrow-level hash join: the operator API of native/prism-executor/src/hash_join.rs
// over the Arrow C Data Interface, executed on the device.
//
//   HashTable::build / probe     hash_join.rs:86-201
//   hash_join                    hash_join.rs:312-365
//   probe_matches_chunked        hash_join.rs:286-309
//   hash_join_probe_chunked      hash_join.rs:206-282
//
// Key paths, as in the reference:
//   one Int64 key on both sides  -> the int64 table (hash_i64_fast is a
//                                   bijection: the key itself decides, nulls are
//                                   not consulted, hash_join.rs:152-173);
//   anything else                -> a 64-bit row hash keys the table and every
//                                   candidate pair is verified on the device with
//                                   keys_equal's rules (hash_join.rs:175-199,
//                                   628-670): a null meets only a null whose
//                                   stored value is equal, f64 compares by ==.
#include "gather.hpp"
#include "program.hpp"

#include <vector>

namespace prism {
namespace ajoin {

constexpr int kB = 256;

static dim3 jgrid(prism_ctx* ctx, uint64_t n) {
  uint64_t b = ceil_div(n ? n : 1, kB);
  uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 16;
  return dim3(static_cast<unsigned>(b < cap ? b : cap));
}

struct KeyCols {
  prism_column c[kMaxKeys];
  int n;
};

__device__ __forceinline__ uint64_t hmix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Placement hash of one row's keys (the reference uses SipHash here; only
// placement depends on it, equality is verified separately).
__global__ void k_row_hash(KeyCols K, uint64_t n, int64_t* __restrict__ out) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t h = 0x243F6A8885A308D3ULL;
    for (int k = 0; k < K.n; ++k) {
      const prism_column& c = K.c[k];
      if (!col_valid(c, r)) {
        h = hmix(h ^ 0x6E756C6CULL);  // null: its own word (a null never meets a value)
        continue;
      }
      const uint64_t i = c.offset + r;
      uint64_t w;
      switch (c.type) {
        case PRISM_TYPE_INT64: w = static_cast<uint64_t>(static_cast<const int64_t*>(c.data)[i]); break;
        case PRISM_TYPE_FLOAT64: w = static_cast<const uint64_t*>(c.data)[i]; break;
        case PRISM_TYPE_BOOL: w = bit_at(static_cast<const uint8_t*>(c.data), i); break;
        case PRISM_TYPE_UTF8: {
          const int32_t b = c.offsets[i], e = c.offsets[i + 1];
          const uint8_t* s = static_cast<const uint8_t*>(c.data) + b;
          uint64_t x = static_cast<uint64_t>(e - b) * 0x9E3779B97F4A7C15ULL;
          for (int32_t p = 0; p < e - b; ++p) x = hmix(x ^ s[p]);
          w = x;
          break;
        }
        default: w = static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int32_t*>(c.data)[i]));
      }
      h = hmix(h ^ w);
    }
    out[r] = static_cast<int64_t>(h);
  }
}

__device__ bool key_eq(const prism_column& a, uint64_t ra, const prism_column& b, uint64_t rb) {
  const bool va = col_valid(a, ra), vb = col_valid(b, rb);
  if (va != vb) return false;
  const uint64_t ia = a.offset + ra, ib = b.offset + rb;
  switch (a.type) {
    case PRISM_TYPE_INT64: return static_cast<const int64_t*>(a.data)[ia] == static_cast<const int64_t*>(b.data)[ib];
    case PRISM_TYPE_FLOAT64: {
      const double x = static_cast<const double*>(a.data)[ia], y = static_cast<const double*>(b.data)[ib];
      return static_cast<const uint64_t*>(a.data)[ia] == static_cast<const uint64_t*>(b.data)[ib] && x == y;
    }
    case PRISM_TYPE_BOOL:
      return bit_at(static_cast<const uint8_t*>(a.data), ia) == bit_at(static_cast<const uint8_t*>(b.data), ib);
    case PRISM_TYPE_UTF8: {
      const int32_t a0 = a.offsets[ia], a1 = a.offsets[ia + 1], b0 = b.offsets[ib], b1 = b.offsets[ib + 1];
      if (a1 - a0 != b1 - b0) return false;
      const uint8_t* x = static_cast<const uint8_t*>(a.data) + a0;
      const uint8_t* y = static_cast<const uint8_t*>(b.data) + b0;
      for (int32_t p = 0; p < a1 - a0; ++p)
        if (x[p] != y[p]) return false;
      return true;
    }
    default: return static_cast<const int32_t*>(a.data)[ia] == static_cast<const int32_t*>(b.data)[ib];
  }
}

__global__ void k_verify(KeyCols P, KeyCols Bk, const uint32_t* __restrict__ pi, const uint32_t* __restrict__ bi,
                         uint64_t n, uint8_t* __restrict__ keep) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    bool ok = true;
    for (int k = 0; k < P.n && ok; ++k) ok = key_eq(P.c[k], pi[i], Bk.c[k], bi[i]);
    keep[i] = ok;
  }
}

__global__ void k_set_bits_prefix(unsigned long long* bits, uint64_t n_total, uint64_t n_set) {
  const uint64_t nw = ceil_div(n_total, 64);
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = w * 64;
    uint64_t word;
    if (lo + 64 <= n_set) word = ~0ULL;
    else if (lo >= n_set) word = 0;
    else word = (1ULL << (n_set - lo)) - 1;
    bits[w] = word;
  }
}

// Matched (probe_row, build_row) pairs on the device.
struct Pairs {
  DeviceBuffer p, b;
  uint64_t n = 0;
};

struct Table {
  Ctx* ctx = nullptr;
  bool fast = false;
  std::vector<DType> key_types;
  std::vector<DevColumn> keys;  // general path: build key columns kept for verification
  prism_join_table* t = nullptr;
  uint64_t n_rows = 0;
  ~Table() {
    if (t) prism_join_table_free(t);
  }
};

static KeyCols key_cols(const std::vector<DevColumn>& cols) {
  KeyCols K;
  memset(&K, 0, sizeof(K));
  K.n = static_cast<int>(cols.size());
  for (size_t i = 0; i < cols.size(); ++i) K.c[i] = cols[i].desc();
  return K;
}

static void check_keys(const HostBatch& b, const uint32_t* keys, uint32_t n, const char* side) {
  if (n == 0) fail(PRISM_ERR_JOIN, std::string("Join error: no ") + side + " key columns");
  if (n > static_cast<uint32_t>(kMaxKeys)) fail(PRISM_ERR_JOIN, "Join error: at most 8 key columns");
  for (uint32_t i = 0; i < n; ++i) {
    if (keys[i] >= b.cols.size())
      fail(PRISM_ERR_INVALID_ARGUMENT, std::string("Invalid argument: ") + side + " key column out of range");
    DType t = b.cols[keys[i]].type;
    if (t == DType::UInt32 || t == DType::Decimal128)
      fail(PRISM_ERR_JOIN, "Join error: unsupported key type " + dtype_name(t));
  }
}

static void row_hash(prism_ctx* ctx, const std::vector<DevColumn>& keys, uint64_t n, DeviceBuffer& out) {
  out.reserve(as_ctx(ctx), (n ? n : 1) * 8);
  if (n) hipLaunchKernelGGL(k_row_hash, jgrid(ctx, n), dim3(kB), 0, ctx->stream, key_cols(keys), n, out.as<int64_t>());
  PRISM_CHECK_LAUNCH();
}

static std::unique_ptr<Table> build_table(prism_ctx* ctx, const HostBatch& b, const uint32_t* keys, uint32_t n_keys) {
  check_keys(b, keys, n_keys, "build");
  auto T = std::make_unique<Table>();
  T->ctx = as_ctx(ctx);
  T->n_rows = static_cast<uint64_t>(b.n_rows);
  for (uint32_t i = 0; i < n_keys; ++i) T->key_types.push_back(b.cols[keys[i]].type);
  T->fast = n_keys == 1 && b.cols[keys[0]].type == DType::Int64;
  if (T->fast) {
    DevColumn k = upload(ctx, b.cols[keys[0]]);
    const int64_t* kp = k.data.as<int64_t>() + k.dev_offset;
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
    prism_status s = prism_join_build_i64(ctx, kp, T->n_rows, PRISM_TABLE_AUTO, &T->t);
    if (s != PRISM_OK) fail(s, prism_last_error());
  } else {
    for (uint32_t i = 0; i < n_keys; ++i) T->keys.push_back(upload(ctx, b.cols[keys[i]]));
    DeviceBuffer h;
    row_hash(ctx, T->keys, T->n_rows, h);
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
    prism_status s = prism_join_build_i64(ctx, h.as<int64_t>(), T->n_rows, PRISM_TABLE_HASHED, &T->t);
    if (s != PRISM_OK) fail(s, prism_last_error());
  }
  return T;
}

// Probe: ordered pairs (probe row ascending, build rows ascending).
static Pairs probe_table(prism_ctx* ctx, const Table& T, const HostBatch& p, const uint32_t* keys, uint32_t n_keys) {
  check_keys(p, keys, n_keys, "probe");
  if (n_keys != T.key_types.size()) fail(PRISM_ERR_JOIN, "Join error: build and probe key counts differ");
  for (uint32_t i = 0; i < n_keys; ++i)
    if (p.cols[keys[i]].type != T.key_types[i])
      fail(PRISM_ERR_JOIN, "Join error: key types differ (" + dtype_name(p.cols[keys[i]].type) + " vs " +
                               dtype_name(T.key_types[i]) + ")");
  Ctx* c = as_ctx(ctx);
  const uint64_t n = static_cast<uint64_t>(p.n_rows);
  if (n >= 0xFFFFFFFFULL) fail(PRISM_ERR_JOIN, "Join error: probe batch exceeds u32 row indices");
  std::vector<DevColumn> pk;
  DeviceBuffer h;
  const int64_t* kp;
  if (T.fast) {
    pk.push_back(upload(ctx, p.cols[keys[0]]));
    kp = pk[0].data.as<int64_t>() + pk[0].dev_offset;
  } else {
    for (uint32_t i = 0; i < n_keys; ++i) pk.push_back(upload(ctx, p.cols[keys[i]]));
    row_hash(ctx, pk, n, h);
    kp = h.as<int64_t>();
  }
  prism_join_table_info info;
  prism_join_table_info_get(T.t, &info);
  uint64_t cap = info.unique ? n : n + 1024;
  DeviceBuffer res(c, sizeof(prism_probe_result));
  Pairs out;
  while (true) {
    out.p.reserve(c, (cap ? cap : 1) * 4);
    out.b.reserve(c, (cap ? cap : 1) * 4);
    prism_status s = prism_join_probe_i64(ctx, T.t, kp, n, 0, nullptr, nullptr, out.p.as<uint32_t>(),
                                          out.b.as<uint32_t>(), cap, res.as<prism_probe_result>());
    if (s != PRISM_OK) fail(s, prism_last_error());
    prism_probe_result r;
    PRISM_HIP(hipMemcpyAsync(&r, res.ptr, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
    if (r.error) fail(PRISM_ERR_EXECUTION, "Execution error: probe look-back wait gave up");
    if (r.matches > cap) {
      cap = r.matches;
      continue;
    }
    out.n = r.matches;
    break;
  }
  if (!T.fast && out.n) {
    DeviceBuffer keep(c, out.n), bits(c, ceil_div(out.n, 64) * 8), sel;
    hipLaunchKernelGGL(k_verify, jgrid(ctx, out.n), dim3(kB), 0, ctx->stream, key_cols(pk), key_cols(T.keys),
                       out.p.as<uint32_t>(), out.b.as<uint32_t>(), out.n, keep.as<uint8_t>());
    PRISM_CHECK_LAUNCH();
    flags_to_bits(ctx, keep.as<uint8_t>(), out.n, true, bits.as<unsigned long long>());
    const uint64_t m = select_bits(ctx, bits.as<unsigned long long>(), out.n, sel);
    if (m != out.n) {
      DevColumn pc, bc;
      pc.type = bc.type = DType::UInt32;
      pc.length = bc.length = static_cast<int64_t>(out.n);
      pc.data = std::move(out.p);
      bc.data = std::move(out.b);
      DevColumn np = take(ctx, pc, sel.as<uint32_t>(), nullptr, m);
      DevColumn nb = take(ctx, bc, sel.as<uint32_t>(), nullptr, m);
      out.p = std::move(np.data);
      out.b = std::move(nb.data);
      out.n = m;
    }
  }
  PRISM_HIP(hipStreamSynchronize(ctx->stream));
  return out;
}

static std::vector<DevColumn> upload_all(prism_ctx* ctx, const HostBatch& b) {
  std::vector<DevColumn> v;
  for (const HostColumn& c : b.cols) v.push_back(upload(ctx, c));
  return v;
}

static OutColumn index_array(prism_ctx* ctx, DeviceBuffer& buf, uint64_t n, const char* name) {
  DevColumn d;
  d.type = DType::UInt32;
  d.name = name;
  d.nullable = false;
  d.length = static_cast<int64_t>(n);
  d.data = std::move(buf);
  OutColumn o = download(ctx, d);
  buf = std::move(d.data);
  return o;
}

// Rows of [0, n_rows) whose flag (1 = present in idx) equals `want`, ordered.
static uint64_t rows_where(prism_ctx* ctx, const uint32_t* idx, uint64_t n_idx, uint64_t n_rows, bool want,
                           DeviceBuffer& out) {
  Ctx* c = as_ctx(ctx);
  DeviceBuffer flags(c, n_rows ? n_rows : 1), bits(c, ceil_div(n_rows ? n_rows : 1, 64) * 8);
  mark_rows(ctx, idx, n_idx, flags.as<uint8_t>(), n_rows);
  flags_to_bits(ctx, flags.as<uint8_t>(), n_rows, want, bits.as<unsigned long long>());
  return select_bits(ctx, bits.as<unsigned long long>(), n_rows, out);
}

struct Assembly {
  DeviceBuffer pidx, bidx, pvalid, bvalid;  // validity bitmaps (empty = all valid)
  uint64_t n = 0;
};

// Join-type index construction (hash_join.rs:455-566) on the device.
static Assembly assemble_indices(prism_ctx* ctx, int32_t jt, Pairs& m, uint64_t n_probe, uint64_t n_build) {
  Ctx* c = as_ctx(ctx);
  Assembly A;
  if (jt == PRISM_JOIN_INNER) {
    A.pidx = std::move(m.p);
    A.bidx = std::move(m.b);
    A.n = m.n;
    return A;
  }
  if (jt == PRISM_JOIN_LEFT) {  // unmatched probe rows get build row 0 (hash_join.rs:472-477)
    DeviceBuffer un;
    uint64_t k = rows_where(ctx, m.p.as<uint32_t>(), m.n, n_probe, false, un);
    A.n = m.n + k;
    A.pidx.alloc(c, (A.n ? A.n : 1) * 4);
    A.bidx.alloc(c, (A.n ? A.n : 1) * 4);
    copy_u32(ctx, A.pidx.as<uint32_t>(), m.p.as<uint32_t>(), m.n);
    copy_u32(ctx, A.bidx.as<uint32_t>(), m.b.as<uint32_t>(), m.n);
    copy_u32(ctx, A.pidx.as<uint32_t>() + m.n, un.as<uint32_t>(), k);
    fill_bytes(ctx, reinterpret_cast<uint8_t*>(A.bidx.as<uint32_t>() + m.n), k * 4, 0);
    return A;
  }
  // RIGHT / FULL: unmatched build rows with a null probe side (hash_join.rs:495-566)
  DeviceBuffer unp, unb;
  uint64_t kp = jt == PRISM_JOIN_FULL ? rows_where(ctx, m.p.as<uint32_t>(), m.n, n_probe, false, unp) : 0;
  uint64_t kb = rows_where(ctx, m.b.as<uint32_t>(), m.n, n_build, false, unb);
  A.n = m.n + kp + kb;
  const uint64_t nn = A.n ? A.n : 1;
  A.pidx.alloc(c, nn * 4);
  A.bidx.alloc(c, nn * 4);
  copy_u32(ctx, A.pidx.as<uint32_t>(), m.p.as<uint32_t>(), m.n);
  copy_u32(ctx, A.bidx.as<uint32_t>(), m.b.as<uint32_t>(), m.n);
  copy_u32(ctx, A.pidx.as<uint32_t>() + m.n, unp.as<uint32_t>(), kp);
  fill_bytes(ctx, reinterpret_cast<uint8_t*>(A.bidx.as<uint32_t>() + m.n), kp * 4, 0);
  fill_bytes(ctx, reinterpret_cast<uint8_t*>(A.pidx.as<uint32_t>() + m.n + kp), kb * 4, 0);
  copy_u32(ctx, A.bidx.as<uint32_t>() + m.n + kp, unb.as<uint32_t>(), kb);
  // probe valid for the first m.n + kp rows; build valid for all but [m.n, m.n+kp)
  A.pvalid.alloc(c, ceil_div(nn, 64) * 8);
  hipLaunchKernelGGL(k_set_bits_prefix, jgrid(ctx, ceil_div(nn, 64)), dim3(kB), 0, ctx->stream,
                     A.pvalid.as<unsigned long long>(), A.n, m.n + kp);
  if (jt == PRISM_JOIN_FULL) {
    DeviceBuffer flags(c, nn);
    fill_bytes(ctx, flags.as<uint8_t>(), A.n, 1);
    fill_bytes(ctx, flags.as<uint8_t>() + m.n, kp, 0);
    A.bvalid.alloc(c, ceil_div(nn, 64) * 8);
    flags_to_bits(ctx, flags.as<uint8_t>(), A.n, true, A.bvalid.as<unsigned long long>());
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
  }
  PRISM_CHECK_LAUNCH();
  return A;
}

// assemble_joined_batch[_nullable_probe/_both_nullable] (hash_join.rs:424-453,569-625)
static void emit_joined(prism_ctx* ctx, int32_t jt, const HostBatch& probe, const HostBatch& build, Assembly& A,
                        ArrowArray* out, ArrowSchema* out_schema) {
  std::vector<DevColumn> pc = upload_all(ctx, probe), bc = upload_all(ctx, build);
  std::vector<OutColumn> cols;
  const uint8_t* pv = A.pvalid.as<uint8_t>();
  const uint8_t* bv = A.bvalid.as<uint8_t>();
  for (size_t i = 0; i < pc.size(); ++i) {
    DevColumn t = take(ctx, pc[i], A.pidx.as<uint32_t>(), pv, A.n);
    OutColumn o = download(ctx, t);
    if (jt == PRISM_JOIN_RIGHT || jt == PRISM_JOIN_FULL) o.nullable = true;
    else o.nullable = probe.cols[i].nullable;
    cols.push_back(std::move(o));
  }
  for (size_t i = 0; i < bc.size(); ++i) {
    DevColumn t = take(ctx, bc[i], A.bidx.as<uint32_t>(), bv, A.n);
    OutColumn o = download(ctx, t);
    bool clash = false;
    for (const HostColumn& q : probe.cols) clash |= q.name == build.cols[i].name;
    if (clash) o.name = build.cols[i].name + "_right";
    o.nullable = jt == PRISM_JOIN_FULL ? true : build.cols[i].nullable;
    cols.push_back(std::move(o));
  }
  export_batch(std::move(cols), static_cast<int64_t>(A.n), out, out_schema);
}

// LeftSemi / LeftAnti: filter the probe batch (hash_join.rs:238-260,329-351)
static void emit_semi(prism_ctx* ctx, bool anti, const HostBatch& probe, Pairs& m, ArrowArray* out,
                      ArrowSchema* out_schema) {
  DeviceBuffer rows;
  uint64_t k = rows_where(ctx, m.p.as<uint32_t>(), m.n, static_cast<uint64_t>(probe.n_rows), !anti, rows);
  std::vector<OutColumn> cols;
  for (const HostColumn& hc : probe.cols) {
    DevColumn d = upload(ctx, hc);
    DevColumn t = take(ctx, d, rows.as<uint32_t>(), nullptr, k);
    OutColumn o = download(ctx, t);
    o.nullable = hc.nullable;
    cols.push_back(std::move(o));
  }
  export_batch(std::move(cols), static_cast<int64_t>(k), out, out_schema);
}

static void join_one(prism_ctx* ctx, int32_t jt, const Table& T, const HostBatch& probe, const HostBatch& build,
                     const uint32_t* probe_keys, uint32_t n_keys, ArrowArray* out, ArrowSchema* out_schema) {
  Pairs m = probe_table(ctx, T, probe, probe_keys, n_keys);
  if (jt == PRISM_JOIN_LEFT_SEMI || jt == PRISM_JOIN_LEFT_ANTI) {
    emit_semi(ctx, jt == PRISM_JOIN_LEFT_ANTI, probe, m, out, out_schema);
    return;
  }
  Assembly A = assemble_indices(ctx, jt, m, static_cast<uint64_t>(probe.n_rows), static_cast<uint64_t>(build.n_rows));
  emit_joined(ctx, jt, probe, build, A, out, out_schema);
}

static void check_join_type(int32_t jt) {
  if (jt < PRISM_JOIN_INNER || jt > PRISM_JOIN_LEFT_ANTI)
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown join type " + std::to_string(jt));
}

// Left pads unmatched probe rows with build row 0 (hash_join.rs:472-477); with
// no build rows that index does not exist and arrow's take would refuse it
// (the reference panics there).  Refused before anything is launched.
static void check_left_pad(int32_t jt, const HostBatch& build, int64_t probe_rows) {
  if (jt == PRISM_JOIN_LEFT && build.n_rows == 0 && probe_rows > 0)
    fail(PRISM_ERR_ARROW, "Arrow error: left join pads unmatched probe rows with build row 0, but the build side "
                          "is empty (take index 0 out of bounds for length 0)");
}

}  // namespace ajoin
}  // namespace prism

using namespace prism;
using namespace prism::ajoin;

struct prism_hash_table {
  std::unique_ptr<Table> t;
};

extern "C" {

prism_status prism_hash_table_build(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                    const uint32_t* key_cols, uint32_t n_keys, prism_hash_table** out) {
  return guarded([&] {
    if (!ctx || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or output");
    *out = nullptr;
    Lane lane(ctx);
    HostBatch b = import_batch(batch, schema);
    auto h = std::make_unique<prism_hash_table>();
    h->t = build_table(ctx, b, key_cols, n_keys);
    *out = h.release();
  });
}

prism_status prism_hash_table_probe(prism_ctx* ctx, prism_hash_table* table, const struct ArrowArray* probe,
                                    const struct ArrowSchema* probe_schema, const uint32_t* probe_keys, uint32_t n_keys,
                                    struct ArrowArray* out_probe_idx, struct ArrowSchema* out_probe_schema,
                                    struct ArrowArray* out_build_idx, struct ArrowSchema* out_build_schema) {
  return guarded([&] {
    if (!ctx || !table) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or table");
    Lane lane(ctx);
    HostBatch p = import_batch(probe, probe_schema);
    Pairs m = probe_table(ctx, *table->t, p, probe_keys, n_keys);
    OutColumn a = index_array(ctx, m.p, m.n, "probe_indices");
    OutColumn b = index_array(ctx, m.b, m.n, "build_indices");
    export_array(std::move(a), out_probe_idx, out_probe_schema);
    export_array(std::move(b), out_build_idx, out_build_schema);
  });
}

void prism_hash_table_free(prism_hash_table* table) { delete table; }

prism_status prism_hash_join(prism_ctx* ctx, const struct ArrowArray* probe, const struct ArrowSchema* probe_schema,
                             const struct ArrowArray* build, const struct ArrowSchema* build_schema,
                             const prism_hash_join_config* cfg, struct ArrowArray* out, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!ctx || !cfg) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or config");
    check_join_type(cfg->join_type);
    Lane lane(ctx);
    HostBatch p = import_batch(probe, probe_schema);
    HostBatch b = import_batch(build, build_schema);
    check_left_pad(cfg->join_type, b, p.n_rows);
    std::unique_ptr<Table> T = build_table(ctx, b, cfg->build_keys, cfg->n_build_keys);
    join_one(ctx, cfg->join_type, *T, p, b, cfg->probe_keys, cfg->n_probe_keys, out, out_schema);
  });
}

prism_status prism_probe_matches_chunked(prism_ctx* ctx, const struct ArrowArray* const* probe_batches,
                                         uint32_t n_batches, const struct ArrowSchema* probe_schema,
                                         const struct ArrowArray* build, const struct ArrowSchema* build_schema,
                                         const uint32_t* probe_keys, const uint32_t* build_keys, uint32_t n_keys,
                                         struct ArrowArray* out_probe_idx, struct ArrowArray* out_build_idx,
                                         struct ArrowSchema* out_idx_schema) {
  return guarded([&] {
    if (!ctx || (n_batches && (!probe_batches || !out_probe_idx || !out_build_idx)))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL inputs/outputs");
    Lane lane(ctx);
    HostBatch b = import_batch(build, build_schema);
    std::unique_ptr<Table> T = build_table(ctx, b, build_keys, n_keys);
    std::vector<HostBatch> ps;
    for (uint32_t i = 0; i < n_batches; ++i) ps.push_back(import_batch(probe_batches[i], probe_schema));
    uint32_t done = 0;
    try {
      for (uint32_t i = 0; i < n_batches; ++i) {
        Pairs m = probe_table(ctx, *T, ps[i], probe_keys, n_keys);
        OutColumn a = index_array(ctx, m.p, m.n, "probe_indices");
        OutColumn bb = index_array(ctx, m.b, m.n, "build_indices");
        ArrowSchema s1, s2;
        export_array(std::move(a), &out_probe_idx[i], &s1);
        export_array(std::move(bb), &out_build_idx[i], &s2);
        s2.release(&s2);
        if (out_idx_schema && i == 0) *out_idx_schema = s1;
        else s1.release(&s1);
        ++done;
      }
    } catch (...) {
      for (uint32_t i = 0; i < done; ++i) {
        out_probe_idx[i].release(&out_probe_idx[i]);
        out_build_idx[i].release(&out_build_idx[i]);
      }
      if (out_idx_schema && done && out_idx_schema->release) out_idx_schema->release(out_idx_schema);
      throw;
    }
  });
}

prism_status prism_hash_join_probe_chunked(prism_ctx* ctx, const struct ArrowArray* const* probe_batches,
                                           uint32_t n_batches, const struct ArrowSchema* probe_schema,
                                           const struct ArrowArray* build, const struct ArrowSchema* build_schema,
                                           const prism_hash_join_config* cfg, struct ArrowArray* out_batches,
                                           struct ArrowSchema* out_schema, uint32_t* n_out) {
  return guarded([&] {
    if (!ctx || !cfg || !n_out || (n_batches && (!probe_batches || !out_batches)))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL inputs/outputs");
    check_join_type(cfg->join_type);
    Lane lane(ctx);
    *n_out = 0;
    HostBatch b = import_batch(build, build_schema);
    std::vector<HostBatch> ps;
    int64_t probe_rows = 0;
    for (uint32_t i = 0; i < n_batches; ++i) {
      ps.push_back(import_batch(probe_batches[i], probe_schema));
      probe_rows += ps.back().n_rows;
    }
    check_left_pad(cfg->join_type, b, probe_rows);
    std::unique_ptr<Table> T = build_table(ctx, b, cfg->build_keys, cfg->n_build_keys);
    // Right/Full cannot be chunked and fall back to per-batch inner (hash_join.rs:261-270).
    const int32_t jt = (cfg->join_type == PRISM_JOIN_RIGHT || cfg->join_type == PRISM_JOIN_FULL) ? PRISM_JOIN_INNER
                                                                                                    : cfg->join_type;
    uint32_t k = 0;
    bool have_schema = false;
    try {
      for (uint32_t i = 0; i < n_batches; ++i) {
        ArrowSchema s;
        join_one(ctx, jt, *T, ps[i], b, cfg->probe_keys, cfg->n_probe_keys, &out_batches[k], &s);
        if (out_batches[k].length == 0) {  // empty outputs are dropped (hash_join.rs:274-280)
          out_batches[k].release(&out_batches[k]);
          s.release(&s);
          continue;
        }
        if (out_schema && !have_schema) {
          *out_schema = s;
          have_schema = true;
        } else {
          s.release(&s);
        }
        ++k;
      }
    } catch (...) {
      for (uint32_t i = 0; i < k; ++i) out_batches[i].release(&out_batches[i]);
      if (have_schema) out_schema->release(out_schema);
      throw;
    }
    *n_out = k;
  });
}

}  // extern "C"
