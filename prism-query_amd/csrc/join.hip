// Hash join build + probe on gfx950 for the int64 single-key path
// (native/prism-executor/src/hash_join.rs:87-200).
//
// Build: one of two table layouts, both keeping each key's build rows in
// ascending row order as the reference's serial partition pass does
// (hash_join.rs:116-118):
//   DENSE  - direct-address array over [min_key, max_key] when that range is
//            at most a few times the row count (4 B/entry for unique keys);
//   HASHED - open addressing over hash_i64_fast(key) (hash_join.rs:371-373,
//            a bijection on u64, so no separate key check is needed beyond the
//            stored key), linear probing, load factor <= 1/2, 16 B slots.
// Duplicate keys: a CSR row list per key built with a stable radix sort.
//
// Probe: one 2048-row tile per workgroup, claimed through an atomic ticket so
// tiles start in order; probe keys stream in with 128-bit non-temporal loads,
// wave64 ballots + popcounts rank the matches inside the tile, and a
// decoupled look-back over per-tile status words gives the tile's output
// offset in one pass.  Pairs therefore come out in the reference's order
// (probe row ascending, then build row ascending).
#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace prism {
namespace join {

constexpr int64_t kEmpty = static_cast<int64_t>(0x8000000000000000ULL);  // INT64_MIN
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct Slot {
  int64_t key;
  uint32_t val;  // unique: build row; duplicates: CSR start
  uint32_t cnt;  // rows of this key
};
static_assert(sizeof(Slot) == 16, "slot is 16 bytes");

struct Stats {
  int64_t min_key;
  int64_t max_key;
  unsigned long long n_distinct;
  unsigned long long max_dup;
};

struct TableView {
  const Slot* slots;        // HASHED: mask + 2 entries; entry mask+1 holds the INT64_MIN key
  uint64_t mask;
  int shift;
  const uint32_t* dense_u;  // DENSE unique: row per key offset, kNone = absent
  const uint2* dense_nu;    // DENSE duplicates: {start, cnt}
  int64_t min_key;
  uint64_t range;
  const uint32_t* rows;     // CSR rows (duplicates)
};

}  // namespace join
}  // namespace prism

struct prism_join_table {
  prism::Ctx* ctx = nullptr;
  prism_join_table_info info{};
  prism::DeviceBuffer slots;  // HASHED
  prism::DeviceBuffer dense;  // DENSE
  prism::DeviceBuffer rows;   // CSR rows
  prism::join::TableView view{};
};

namespace prism {
namespace join {

constexpr int kBuildBlock = 256;

__global__ void k_minmax(const int64_t* __restrict__ keys, uint64_t n, Stats* st) {
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    int64_t k = keys[i];
    mn = k < mn ? k : mn;
    mx = k > mx ? k : mx;
  }
  for (int d = 32; d > 0; d >>= 1) {
    int64_t a = __shfl_xor(mn, d, 64), b = __shfl_xor(mx, d, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(reinterpret_cast<long long*>(&st->min_key), static_cast<long long>(mn));
    atomicMax(reinterpret_cast<long long*>(&st->max_key), static_cast<long long>(mx));
  }
}

__global__ void k_init_stats(Stats* st) {
  st->min_key = INT64_MAX;
  st->max_key = INT64_MIN;
  st->n_distinct = 0;
  st->max_dup = 0;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_init_slots(Slot* s, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    s[i].key = kEmpty;
    s[i].val = 0;
    s[i].cnt = 0;
  }
}

// DENSE: count rows per key offset; remember each row's offset.
__global__ void k_dense_count(const int64_t* __restrict__ keys, uint64_t n, int64_t min_key,
                              uint32_t* cnt, uint32_t* slot_of_row) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t d = static_cast<uint32_t>(static_cast<uint64_t>(keys[i]) - static_cast<uint64_t>(min_key));
    slot_of_row[i] = d;
    atomicAdd(&cnt[d], 1u);
  }
}

// HASHED: claim a slot per distinct key (CAS on the key word), count rows.
__global__ void k_hashed_insert(const int64_t* __restrict__ keys, uint64_t n, Slot* slots, uint64_t mask,
                                int shift, uint32_t* slot_of_row) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    int64_t k = keys[i];
    uint64_t s;
    if (k == kEmpty) {
      s = mask + 1;
    } else {
      s = hash_i64_fast(k) >> shift;
      while (true) {
        unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long*>(&slots[s].key),
                                           static_cast<unsigned long long>(kEmpty),
                                           static_cast<unsigned long long>(k));
        if (old == static_cast<unsigned long long>(kEmpty) || old == static_cast<unsigned long long>(k)) break;
        s = (s + 1) & mask;
      }
    }
    slot_of_row[i] = static_cast<uint32_t>(s);
    atomicAdd(&slots[s].cnt, 1u);
  }
}

__global__ void k_count_stats_u32(const uint32_t* __restrict__ cnt, uint64_t n, Stats* st) {
  unsigned long long nd = 0, mx = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = cnt[i];
    nd += c ? 1 : 0;
    mx = c > mx ? c : mx;
  }
  for (int d = 32; d > 0; d >>= 1) {
    nd += __shfl_xor(nd, d, 64);
    unsigned long long o = __shfl_xor(mx, d, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    if (nd) atomicAdd(&st->n_distinct, nd);
    atomicMax(&st->max_dup, mx);
  }
}

__global__ void k_slot_counts(const Slot* __restrict__ slots, uint64_t n, uint32_t* cnt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    cnt[i] = slots[i].cnt;
}

__global__ void k_slot_starts(Slot* slots, uint64_t n, const uint32_t* __restrict__ start) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    slots[i].val = start[i];
}

__global__ void k_hashed_unique_rows(Slot* slots, const uint32_t* __restrict__ slot_of_row, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    slots[slot_of_row[i]].val = static_cast<uint32_t>(i);
}

__global__ void k_dense_unique_rows(uint32_t* dense, const uint32_t* __restrict__ slot_of_row, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dense[slot_of_row[i]] = static_cast<uint32_t>(i);
}

__global__ void k_dense_pack(uint2* dense, const uint32_t* __restrict__ start, const uint32_t* __restrict__ cnt,
                             uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dense[i] = make_uint2(start[i], cnt[i]);
}

__global__ void k_iota_u32(uint32_t* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = static_cast<uint32_t>(i);
}

static dim3 grid_for(prism_ctx* ctx, uint64_t n) {
  uint64_t blocks = ceil_div(n ? n : 1, kBuildBlock);
  uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 8;
  return dim3(static_cast<unsigned>(blocks < cap ? blocks : cap));
}

// Stable CSR: sort (slot, row) pairs by slot (LSD radix sort keeps rows
// ascending within a slot), then starts = exclusive scan of counts.
static void build_csr(prism_ctx* ctx, uint32_t* slot_of_row, uint64_t n, const uint32_t* cnt,
                      uint64_t n_slots, DeviceBuffer& rows_out, uint32_t* start_out) {
  Ctx* c = as_ctx(ctx);
  hipStream_t st = ctx->stream;
  DeviceBuffer rows_in(c, n * 4), keys_out(c, n * 4);
  rows_out.alloc(c, n * 4);
  hipLaunchKernelGGL(k_iota_u32, grid_for(ctx, n), dim3(kBuildBlock), 0, st, rows_in.as<uint32_t>(), n);
  PRISM_CHECK_LAUNCH();
  int end_bit = 1;
  while (end_bit < 32 && (1ULL << end_bit) <= n_slots) ++end_bit;
  size_t tmp_sort = 0, tmp_scan = 0;
  PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, slot_of_row, keys_out.as<uint32_t>(),
                                               rows_in.as<uint32_t>(), rows_out.as<uint32_t>(),
                                               static_cast<int>(n), 0, end_bit, st));
  PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, cnt, start_out, static_cast<int>(n_slots), st));
  DeviceBuffer tmp(c, tmp_sort > tmp_scan ? tmp_sort : tmp_scan);
  PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_sort, slot_of_row, keys_out.as<uint32_t>(),
                                               rows_in.as<uint32_t>(), rows_out.as<uint32_t>(),
                                               static_cast<int>(n), 0, end_bit, st));
  PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, tmp_scan, cnt, start_out, static_cast<int>(n_slots), st));
  PRISM_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Probe.
// ---------------------------------------------------------------------------
constexpr int kPB = 256;                  // threads per workgroup (4 waves)
constexpr int kPairs = 4;                 // 16-byte key pairs per thread
constexpr int kChunk = kPB * 2;           // rows per chunk (one pair per thread)
constexpr int kTileRows = kChunk * kPairs;  // 2048 rows per tile
constexpr int kWaves = kPB / 64;

constexpr uint64_t kStAgg = 1ULL << 62;
constexpr uint64_t kStPrefix = 2ULL << 62;
constexpr uint64_t kStMask = (1ULL << 62) - 1;

template <int KIND, bool UNIQUE>
__device__ __forceinline__ bool lookup(const TableView& T, int64_t k, uint32_t& val, uint32_t& cnt) {
  if constexpr (KIND == PRISM_TABLE_DENSE) {
    uint64_t d = static_cast<uint64_t>(k) - static_cast<uint64_t>(T.min_key);
    if (d >= T.range) return false;
    if constexpr (UNIQUE) {
      uint32_t v = T.dense_u[d];
      val = v;
      cnt = 1;
      return v != kNone;
    } else {
      uint2 e = T.dense_nu[d];
      val = e.x;
      cnt = e.y;
      return e.y != 0;
    }
  } else {
    uint64_t s;
    if (k == kEmpty) {
      const Slot e = T.slots[T.mask + 1];
      val = e.val;
      cnt = e.cnt;
      return e.cnt != 0;
    }
    s = hash_i64_fast(k) >> T.shift;
    while (true) {  // load factor <= 1/2: an empty slot always ends the run
      const Slot e = T.slots[s];
      if (e.key == k) {
        val = e.val;
        cnt = e.cnt;
        return true;
      }
      if (e.key == kEmpty) return false;
      s = (s + 1) & T.mask;
    }
  }
}

// Wave-parallel decoupled look-back (executed by one whole wave).  Returns the
// exclusive prefix of `tile`; publishes the tile's inclusive prefix.
__device__ uint64_t lookback(uint64_t* status, uint32_t tile, uint64_t total, uint32_t* err) {
  const uint32_t lane = lane_id();
  if (tile == 0) {
    if (lane == 0) __hip_atomic_store(&status[0], kStPrefix | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (lane == 0) __hip_atomic_store(&status[tile], kStAgg | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t excl = 0;
  int64_t base = static_cast<int64_t>(tile) - 1;
  uint32_t spins = 0;
  while (true) {
    int64_t idx = base - static_cast<int64_t>(lane);
    uint64_t v = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : kStPrefix;
    uint64_t st = v >> 62;
    uint64_t pm = ballot64(st == 2);
    uint64_t zm = ballot64(st == 0);
    int first_p = pm ? __builtin_ctzll(pm) : 64;
    uint64_t before = first_p == 64 ? ~0ULL : ((1ULL << first_p) - 1);
    if (zm & before) {
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint64_t mine = (static_cast<int>(lane) <= first_p) ? (v & kStMask) : 0;
    for (int d = 32; d > 0; d >>= 1) mine += __shfl_xor(mine, d, 64);
    excl += mine;
    if (first_p < 64) break;
    base -= 64;
  }
  if (lane == 0)
    __hip_atomic_store(&status[tile], kStPrefix | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

__device__ __forceinline__ void wave_reduce_i128(__int128& v) {
  for (int d = 32; d > 0; d >>= 1) {
    uint64_t lo = static_cast<uint64_t>(v);
    uint64_t hi = static_cast<uint64_t>(v >> 64);
    uint64_t olo = __shfl_xor(lo, d, 64), ohi = __shfl_xor(hi, d, 64);
    v += static_cast<__int128>((static_cast<unsigned __int128>(ohi) << 64) | olo);
  }
}

__global__ void k_probe_clear(uint64_t* status, uint64_t n_tiles, uint32_t* ticket, prism_probe_result* res) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (; i < n_tiles; i += (uint64_t)gridDim.x * blockDim.x) status[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *ticket = 0;
    res->matches = 0;
    res->probe_sum_lo = 0;
    res->probe_sum_hi = 0;
    res->build_sum_lo = 0;
    res->build_sum_hi = 0;
    res->error = 0;
    res->reserved = 0;
  }
}

template <int KIND, bool UNIQUE, bool VEC>
__global__ __launch_bounds__(kPB) void k_probe(TableView T, const int64_t* __restrict__ keys, uint64_t n,
                                               uint32_t n_tiles, uint32_t probe_base,
                                               const int64_t* __restrict__ ppay, const int32_t* __restrict__ bpay,
                                               uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                               uint64_t capacity, uint64_t* status, uint32_t* ticket,
                                               prism_probe_result* res) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wave_tot[kPairs][kWaves];
  __shared__ uint64_t s_excl;
  __shared__ __int128 s_red[2][kWaves];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  if (t == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t base = static_cast<uint64_t>(tile) * kTileRows;

  int64_t k[kPairs][2];
  bool ok[kPairs][2];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    uint64_t r0 = base + static_cast<uint64_t>(j) * kChunk + 2 * t;
    if (VEC && r0 + 1 < n) {
      const v2i64 v = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(keys + r0));
      k[j][0] = v.x;
      k[j][1] = v.y;
      ok[j][0] = ok[j][1] = true;
    } else {
      ok[j][0] = r0 < n;
      ok[j][1] = r0 + 1 < n;
      k[j][0] = ok[j][0] ? keys[r0] : 0;
      k[j][1] = ok[j][1] ? keys[r0 + 1] : 0;
    }
  }

  uint32_t val[kPairs][2], cnt[kPairs][2];
#pragma unroll
  for (int j = 0; j < kPairs; ++j)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      uint32_t v = 0, c = 0;
      bool m = ok[j][e] && lookup<KIND, UNIQUE>(T, k[j][e], v, c);
      val[j][e] = v;
      cnt[j][e] = m ? c : 0;
    }

  // Rank matches inside the tile: row order is (chunk, thread, element).
  uint64_t in_wave[kPairs];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    if constexpr (UNIQUE) {
      uint64_t b0 = ballot64(cnt[j][0] != 0), b1 = ballot64(cnt[j][1] != 0);
      in_wave[j] = lanes_below(b0) + lanes_below(b1);
      if (lane == 0) s_wave_tot[j][wave] = __popcll(b0) + __popcll(b1);
    } else {
      uint64_t c = static_cast<uint64_t>(cnt[j][0]) + cnt[j][1];
      uint64_t x = c;
      for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
      }
      in_wave[j] = x - c;
      if (lane == 63) s_wave_tot[j][wave] = x;
    }
  }
  __syncthreads();
  uint64_t chunk_base[kPairs], wave_base[kPairs];
  uint64_t tile_total = 0;
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    chunk_base[j] = tile_total;
    uint64_t wb = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      uint64_t v = s_wave_tot[j][w];
      wb += (w < wave) ? v : 0;
      tile_total += v;
    }
    wave_base[j] = wb;
  }
  if (wave == 0) {
    uint64_t ex = lookback(status, tile, tile_total, &res->error);
    if (lane == 0) s_excl = ex;
  }
  __syncthreads();
  const uint64_t gbase = s_excl;

  const bool want_pairs = out_p != nullptr;
  __int128 psum = 0, bsum = 0;
#pragma unroll
  for (int j = 0; j < kPairs; ++j)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      if (cnt[j][e] == 0) continue;
      const uint64_t row = base + static_cast<uint64_t>(j) * kChunk + 2 * t + e;
      uint64_t pos = gbase + chunk_base[j] + wave_base[j] + in_wave[j] + (e ? cnt[j][0] : 0);
      if constexpr (UNIQUE) {
        if (want_pairs && pos < capacity) {
          __builtin_nontemporal_store(static_cast<uint32_t>(row) + probe_base, out_p + pos);
          __builtin_nontemporal_store(val[j][e], out_b + pos);
        }
        if (ppay) psum += ppay[row];
        if (bpay) bsum += bpay[val[j][e]];
      } else {
        const uint32_t c = cnt[j][e];
        const uint32_t start = val[j][e];
        for (uint32_t r = 0; r < c; ++r) {
          uint32_t brow = T.rows[start + r];
          if (want_pairs && pos + r < capacity) {
            out_p[pos + r] = static_cast<uint32_t>(row) + probe_base;
            out_b[pos + r] = brow;
          }
          if (bpay) bsum += bpay[brow];
        }
        if (ppay) psum += static_cast<__int128>(ppay[row]) * static_cast<__int128>(c);
      }
    }

  if (ppay || bpay) {
    wave_reduce_i128(psum);
    wave_reduce_i128(bsum);
    if (lane == 0) {
      s_red[0][wave] = psum;
      s_red[1][wave] = bsum;
    }
    __syncthreads();
    if (t == 0) {
      __int128 p = 0, b = 0;
      for (int w = 0; w < kWaves; ++w) {
        p += s_red[0][w];
        b += s_red[1][w];
      }
      if (p != 0)
        atomic_add_i128(reinterpret_cast<unsigned long long*>(&res->probe_sum_lo),
                        reinterpret_cast<unsigned long long*>(&res->probe_sum_hi), static_cast<uint64_t>(p),
                        static_cast<int64_t>(p >> 64));
      if (b != 0)
        atomic_add_i128(reinterpret_cast<unsigned long long*>(&res->build_sum_lo),
                        reinterpret_cast<unsigned long long*>(&res->build_sum_hi), static_cast<uint64_t>(b),
                        static_cast<int64_t>(b >> 64));
    }
  }
  if (tile == n_tiles - 1 && t == 0) res->matches = gbase + tile_total;
}

template <int KIND, bool UNIQUE>
static void launch_probe_vec(prism_ctx* ctx, bool vec, dim3 g, const TableView& T, const int64_t* keys,
                             uint64_t n, uint32_t n_tiles, uint32_t probe_base, const int64_t* ppay,
                             const int32_t* bpay, uint32_t* op, uint32_t* ob, uint64_t cap, uint64_t* status,
                             uint32_t* ticket, prism_probe_result* res) {
  if (vec)
    hipLaunchKernelGGL((k_probe<KIND, UNIQUE, true>), g, dim3(kPB), 0, ctx->stream, T, keys, n, n_tiles,
                       probe_base, ppay, bpay, op, ob, cap, status, ticket, res);
  else
    hipLaunchKernelGGL((k_probe<KIND, UNIQUE, false>), g, dim3(kPB), 0, ctx->stream, T, keys, n, n_tiles,
                       probe_base, ppay, bpay, op, ob, cap, status, ticket, res);
}

}  // namespace join
}  // namespace prism

using namespace prism;
using namespace prism::join;

extern "C" {

prism_status prism_join_build_i64(prism_ctx* ctx, const int64_t* d_keys, uint64_t n_rows, int32_t kind,
                                  prism_join_table** out) {
  return guarded([&] {
    if (!ctx || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or output");
    *out = nullptr;
    if (n_rows >= 0xFFFFFFFFULL)
      fail(PRISM_ERR_JOIN, "Join error: build side exceeds u32 row indices (UInt32Array, hash_join.rs:68-69)");
    if (n_rows && !d_keys) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL build keys");
    if (kind != PRISM_TABLE_AUTO && kind != PRISM_TABLE_HASHED && kind != PRISM_TABLE_DENSE)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown table kind");
    use_device(ctx);
    Ctx* c = as_ctx(ctx);
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t st = ctx->stream;
    auto* T = new prism_join_table();
    std::unique_ptr<prism_join_table> guard(T);
    T->ctx = c;
    T->info.n_rows = n_rows;

    DeviceBuffer sbuf(c, sizeof(Stats));
    Stats* dst = sbuf.as<Stats>();
    hipLaunchKernelGGL(k_init_stats, dim3(1), dim3(1), 0, st, dst);
    if (n_rows) hipLaunchKernelGGL(k_minmax, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows, dst);
    PRISM_CHECK_LAUNCH();
    Stats hs;
    PRISM_HIP(hipMemcpyAsync(&hs, dst, sizeof(Stats), hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    if (n_rows == 0) hs.min_key = hs.max_key = 0;
    const uint64_t range = static_cast<uint64_t>(hs.max_key) - static_cast<uint64_t>(hs.min_key) + 1;
    bool dense;
    if (kind == PRISM_TABLE_DENSE) {
      if (range == 0 || range > (1ULL << 31))
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: key range too wide for a dense table");
      dense = true;
    } else if (kind == PRISM_TABLE_HASHED) {
      dense = false;
    } else {
      dense = range != 0 && range <= 4 * n_rows + 1024 && range <= (1ULL << 31);
    }
    T->info.min_key = hs.min_key;
    T->info.max_key = hs.max_key;

    DeviceBuffer slot_of_row(c, (n_rows ? n_rows : 1) * 4);
    if (dense) {
      T->info.kind = PRISM_TABLE_DENSE;
      T->info.slots = range;
      DeviceBuffer cnt(c, range * 4);
      hipLaunchKernelGGL(k_fill_u32, grid_for(ctx, range), dim3(kBuildBlock), 0, st, cnt.as<uint32_t>(), range, 0u);
      if (n_rows)
        hipLaunchKernelGGL(k_dense_count, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows,
                           hs.min_key, cnt.as<uint32_t>(), slot_of_row.as<uint32_t>());
      hipLaunchKernelGGL(k_init_stats, dim3(1), dim3(1), 0, st, dst);
      hipLaunchKernelGGL(k_count_stats_u32, grid_for(ctx, range), dim3(kBuildBlock), 0, st, cnt.as<uint32_t>(),
                         range, dst);
      PRISM_CHECK_LAUNCH();
      PRISM_HIP(hipMemcpyAsync(&hs, dst, sizeof(Stats), hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      T->info.n_distinct = hs.n_distinct;
      T->info.max_dup = hs.max_dup;
      T->info.unique = hs.max_dup <= 1;
      if (T->info.unique) {
        T->dense = std::move(cnt);  // reuse as the row array
        uint32_t* d = T->dense.as<uint32_t>();
        hipLaunchKernelGGL(k_fill_u32, grid_for(ctx, range), dim3(kBuildBlock), 0, st, d, range, kNone);
        if (n_rows)
          hipLaunchKernelGGL(k_dense_unique_rows, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d,
                             slot_of_row.as<uint32_t>(), n_rows);
        PRISM_CHECK_LAUNCH();
        T->view.dense_u = d;
      } else {
        DeviceBuffer start(c, range * 4);
        build_csr(ctx, slot_of_row.as<uint32_t>(), n_rows, cnt.as<uint32_t>(), range, T->rows,
                  start.as<uint32_t>());
        T->dense.alloc(c, range * 8);
        hipLaunchKernelGGL(k_dense_pack, grid_for(ctx, range), dim3(kBuildBlock), 0, st, T->dense.as<uint2>(),
                           start.as<uint32_t>(), cnt.as<uint32_t>(), range);
        PRISM_CHECK_LAUNCH();
        T->view.dense_nu = T->dense.as<uint2>();
        T->view.rows = T->rows.as<uint32_t>();
      }
      T->view.min_key = hs.min_key;
      T->view.range = range;
    } else {
      T->info.kind = PRISM_TABLE_HASHED;
      uint64_t S = next_pow2(n_rows * 2 > 64 ? n_rows * 2 : 64);
      T->info.slots = S;
      const int shift = 64 - log2_pow2(S);
      T->slots.alloc(c, (S + 1) * sizeof(Slot));
      Slot* slots = T->slots.as<Slot>();
      hipLaunchKernelGGL(k_init_slots, grid_for(ctx, S + 1), dim3(kBuildBlock), 0, st, slots, S + 1);
      if (n_rows)
        hipLaunchKernelGGL(k_hashed_insert, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows,
                           slots, S - 1, shift, slot_of_row.as<uint32_t>());
      DeviceBuffer cnt(c, (S + 1) * 4);
      hipLaunchKernelGGL(k_slot_counts, grid_for(ctx, S + 1), dim3(kBuildBlock), 0, st, slots, S + 1,
                         cnt.as<uint32_t>());
      hipLaunchKernelGGL(k_init_stats, dim3(1), dim3(1), 0, st, dst);
      hipLaunchKernelGGL(k_count_stats_u32, grid_for(ctx, S + 1), dim3(kBuildBlock), 0, st, cnt.as<uint32_t>(),
                         S + 1, dst);
      PRISM_CHECK_LAUNCH();
      PRISM_HIP(hipMemcpyAsync(&hs, dst, sizeof(Stats), hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      T->info.n_distinct = hs.n_distinct;
      T->info.max_dup = hs.max_dup;
      T->info.unique = hs.max_dup <= 1;
      if (T->info.unique) {
        if (n_rows)
          hipLaunchKernelGGL(k_hashed_unique_rows, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, slots,
                             slot_of_row.as<uint32_t>(), n_rows);
        PRISM_CHECK_LAUNCH();
      } else {
        DeviceBuffer start(c, (S + 1) * 4);
        build_csr(ctx, slot_of_row.as<uint32_t>(), n_rows, cnt.as<uint32_t>(), S + 1, T->rows,
                  start.as<uint32_t>());
        hipLaunchKernelGGL(k_slot_starts, grid_for(ctx, S + 1), dim3(kBuildBlock), 0, st, slots, S + 1,
                           start.as<uint32_t>());
        PRISM_CHECK_LAUNCH();
        T->view.rows = T->rows.as<uint32_t>();
      }
      T->view.slots = slots;
      T->view.mask = S - 1;
      T->view.shift = shift;
    }
    PRISM_HIP(hipStreamSynchronize(st));
    T->info.device_bytes = T->slots.bytes + T->dense.bytes + T->rows.bytes;
    *out = guard.release();
  });
}

prism_status prism_join_table_info_get(const prism_join_table* t, prism_join_table_info* out) {
  return guarded([&] {
    if (!t || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL table or output");
    *out = t->info;
  });
}

void prism_join_table_free(prism_join_table* t) {
  if (!t) return;
  if (t->ctx) (void)hipSetDevice(t->ctx->device);
  delete t;
}

prism_status prism_join_probe_i64(prism_ctx* ctx, const prism_join_table* t, const int64_t* d_probe_keys,
                                  uint64_t n_probe, uint32_t probe_row_base, const int64_t* d_probe_payload,
                                  const int32_t* d_build_payload, uint32_t* d_probe_idx, uint32_t* d_build_idx,
                                  uint64_t capacity, prism_probe_result* d_result) {
  return guarded([&] {
    if (!ctx || !t || !d_result) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/table/result");
    if ((d_probe_idx == nullptr) != (d_build_idx == nullptr))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: give both index outputs or neither");
    if (n_probe && !d_probe_keys) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL probe keys");
    if (n_probe + probe_row_base > 0xFFFFFFFFULL)
      fail(PRISM_ERR_JOIN, "Join error: probe rows exceed u32 row indices (UInt32Array, hash_join.rs:68-69)");
    use_device(ctx);
    Ctx* c = as_ctx(ctx);
    std::lock_guard<std::mutex> lock(ctx->mu);
    const uint64_t n_tiles = ceil_div(n_probe, kTileRows);
    c->scratch.reserve(c, 256 + (n_tiles + 1) * 8);
    uint32_t* ticket = c->scratch.as<uint32_t>();
    uint64_t* status = reinterpret_cast<uint64_t*>(c->scratch.as<char>() + 256);
    hipLaunchKernelGGL(k_probe_clear, grid_for(ctx, n_tiles + 1), dim3(kBuildBlock), 0, ctx->stream, status,
                       n_tiles, ticket, d_result);
    PRISM_CHECK_LAUNCH();
    if (n_probe == 0) return;
    const bool vec = (reinterpret_cast<uintptr_t>(d_probe_keys) & 15) == 0;
    const TableView& T = t->view;
    dim3 g(static_cast<unsigned>(n_tiles));
    const uint32_t nt = static_cast<uint32_t>(n_tiles);
    if (t->info.kind == PRISM_TABLE_DENSE) {
      if (t->info.unique)
        launch_probe_vec<PRISM_TABLE_DENSE, true>(ctx, vec, g, T, d_probe_keys, n_probe, nt, probe_row_base,
                                                  d_probe_payload, d_build_payload, d_probe_idx, d_build_idx,
                                                  capacity, status, ticket, d_result);
      else
        launch_probe_vec<PRISM_TABLE_DENSE, false>(ctx, vec, g, T, d_probe_keys, n_probe, nt, probe_row_base,
                                                   d_probe_payload, d_build_payload, d_probe_idx, d_build_idx,
                                                   capacity, status, ticket, d_result);
    } else {
      if (t->info.unique)
        launch_probe_vec<PRISM_TABLE_HASHED, true>(ctx, vec, g, T, d_probe_keys, n_probe, nt, probe_row_base,
                                                   d_probe_payload, d_build_payload, d_probe_idx, d_build_idx,
                                                   capacity, status, ticket, d_result);
      else
        launch_probe_vec<PRISM_TABLE_HASHED, false>(ctx, vec, g, T, d_probe_keys, n_probe, nt, probe_row_base,
                                                    d_probe_payload, d_build_payload, d_probe_idx, d_build_idx,
                                                    capacity, status, ticket, d_result);
    }
    PRISM_CHECK_LAUNCH();
  });
}

}  // extern "C"
