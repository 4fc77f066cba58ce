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
// Probe: one 2048-row tile per workgroup (tile = workgroup index); probe keys
// stream in with 128-bit non-temporal loads, wave64 ballots + popcounts rank
// the matches inside the tile, and a decoupled look-back over per-tile status
// words gives the tile's output offset in one pass.  The look-back never waits
// unboundedly: a predecessor that has not published is counted by the waiter.  Pairs therefore come out in the reference's order
// (probe row ascending, then build row ascending).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "join.hpp"

namespace prism {
namespace join {

struct Stats {
  int64_t min_key;
  int64_t max_key;
  unsigned long long n_distinct;
  unsigned long long max_dup;
  unsigned long long sentinel;  // HASHED: build rows holding INT64_MIN
  unsigned long long dups;      // DENSE speculative build: second claims of an entry
};


}  // namespace join
}  // namespace prism


namespace prism {
namespace join {

constexpr int kBuildBlock = 256;

// Reductions run in two stages (per-block partials, then one block) so that
// no global address takes more than one atomic per launch.
constexpr int kRedBlocks = 1024;

template <class T, class Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T* lds) {
  for (int d = 32; d > 0; d >>= 1) v = op(v, __shfl_xor(v, d, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  T r = lds[0];
  for (int i = 1; i < static_cast<int>(blockDim.x >> 6); ++i) r = op(r, lds[i]);
  __syncthreads();
  return r;
}

struct MinOp { __device__ int64_t operator()(int64_t a, int64_t b) const { return a < b ? a : b; } };
struct MaxOp { __device__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; } };
struct AddOp { __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a + b; } };
struct UMaxOp { __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a > b ? a : b; } };

// Also clears the speculative dense table (fill_n entries of fill_u32, or of
// fill_up when a payload rides in the entries) in the same pass.
__global__ void k_minmax_part(const int64_t* __restrict__ keys, uint64_t n, int64_t* __restrict__ part,
                              uint32_t* fill_u32, uint2* fill_up, uint64_t fill_n) {
  __shared__ int64_t lds[16];
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < fill_n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (fill_up) fill_up[i] = make_uint2(kNone, 0u);
    else fill_u32[i] = kNone;
  }
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    int64_t k = keys[i];
    mn = k < mn ? k : mn;
    mx = k > mx ? k : mx;
  }
  mn = block_reduce(mn, MinOp(), lds);
  mx = block_reduce(mx, MaxOp(), lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = mn;
    part[2 * blockIdx.x + 1] = mx;
  }
}

__global__ void k_minmax_final(const int64_t* __restrict__ part, int n_part, Stats* st) {
  __shared__ int64_t lds[16];
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) {
    mn = part[2 * i] < mn ? part[2 * i] : mn;
    mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
  }
  mn = block_reduce(mn, MinOp(), lds);
  mx = block_reduce(mx, MaxOp(), lds);
  if (threadIdx.x == 0) {  // every field: no separate initialisation
    st->min_key = mn;
    st->max_key = mx;
    st->n_distinct = 0;
    st->max_dup = 0;
    st->sentinel = 0;
    st->dups = 0;
  }
}

__global__ void k_init_stats(Stats* st) {
  st->min_key = INT64_MAX;
  st->max_key = INT64_MIN;
  st->n_distinct = 0;
  st->max_dup = 0;
  st->sentinel = 0;
  st->dups = 0;
}

__global__ void k_copy_sentinel(const uint32_t* cnt_of_sentinel, Stats* st) { st->sentinel = *cnt_of_sentinel; }

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

// DENSE: count rows per key offset; remember each row's offset.  Counts
// gather in a workgroup LDS table first (skewed keys would otherwise
// serialise on one counter); offsets it cannot place are counted in HBM.
// With `sample_shift` > 0 only every 2^sample_shift-th run of kCountRun rows
// is counted (and no row offsets are written).
constexpr int kCountSlots = 4096, kCountProbes = 4, kCountRun = 256;
__global__ __launch_bounds__(256) void k_dense_count(const int64_t* __restrict__ keys, uint64_t n, int64_t min_key,
                                                     uint32_t* cnt, uint32_t* slot_of_row, int sample_shift = 0) {
  __shared__ uint32_t sk[kCountSlots], sc[kCountSlots];
  for (int i = threadIdx.x; i < kCountSlots; i += blockDim.x) {
    sk[i] = kNone;
    sc[i] = 0;
  }
  __syncthreads();
  // sampled rows: every 2^sample_shift-th run of kCountRun rows
  const uint64_t ns = sample_shift ? ceil_div(n, static_cast<uint64_t>(kCountRun) << sample_shift) * kCountRun : n;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < ns; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = ((j / kCountRun) << sample_shift) * kCountRun + j % kCountRun;
    if (i >= n) continue;
    const uint32_t d = static_cast<uint32_t>(static_cast<uint64_t>(keys[i]) - static_cast<uint64_t>(min_key));
    if (slot_of_row) slot_of_row[i] = d;
    uint32_t h = static_cast<uint32_t>((static_cast<uint64_t>(d) * 0x9E3779B97F4A7C15ULL) >> 52);  // 12 bits
    bool placed = false;
    for (int p = 0; p < kCountProbes && d != kNone; ++p) {
      const uint32_t cur = __hip_atomic_load(&sk[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (cur == kNone) {
        const uint32_t old = atomicCAS(&sk[h], kNone, d);
        if (old == kNone || old == d) {
          atomicAdd(&sc[h], 1u);
          placed = true;
          break;
        }
      } else if (cur == d) {
        atomicAdd(&sc[h], 1u);
        placed = true;
        break;
      }
      h = (h + 1) & (kCountSlots - 1);
    }
    if (!placed) atomicAdd(&cnt[d], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kCountSlots; i += blockDim.x)
    if (sk[i] != kNone) atomicAdd(&cnt[sk[i]], sc[i]);
}

// DENSE key marks: map[d] = 1 for every key offset d of a build row (plain
// byte stores: every writer stores the same value).
__global__ __launch_bounds__(256) void k_dense_mark(const int64_t* __restrict__ keys, uint64_t n, int64_t min_key,
                                                    uint8_t* map) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    map[static_cast<uint64_t>(keys[i]) - static_cast<uint64_t>(min_key)] = 1;
}

// Byte marks (0 / 1, 32 per word) packed into bits: bit j of word w = map[32 w + j].
__global__ __launch_bounds__(256) void k_pack_marks(const uint4* __restrict__ map, uint64_t n_words, uint32_t* bits) {
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < n_words; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 a = map[2 * w], b = map[2 * w + 1];
    const uint32_t x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      r |= ((x[k] & 1u) | ((x[k] >> 7) & 2u) | ((x[k] >> 14) & 4u) | ((x[k] >> 21) & 8u)) << (4 * k);
    bits[w] = r;
  }
}

// Rows per partition of a partitioned HASHED table (partition = top hash bits).
__global__ void k_pt_count(const int64_t* __restrict__ keys, uint64_t n, int pshift, uint32_t* __restrict__ cnt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[hash_i64_fast(keys[i]) >> pshift], 1u);
}

// HASHED: claim a slot per distinct key (CAS on the key word), count rows.
__global__ void k_hashed_insert(const int64_t* __restrict__ keys, uint64_t n, Slot* slots, uint64_t mask,
                                uint64_t pmask, int shift, uint32_t* slot_of_row) {
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
        s = (s & ~pmask) | ((s + 1) & pmask);  // wraps inside the slot's partition
      }
    }
    slot_of_row[i] = static_cast<uint32_t>(s);
    atomicAdd(&slots[s].cnt, 1u);
  }
}

__global__ void k_count_part(const uint32_t* __restrict__ cnt, uint64_t n, uint64_t* __restrict__ part) {
  __shared__ uint64_t lds[16];
  uint64_t nd = 0, mx = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = cnt[i];
    nd += c ? 1 : 0;
    mx = c > mx ? c : mx;
  }
  nd = block_reduce(nd, AddOp(), lds);
  mx = block_reduce(mx, UMaxOp(), lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = nd;
    part[2 * blockIdx.x + 1] = mx;
  }
}

__global__ void k_count_final(const uint64_t* __restrict__ part, int n_part, Stats* st) {
  __shared__ uint64_t lds[16];
  uint64_t nd = 0, mx = 0;
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) {
    nd += part[2 * i];
    mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
  }
  nd = block_reduce(nd, AddOp(), lds);
  mx = block_reduce(mx, UMaxOp(), lds);
  if (threadIdx.x == 0) {
    st->n_distinct = nd;
    st->max_dup = mx;
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

// DENSE build speculating unique keys: entries start absent ({kNone, 0} in the
// payload form), every row stores its index (and payload) into its key's
// entry, then a second pass checks that each row's entry holds that row.  With
// unique keys every entry has one writer; a duplicate key leaves some row
// whose entry holds another row, which counts a duplicate (the table is then
// built the general way).  Plain stores: no atomics on the table.
// Key range from the device statistics; 0 when it exceeds `cap` entries.
__device__ __forceinline__ uint64_t dense_range(const Stats* st, uint64_t cap) {
  const uint64_t r = static_cast<uint64_t>(st->max_key) - static_cast<uint64_t>(st->min_key) + 1;  // 0: 2^64 wrapped
  return (st->max_key >= st->min_key && r != 0 && r <= cap) ? r : 0;
}
// The speculative build keeps one array: `dense` (row per key) without a
// payload, `dense_up` ({row, payload} per key) with one (dense is then NULL);
// k_minmax_part clears it.
__global__ void k_dense_scatter_dev(const int64_t* __restrict__ keys, uint64_t n, uint64_t cap, const Stats* st,
                                    uint32_t* dense, uint2* dense_up, const int32_t* __restrict__ pay) {
  if (dense_range(st, cap) == 0) return;
  const int64_t min_key = st->min_key;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = static_cast<uint32_t>(static_cast<uint64_t>(keys[i]) - static_cast<uint64_t>(min_key));
    if (dense_up) dense_up[d] = make_uint2(static_cast<uint32_t>(i), static_cast<uint32_t>(pay[i]));
    else dense[d] = static_cast<uint32_t>(i);
  }
}
__global__ void k_dense_verify_dev(const int64_t* __restrict__ keys, uint64_t n, uint64_t cap, Stats* st,
                                   const uint32_t* __restrict__ dense, const uint2* __restrict__ dense_up) {
  if (dense_range(st, cap) == 0) return;
  const int64_t min_key = st->min_key;
  uint32_t bad = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = static_cast<uint32_t>(static_cast<uint64_t>(keys[i]) - static_cast<uint64_t>(min_key));
    bad += ((dense_up ? dense_up[d].x : dense[d]) != static_cast<uint32_t>(i)) ? 1u : 0u;
  }
  if (bad) atomicAdd(&st->dups, static_cast<unsigned long long>(bad));
}

__global__ void k_dense_fill2(uint32_t* dense, uint2* dense_up, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    dense[i] = kNone;
    if (dense_up) dense_up[i] = make_uint2(kNone, 0u);
  }
}
__global__ void k_dense_scatter(const int64_t* __restrict__ keys, uint64_t n, int64_t min_key, uint32_t* dense,
                                uint2* dense_up, const int32_t* __restrict__ pay) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = static_cast<uint32_t>(static_cast<uint64_t>(keys[i]) - static_cast<uint64_t>(min_key));
    dense[d] = static_cast<uint32_t>(i);
    if (dense_up) dense_up[d] = make_uint2(static_cast<uint32_t>(i), static_cast<uint32_t>(pay[i]));
  }
}
__global__ void k_dense_verify(const int64_t* __restrict__ keys, uint64_t n, int64_t min_key,
                               const uint32_t* __restrict__ dense, const uint2* __restrict__ dense_up,
                               unsigned long long* dups) {
  uint32_t bad = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = static_cast<uint32_t>(static_cast<uint64_t>(keys[i]) - static_cast<uint64_t>(min_key));
    bad += (dense[d] != static_cast<uint32_t>(i) || (dense_up && dense_up[d].x != static_cast<uint32_t>(i))) ? 1u : 0u;
  }
  if (bad) atomicAdd(dups, static_cast<unsigned long long>(bad));
}

__global__ void k_dense_pack(uint2* dense, const uint32_t* __restrict__ start, const uint32_t* __restrict__ cnt,
                             uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dense[i] = make_uint2(start[i], cnt[i]);
}

__global__ void k_dense_pack_payload(uint2* out, const uint32_t* __restrict__ rows, const int32_t* __restrict__ pay,
                                     uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rows[i];
    out[i] = make_uint2(r, r != kNone ? static_cast<uint32_t>(pay[r]) : 0u);
  }
}

__global__ void k_slot_payload(Slot* slots, uint64_t n, const int32_t* __restrict__ pay) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (slots[i].cnt) slots[i].cnt = static_cast<uint32_t>(pay[slots[i].val]);
}

__global__ void k_gather_i32(int32_t* out, const uint32_t* __restrict__ idx, const int32_t* __restrict__ src,
                             uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = src[idx[i]];
}

__global__ void k_widen_u32(uint64_t* out, const uint32_t* __restrict__ in, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
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

static void launch_minmax(prism_ctx* ctx, const int64_t* keys, uint64_t n, int64_t* part, Stats* st,
                          uint32_t* fill_u32 = nullptr, uint2* fill_up = nullptr, uint64_t fill_n = 0) {
  uint64_t b = ceil_div(n ? n : 1, kBuildBlock);
  int nb = static_cast<int>(b < kRedBlocks ? b : kRedBlocks);
  hipLaunchKernelGGL(k_minmax_part, dim3(nb), dim3(kBuildBlock), 0, ctx->stream, keys, n, part, fill_u32, fill_up,
                     fill_n);
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(kBuildBlock), 0, ctx->stream, part, nb, st);
  PRISM_CHECK_LAUNCH();
}

static void launch_count_stats(prism_ctx* ctx, const uint32_t* cnt, uint64_t n, uint64_t* part, Stats* st) {
  uint64_t b = ceil_div(n ? n : 1, kBuildBlock);
  int nb = static_cast<int>(b < kRedBlocks ? b : kRedBlocks);
  hipLaunchKernelGGL(k_count_part, dim3(nb), dim3(kBuildBlock), 0, ctx->stream, cnt, n, part);
  hipLaunchKernelGGL(k_count_final, dim3(1), dim3(kBuildBlock), 0, ctx->stream, part, nb, st);
  PRISM_CHECK_LAUNCH();
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
constexpr int kPB = 512;                  // threads per workgroup (8 waves)
constexpr int kPairs = 4;                 // 16-byte key pairs per thread
constexpr int kChunk = kPB * 2;           // rows per chunk (one pair per thread)
constexpr int kTileRows = kChunk * kPairs;  // 4096 rows per tile
constexpr int kWaves = kPB / 64;

constexpr uint64_t kStAgg = 1ULL << 62;
constexpr uint64_t kStPrefix = 2ULL << 62;
constexpr uint64_t kStMask = (1ULL << 62) - 1;


// Matches of one tile counted by a whole wave (used when a predecessor tile
// has not published its count yet, so no tile ever waits on another's progress).
template <int KIND, bool UNIQUE>
__device__ uint64_t tile_count(const TableView& T, const int64_t* __restrict__ keys, uint64_t n, uint64_t tile) {
  const uint32_t lane = lane_id();
  uint64_t c = 0;
  const uint64_t base = tile * kTileRows;
  for (uint32_t i = lane; i < static_cast<uint32_t>(kTileRows); i += 64) {
    const uint64_t r = base + i;
    if (r < n) {
      uint32_t v, cn;
      if (lookup<KIND, UNIQUE>(T, keys[r], v, cn)) c += cn;
    }
  }
  for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
  return c;
}

// Block-wide decoupled look-back (every thread of the workgroup takes part:
// one status word per thread, so one memory round trip covers kPB
// predecessors).  Returns the exclusive prefix of `tile` and publishes its
// inclusive prefix.  A predecessor that stays unpublished for kStallPolls
// polls is counted here instead of waited for (and its count published), so
// the scan finishes whatever order the workgroups are scheduled in.
constexpr uint32_t kStallPolls = 256;

struct LookbackLds {
  uint64_t zero[kWaves];
  uint64_t sum[kWaves];
  int first[kWaves];
};

template <int KIND, bool UNIQUE>
__device__ uint64_t lookback(uint64_t* status, uint32_t tile, uint64_t total, const TableView& T,
                             const int64_t* __restrict__ keys, uint64_t n, LookbackLds& L) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (tile == 0) {
    if (t == 0) __hip_atomic_store(&status[0], kStPrefix | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (t == 0) __hip_atomic_store(&status[tile], kStAgg | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t excl = 0;
  int64_t base = static_cast<int64_t>(tile) - 1;
  uint32_t polls = 0;
  while (true) {
    const int64_t idx = base - t;
    const uint64_t v = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : kStPrefix;
    const uint64_t st = v >> 62;
    const uint64_t pm = ballot64(st == 2), zm = ballot64(st == 0);
    if (lane == 0) {
      L.first[wave] = pm ? __builtin_ctzll(pm) : 64;
      L.zero[wave] = zm;
    }
    __syncthreads();
    int fp = kPB;  // thread holding the nearest inclusive prefix
    for (int w = 0; w < kWaves; ++w)
      if (L.first[w] < 64) {
        fp = w * 64 + L.first[w];
        break;
      }
    int zfirst = -1;  // nearest unpublished predecessor before it
    for (int w = 0; w < kWaves && w * 64 < fp && zfirst < 0; ++w) {
      const int lim = fp - w * 64 < 64 ? fp - w * 64 : 64;
      const uint64_t mask = lim == 64 ? ~0ULL : ((1ULL << lim) - 1);
      const uint64_t z = L.zero[w] & mask;
      if (z) zfirst = w * 64 + __builtin_ctzll(z);
    }
    if (zfirst >= 0) {
      __syncthreads();
      if (++polls < kStallPolls) {
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      if (wave == 0) {
        const int64_t q = base - zfirst;
        const uint64_t c = tile_count<KIND, UNIQUE>(T, keys, n, static_cast<uint64_t>(q));
        if (lane == 0) {
          unsigned long long expect = 0;
          __hip_atomic_compare_exchange_strong(reinterpret_cast<unsigned long long*>(&status[q]), &expect,
                                               static_cast<unsigned long long>(kStAgg | c), __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      polls = 0;
      continue;
    }
    uint64_t mine = (t <= fp) ? (v & kStMask) : 0;
    for (int d = 32; d > 0; d >>= 1) mine += __shfl_xor(mine, d, 64);
    if (lane == 0) L.sum[wave] = mine;
    __syncthreads();
    uint64_t sum = 0;
    for (int w = 0; w < kWaves; ++w) sum += L.sum[w];
    excl += sum;
    __syncthreads();
    if (fp < kPB) break;
    base -= kPB;
  }
  if (t == 0)
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

__global__ void k_probe_clear(uint64_t* status, uint64_t n_tiles, prism_probe_result* res) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (; i < n_tiles; i += (uint64_t)gridDim.x * blockDim.x) status[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    res->matches = 0;
    res->probe_sum_lo = 0;
    res->probe_sum_hi = 0;
    res->build_sum_lo = 0;
    res->build_sum_hi = 0;
    res->error = 0;
    res->reserved = 0;
  }
}

// Staging of the probe's matches between the lookup pass and the move pass:
// the build row of each match, in row order at tile * kTileRows, and the
// tile's match ballots, two 64-bit words (element 0, element 1 of each lane's
// row pair) per (chunk, wave), from which the move pass recovers the probe
// rows: 4 B per match + 512 B per tile instead of an 8 B pair per match.
constexpr int kTileBallots = kPairs * kWaves * 2;
struct PairStage {
  uint32_t* brow;
  unsigned long long* bits;
};

__device__ __forceinline__ void stage_ballots(const PairStage& S, uint32_t tile, int j, int wave, uint64_t b0,
                                              uint64_t b1) {
  unsigned long long* w = S.bits + static_cast<uint64_t>(tile) * kTileBallots + (j * kWaves + wave) * 2;
  w[0] = b0;
  w[1] = b1;
}

// Unique-key probe, phase 1: each tile looks its keys up, ranks its matches in
// LDS (row order), writes them contiguously at tile * kTileRows of a staging
// buffer and records its match count and payload partials.  No tile waits on
// another.  Phase 2 scans the counts; phase 3 moves each tile's pairs to their
// final offset with coalesced copies.
template <int KIND, bool VEC>
__global__ __launch_bounds__(kPB) void k_probe_unique(TableView T, const int64_t* __restrict__ keys, uint64_t n,
                                                      const int64_t* __restrict__ ppay,
                                                      const int32_t* __restrict__ bpay, PairStage stage,
                                                      uint32_t* __restrict__ tile_count,
                                                      __int128* __restrict__ tile_sums) {
  __shared__ uint32_t s_b[kTileRows];
  __shared__ uint32_t s_wave_tot[kPairs][kWaves];
  __shared__ __int128 s_red[2][kWaves];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const uint32_t tile = blockIdx.x;
  const uint64_t base = static_cast<uint64_t>(tile) * kTileRows;
  const bool table_pay = T.has_payload && bpay == nullptr;
  const bool sums = ppay || bpay || table_pay;

  uint32_t val[kPairs][2];
  bool hit[kPairs][2];
  __int128 psum = 0, bsum = 0;
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const uint64_t r0 = base + static_cast<uint64_t>(j) * kChunk + 2 * t;
    int64_t k0, k1, p0 = 0, p1 = 0;
    bool ok0, ok1;
    if (VEC && r0 + 1 < n) {
      const v2i64 v = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(keys + r0));
      k0 = v.x;
      k1 = v.y;
      ok0 = ok1 = true;
      if (ppay) {
        const v2i64 p = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(ppay + r0));
        p0 = p.x;
        p1 = p.y;
      }
    } else {
      ok0 = r0 < n;
      ok1 = r0 + 1 < n;
      k0 = ok0 ? keys[r0] : 0;
      k1 = ok1 ? keys[r0 + 1] : 0;
      if (ppay) {
        p0 = ok0 ? ppay[r0] : 0;
        p1 = ok1 ? ppay[r0 + 1] : 0;
      }
    }
    uint32_t v0 = 0, v1 = 0, c;
    int32_t y0 = 0, y1 = 0;
    hit[j][0] = ok0 && lookup<KIND, true>(T, k0, v0, c, y0);
    hit[j][1] = ok1 && lookup<KIND, true>(T, k1, v1, c, y1);
    val[j][0] = v0;
    val[j][1] = v1;
    if (hit[j][0]) {
      psum += p0;
      bsum += bpay ? bpay[v0] : (table_pay ? y0 : 0);
    }
    if (hit[j][1]) {
      psum += p1;
      bsum += bpay ? bpay[v1] : (table_pay ? y1 : 0);
    }
  }
  if (sums) {
    wave_reduce_i128(psum);
    wave_reduce_i128(bsum);
    if (lane == 0) {
      s_red[0][wave] = psum;
      s_red[1][wave] = bsum;
    }
  }
  // Rank matches inside the tile: row order is (chunk, thread, element).
  uint32_t local[kPairs];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const uint64_t b0 = ballot64(hit[j][0]), b1 = ballot64(hit[j][1]);
    local[j] = lanes_below(b0) + lanes_below(b1);
    if (lane == 0) {
      s_wave_tot[j][wave] = __popcll(b0) + __popcll(b1);
      if (stage.bits) stage_ballots(stage, tile, j, wave, b0, b1);
    }
  }
  __syncthreads();
  uint32_t tile_total = 0;
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    uint32_t before = tile_total;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t v = s_wave_tot[j][w];
      before += (w < wave) ? v : 0;
      tile_total += v;
    }
    local[j] += before;
  }
  if (t == 0) {
    tile_count[tile] = tile_total;
    if (sums) {
      __int128 p = 0, b = 0;
      for (int w = 0; w < kWaves; ++w) {
        p += s_red[0][w];
        b += s_red[1][w];
      }
      tile_sums[2 * tile] = p;
      tile_sums[2 * tile + 1] = b;
    }
  }
  if (!stage.brow) return;
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    if (hit[j][0]) s_b[local[j]] = val[j][0];
    if (hit[j][1]) s_b[local[j] + (hit[j][0] ? 1 : 0)] = val[j][1];
  }
  __syncthreads();
  uint32_t* dst = stage.brow + base;
  // Plain stores: the move pass reads the staged rows back from the
  // Infinity Cache (non-temporal stores made it slower).
  for (uint32_t i = t; i < tile_total; i += kPB) dst[i] = s_b[i];
}

// Look-back of one wave over the status words of the tiles before `tile`
// (lane i reads tile - 1 - i, 64 predecessors per memory round trip).
// Returns the exclusive match prefix of `tile`.  A predecessor still
// unpublished after kStallPolls polls is counted by this wave and its count
// published, so the scan ends whatever order the workgroups run in.
template <int KIND>
__device__ uint64_t wave_lookback(uint64_t* status, uint32_t tile, const TableView& T,
                                  const int64_t* __restrict__ keys, uint64_t n, uint32_t stall_polls) {
  const uint32_t lane = lane_id();
  uint64_t excl = 0;
  int64_t base = static_cast<int64_t>(tile) - 1;
  uint32_t polls = 0;
  while (true) {
    const int64_t idx = base - static_cast<int64_t>(lane);
    const uint64_t v = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : kStPrefix;  // before tile 0: an inclusive prefix of 0
    const uint64_t st = v >> 62;
    const uint64_t pm = ballot64(st == 2), zm = ballot64(st == 0);
    const uint32_t fp = pm ? static_cast<uint32_t>(__builtin_ctzll(pm)) : 64u;  // nearest inclusive prefix
    const uint64_t upto = fp >= 63 ? ~0ULL : ((2ULL << fp) - 1);               // lanes 0..fp
    const uint64_t waiting = zm & upto;
    if (waiting) {
      if (++polls < stall_polls) {
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      const int64_t q = base - static_cast<int64_t>(__builtin_ctzll(waiting));
      const uint64_t c = tile_count<KIND, true>(T, keys, n, static_cast<uint64_t>(q));
      if (lane == 0) {
        unsigned long long expect = 0;
        __hip_atomic_compare_exchange_strong(reinterpret_cast<unsigned long long*>(&status[q]), &expect,
                                             static_cast<unsigned long long>(kStAgg | c), __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      polls = 0;
      continue;
    }
    uint64_t mine = (lane <= fp) ? (v & kStMask) : 0;
    for (int d = 32; d > 0; d >>= 1) mine += __shfl_xor(mine, d, 64);
    excl += mine;
    if (fp < 64) return excl;
    base -= 64;
  }
}

// Unique-key probe in one pass: as k_probe_unique up to the tile's ranked
// pairs in LDS, then the tile publishes its match count, one wave looks back
// over the predecessors' counts for the tile's output offset, and the
// workgroup writes its pairs there with coalesced stores.  The pairs are
// written once, with no staging copy and no move pass.
template <int KIND, bool VEC>
__global__ __launch_bounds__(kPB) void k_probe_unique_lb(TableView T, const int64_t* __restrict__ keys, uint64_t n,
                                                         uint32_t n_tiles, uint32_t probe_base,
                                                         const int64_t* __restrict__ ppay,
                                                         const int32_t* __restrict__ bpay, uint32_t* __restrict__ out_p,
                                                         uint32_t* __restrict__ out_b, uint64_t capacity,
                                                         uint64_t* status, __int128* __restrict__ tile_sums,
                                                         prism_probe_result* res, uint32_t stall_polls) {
  __shared__ uint32_t s_p[kTileRows];
  __shared__ uint32_t s_b[kTileRows];
  __shared__ uint32_t s_wave_tot[kPairs][kWaves];
  __shared__ __int128 s_red[2][kWaves];
  __shared__ uint64_t s_excl;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const uint32_t tile = blockIdx.x;
  const uint64_t base = static_cast<uint64_t>(tile) * kTileRows;
  const bool table_pay = T.has_payload && bpay == nullptr;
  const bool sums = ppay || bpay || table_pay;

  uint32_t val[kPairs][2];
  bool hit[kPairs][2];
  __int128 psum = 0, bsum = 0;
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const uint64_t r0 = base + static_cast<uint64_t>(j) * kChunk + 2 * t;
    int64_t k0, k1, p0 = 0, p1 = 0;
    bool ok0, ok1;
    if (VEC && r0 + 1 < n) {
      const v2i64 v = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(keys + r0));
      k0 = v.x;
      k1 = v.y;
      ok0 = ok1 = true;
      if (ppay) {
        const v2i64 p = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(ppay + r0));
        p0 = p.x;
        p1 = p.y;
      }
    } else {
      ok0 = r0 < n;
      ok1 = r0 + 1 < n;
      k0 = ok0 ? keys[r0] : 0;
      k1 = ok1 ? keys[r0 + 1] : 0;
      if (ppay) {
        p0 = ok0 ? ppay[r0] : 0;
        p1 = ok1 ? ppay[r0 + 1] : 0;
      }
    }
    uint32_t v0 = 0, v1 = 0, c;
    int32_t y0 = 0, y1 = 0;
    hit[j][0] = ok0 && lookup<KIND, true>(T, k0, v0, c, y0);
    hit[j][1] = ok1 && lookup<KIND, true>(T, k1, v1, c, y1);
    val[j][0] = v0;
    val[j][1] = v1;
    if (hit[j][0]) {
      psum += p0;
      bsum += bpay ? bpay[v0] : (table_pay ? y0 : 0);
    }
    if (hit[j][1]) {
      psum += p1;
      bsum += bpay ? bpay[v1] : (table_pay ? y1 : 0);
    }
  }
  // Rank matches inside the tile: row order is (chunk, thread, element).
  uint32_t local[kPairs];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const uint64_t b0 = ballot64(hit[j][0]), b1 = ballot64(hit[j][1]);
    local[j] = lanes_below(b0) + lanes_below(b1);
    if (lane == 0) s_wave_tot[j][wave] = __popcll(b0) + __popcll(b1);
  }
  if (sums) {
    wave_reduce_i128(psum);
    wave_reduce_i128(bsum);
    if (lane == 0) {
      s_red[0][wave] = psum;
      s_red[1][wave] = bsum;
    }
  }
  __syncthreads();
  uint32_t tile_total = 0;
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    uint32_t before = tile_total;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t v = s_wave_tot[j][w];
      before += (w < wave) ? v : 0;
      tile_total += v;
    }
    local[j] += before;
  }
  // Publish the count as early as possible: successors are looking back.
  if (t == 0)
    __hip_atomic_store(&status[tile], (tile == 0 ? kStPrefix : kStAgg) | tile_total, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const uint32_t row0 = static_cast<uint32_t>(base + static_cast<uint64_t>(j) * kChunk + 2 * t) + probe_base;
    if (hit[j][0]) {
      s_p[local[j]] = row0;
      s_b[local[j]] = val[j][0];
    }
    if (hit[j][1]) {
      const uint32_t q = local[j] + (hit[j][0] ? 1 : 0);
      s_p[q] = row0 + 1;
      s_b[q] = val[j][1];
    }
  }
  if (wave == 0) {
    const uint64_t excl = tile == 0 ? 0 : wave_lookback<KIND>(status, tile, T, keys, n, stall_polls);
    if (lane == 0) {
      if (tile != 0)
        __hip_atomic_store(&status[tile], kStPrefix | (excl + tile_total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      s_excl = excl;
      if (tile == n_tiles - 1) res->matches = excl + tile_total;
      if (sums) {
        __int128 p = 0, b = 0;
        for (int w = 0; w < kWaves; ++w) {
          p += s_red[0][w];
          b += s_red[1][w];
        }
        tile_sums[2 * tile] = p;
        tile_sums[2 * tile + 1] = b;
      }
    }
  }
  __syncthreads();
  if (!out_p) return;
  const uint64_t gbase = s_excl;
  for (uint32_t i = t; i < tile_total; i += kPB) {
    const uint64_t pos = gbase + i;
    if (pos < capacity) {
      __builtin_nontemporal_store(s_p[i], out_p + pos);
      __builtin_nontemporal_store(s_b[i], out_b + pos);
    }
  }
}

// Phase 3: move each tile's matches to their final offset.  The ballots give
// every match's probe row and rank in the tile (as the lookup pass ranked
// them); the probe rows are ranked into LDS, then both index arrays are
// written with coalesced stores, the build rows read in rank order.
__global__ __launch_bounds__(kPB) void k_probe_move(PairStage S, const uint32_t* __restrict__ tile_count,
                                                    const uint64_t* __restrict__ tile_off,
                                                    const uint64_t* __restrict__ slice_off, int slice_shift,
                                                    uint32_t probe_base, uint32_t* __restrict__ out_p,
                                                    uint32_t* __restrict__ out_b, uint64_t capacity) {
  __shared__ uint32_t s_p[kTileRows];
  __shared__ unsigned long long s_bits[kTileBallots];
  __shared__ uint32_t s_pre[kTileBallots / 2];  // matches before each (chunk, wave) group
  static_assert(kTileBallots == 64, "one wave scans the tile's groups");
  const uint32_t tile = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t cnt = tile_count[tile];
  const uint64_t off = slice_off[tile >> slice_shift] + tile_off[tile];
  const uint64_t base = static_cast<uint64_t>(tile) * kTileRows;
  if (wave == 0) {
    const unsigned long long w = S.bits[static_cast<uint64_t>(tile) * kTileBallots + lane];
    s_bits[lane] = w;
    // group g = lane / 2 holds words 2g, 2g + 1: the exclusive scan of the
    // words' match counts at word 2g is the matches before group g
    const uint32_t c = __popcll(w);
    uint32_t inc = c;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(inc, d, 64);
      if (lane >= d) inc += o;
    }
    if ((lane & 1) == 0) s_pre[lane >> 1] = inc - c;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const int g = j * kWaves + wave;
    const uint64_t b0 = s_bits[2 * g], b1 = s_bits[2 * g + 1];
    const uint32_t r0 = s_pre[g] + lanes_below(b0) + lanes_below(b1);
    const bool h0 = (b0 >> lane) & 1, h1 = (b1 >> lane) & 1;
    const uint32_t row0 = static_cast<uint32_t>(base + static_cast<uint64_t>(j) * kChunk + 2 * t) + probe_base;
    if (h0) s_p[r0] = row0;
    if (h1) s_p[r0 + (h0 ? 1 : 0)] = row0 + 1;
  }
  __syncthreads();
  const uint32_t* src = S.brow + base;
  for (uint32_t i = t; i < cnt; i += kPB) {
    const uint64_t pos = off + i;
    if (pos < capacity) {
      __builtin_nontemporal_store(s_p[i], out_p + pos);
      __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), out_b + pos);
    }
  }
}


// ---------------------------------------------------------------------------
// Partitioned probe of a unique HASHED table cut into LDS-sized partitions.
//   A  k_pt_hist     rows of each 4096-row tile per partition (partition-major)
//   -  scan          record offset of each (partition, tile) segment
//   C  k_pt_scatter  (key, row in tile) records into their segments
//   D  k_pt_join     per partition: its slots into LDS, its records looked up
//                    there, (build row, payload) or none per record
//   E  k_pt_restore  per tile: the results of its segments back in row order,
//                    then ranked and staged as the direct probe stages them
// followed by the direct probe's tile scan and move.  One partition (the
// whole table in LDS) skips A-D: k_pt_restore<true> looks rows up itself.
// ---------------------------------------------------------------------------
constexpr uint32_t kPtChunk = 16384;    // records a join workgroup takes at a time
constexpr uint32_t kPtMaxParts = 2048;  // partitions (segment table = partitions x tiles)

__global__ __launch_bounds__(kPB) void k_pt_hist(const int64_t* __restrict__ keys, uint64_t n, int pshift, uint32_t P,
                                                 uint32_t n_tiles, uint32_t* __restrict__ hist) {
  extern __shared__ uint32_t s_h[];
  for (uint32_t i = threadIdx.x; i < P; i += kPB) s_h[i] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTileRows;
  for (uint32_t i = threadIdx.x; i < static_cast<uint32_t>(kTileRows); i += kPB) {
    const uint64_t r = base + i;
    if (r < n) atomicAdd(&s_h[hash_i64_fast(__builtin_nontemporal_load(keys + r)) >> pshift], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += kPB) hist[static_cast<uint64_t>(p) * n_tiles + blockIdx.x] = s_h[p];
}

__global__ __launch_bounds__(kPB) void k_pt_scatter(const int64_t* __restrict__ keys, uint64_t n, int pshift,
                                                    uint32_t P, uint32_t n_tiles, const uint32_t* __restrict__ offs,
                                                    int64_t* __restrict__ rec_key, uint16_t* __restrict__ rec_row) {
  extern __shared__ uint32_t s_c[];
  for (uint32_t p = threadIdx.x; p < P; p += kPB) s_c[p] = offs[static_cast<uint64_t>(p) * n_tiles + blockIdx.x];
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTileRows;
  for (uint32_t i = threadIdx.x; i < static_cast<uint32_t>(kTileRows); i += kPB) {
    const uint64_t r = base + i;
    if (r >= n) continue;
    const int64_t k = __builtin_nontemporal_load(keys + r);
    const uint32_t at = atomicAdd(&s_c[hash_i64_fast(k) >> pshift], 1u);
    rec_key[at] = k;
    rec_row[at] = static_cast<uint16_t>(i);
  }
}

// (build row, slot payload word) of key k in a partition's slots held in LDS;
// false when absent.
__device__ __forceinline__ bool pt_lookup(const TableView& T, const Slot* __restrict__ s_t, int64_t k, uint2& out) {
  if (k == kEmpty) {
    if (!T.sentinel_present) return false;
    const Slot e = T.slots[T.mask + 1];
    out = make_uint2(e.val, e.cnt);
    return true;
  }
  uint64_t s = (hash_i64_fast(k) >> T.shift) & T.pmask;
  while (true) {  // the build left an empty slot in every partition
    const Slot e = s_t[s];
    if (e.key == k) {
      out = make_uint2(e.val, e.cnt);
      return true;
    }
    if (e.key == kEmpty) return false;
    s = (s + 1) & T.pmask;
  }
}

// D: workgroup (p, m) of M per partition loads partition p's slots and looks
// up records m, m + M, ... chunks of that partition's segment range.
__global__ __launch_bounds__(kPB) void k_pt_join(TableView T, const int64_t* __restrict__ rec_key,
                                                 const uint32_t* __restrict__ offs, uint32_t n_tiles, uint32_t M,
                                                 uint2* __restrict__ result) {
  extern __shared__ Slot s_t[];
  const uint64_t PS = T.pmask + 1;
  const uint32_t p = blockIdx.x / M, m = blockIdx.x % M;
  const uint32_t start = offs[static_cast<uint64_t>(p) * n_tiles];
  const uint32_t end = offs[static_cast<uint64_t>(p + 1) * n_tiles];
  if (static_cast<uint64_t>(start) + static_cast<uint64_t>(m) * kPtChunk >= end) return;
  const Slot* src = T.slots + static_cast<uint64_t>(p) * PS;
  for (uint64_t i = threadIdx.x; i < PS; i += kPB) s_t[i] = src[i];
  __syncthreads();
  for (uint64_t c0 = start + static_cast<uint64_t>(m) * kPtChunk; c0 < end; c0 += static_cast<uint64_t>(M) * kPtChunk) {
    const uint64_t c1 = c0 + kPtChunk < end ? c0 + kPtChunk : end;
    for (uint64_t i = c0 + threadIdx.x; i < c1; i += kPB) {
      uint2 out = make_uint2(kNone, 0);
      pt_lookup(T, s_t, __builtin_nontemporal_load(rec_key + i), out);
      result[i] = out;
    }
  }
}

// E: one tile's matches in row order, ranked and staged exactly as
// k_probe_unique stages them.  DIRECT: the table is one partition held in LDS
// and the workgroup looks its tiles' keys up itself (tiles blockIdx.x,
// + gridDim.x, ...); otherwise one tile per workgroup, gathered from the
// tile's segment of every partition.
template <bool DIRECT>
__global__ __launch_bounds__(kPB) void k_pt_restore(TableView T, const int64_t* __restrict__ keys, uint64_t n,
                                                    uint32_t n_tiles, const uint32_t* __restrict__ offs, uint32_t P,
                                                    const uint16_t* __restrict__ rec_row,
                                                    const uint2* __restrict__ result,
                                                    const int64_t* __restrict__ ppay, const int32_t* __restrict__ bpay,
                                                    PairStage stage, uint32_t* __restrict__ tile_count,
                                                    __int128* __restrict__ tile_sums) {
  extern __shared__ Slot s_t[];  // DIRECT: the table's slots
  __shared__ uint2 s_m[kTileRows];  // (build row, payload) per row of the tile, then its staged build rows
  __shared__ uint32_t s_wave_tot[kPairs][kWaves];
  __shared__ __int128 s_red[2][kWaves];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const bool table_pay = T.has_payload && bpay == nullptr;
  const bool sums = ppay || bpay || table_pay;
  if (DIRECT) {
    for (uint64_t i = t; i <= T.pmask; i += kPB) s_t[i] = T.slots[i];
    __syncthreads();
  }
  for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += DIRECT ? gridDim.x : n_tiles) {
    const uint64_t base = static_cast<uint64_t>(tile) * kTileRows;
    if (!DIRECT) {
      for (uint32_t i = t; i < static_cast<uint32_t>(kTileRows); i += kPB) s_m[i] = make_uint2(kNone, 0);
      __syncthreads();
      for (uint32_t p = t; p < P; p += kPB) {
        const uint64_t g = static_cast<uint64_t>(p) * n_tiles + tile;
        const uint32_t a = offs[g], b = offs[g + 1];
        for (uint32_t i = a; i < b; ++i) s_m[rec_row[i]] = result[i];
      }
      __syncthreads();
    }
    uint2 val[kPairs][2];
    bool hit[kPairs][2];
    __int128 psum = 0, bsum = 0;
#pragma unroll
    for (int j = 0; j < kPairs; ++j) {
      const uint32_t l0 = static_cast<uint32_t>(j) * kChunk + 2 * t;
      const uint64_t r0 = base + l0;
      const bool ok0 = r0 < n, ok1 = r0 + 1 < n;
      int64_t p0 = 0, p1 = 0;
      if (ppay) {
        p0 = ok0 ? ppay[r0] : 0;
        p1 = ok1 ? ppay[r0 + 1] : 0;
      }
      uint2 v0 = make_uint2(kNone, 0), v1 = make_uint2(kNone, 0);
      if (DIRECT) {
        if (ok0) pt_lookup(T, s_t, keys[r0], v0);
        if (ok1) pt_lookup(T, s_t, keys[r0 + 1], v1);
      } else {
        if (ok0) v0 = s_m[l0];
        if (ok1) v1 = s_m[l0 + 1];
      }
      hit[j][0] = v0.x != kNone;
      hit[j][1] = v1.x != kNone;
      val[j][0] = v0;
      val[j][1] = v1;
      if (hit[j][0]) {
        psum += p0;
        bsum += bpay ? bpay[v0.x] : (table_pay ? static_cast<int32_t>(v0.y) : 0);
      }
      if (hit[j][1]) {
        psum += p1;
        bsum += bpay ? bpay[v1.x] : (table_pay ? static_cast<int32_t>(v1.y) : 0);
      }
    }
    if (sums) {
      wave_reduce_i128(psum);
      wave_reduce_i128(bsum);
      if (lane == 0) {
        s_red[0][wave] = psum;
        s_red[1][wave] = bsum;
      }
    }
    uint32_t local[kPairs];
#pragma unroll
    for (int j = 0; j < kPairs; ++j) {
      const uint64_t b0 = ballot64(hit[j][0]), b1 = ballot64(hit[j][1]);
      local[j] = lanes_below(b0) + lanes_below(b1);
      if (lane == 0) {
        s_wave_tot[j][wave] = __popcll(b0) + __popcll(b1);
        if (stage.bits) stage_ballots(stage, tile, j, wave, b0, b1);
      }
    }
    __syncthreads();  // also: every s_m read above is done before the build rows overwrite it
    uint32_t tile_total = 0;
#pragma unroll
    for (int j = 0; j < kPairs; ++j) {
      uint32_t before = tile_total;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t v = s_wave_tot[j][w];
        before += (w < wave) ? v : 0;
        tile_total += v;
      }
      local[j] += before;
    }
    if (t == 0) {
      tile_count[tile] = tile_total;
      if (sums) {
        __int128 pp = 0, bb = 0;
        for (int w = 0; w < kWaves; ++w) {
          pp += s_red[0][w];
          bb += s_red[1][w];
        }
        tile_sums[2 * tile] = pp;
        tile_sums[2 * tile + 1] = bb;
      }
    }
    if (stage.brow) {
      uint32_t* s_mb = reinterpret_cast<uint32_t*>(s_m);
#pragma unroll
      for (int j = 0; j < kPairs; ++j) {
        if (hit[j][0]) s_mb[local[j]] = val[j][0].x;
        if (hit[j][1]) s_mb[local[j] + (hit[j][0] ? 1 : 0)] = val[j][1].x;
      }
      __syncthreads();
      uint32_t* dst = stage.brow + base;
      for (uint32_t i = t; i < tile_total; i += kPB) dst[i] = s_mb[i];
    }
    __syncthreads();  // s_m, s_wave_tot and s_red are reused by the next tile
  }
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

template <int KIND, bool UNIQUE, bool VEC>
__global__ __launch_bounds__(kPB, 2) void k_probe(TableView T, const int64_t* __restrict__ keys, uint64_t n,
                                               uint32_t n_tiles, uint32_t probe_base,
                                               const int64_t* __restrict__ ppay, const int32_t* __restrict__ bpay,
                                               uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                               uint64_t capacity, uint64_t* status, __int128* tile_sums,
                                               prism_probe_result* res) {
  __shared__ uint64_t s_wave_tot[kPairs][kWaves];
  __shared__ LookbackLds s_lb;
  __shared__ __int128 s_red[2][kWaves];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const uint32_t tile = blockIdx.x;
  const uint64_t base = static_cast<uint64_t>(tile) * kTileRows;

  // Keys (and probe payloads) for 8 rows: pairs of consecutive rows, one
  // 128-bit load per pair, consecutive lanes on consecutive pairs.
  int64_t k[kPairs][2], pv[kPairs][2];
  bool ok[kPairs][2];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    uint64_t r0 = base + static_cast<uint64_t>(j) * kChunk + 2 * t;
    if (VEC && r0 + 1 < n) {
      const v2i64 v = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(keys + r0));
      k[j][0] = v.x;
      k[j][1] = v.y;
      ok[j][0] = ok[j][1] = true;
      if (ppay) {
        const v2i64 p = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(ppay + r0));
        pv[j][0] = p.x;
        pv[j][1] = p.y;
      } else {
        pv[j][0] = pv[j][1] = 0;
      }
    } else {
      ok[j][0] = r0 < n;
      ok[j][1] = r0 + 1 < n;
      k[j][0] = ok[j][0] ? keys[r0] : 0;
      k[j][1] = ok[j][1] ? keys[r0 + 1] : 0;
      pv[j][0] = (ppay && ok[j][0]) ? ppay[r0] : 0;
      pv[j][1] = (ppay && ok[j][1]) ? ppay[r0 + 1] : 0;
    }
  }

  uint32_t val[kPairs][2], cnt[kPairs][2];
  int32_t tpay[kPairs][2];
#pragma unroll
  for (int j = 0; j < kPairs; ++j)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      uint32_t v = 0, c = 0;
      int32_t py = 0;
      bool m = ok[j][e] && lookup<KIND, UNIQUE>(T, k[j][e], v, c, py);
      val[j][e] = v;
      cnt[j][e] = m ? c : 0;
      tpay[j][e] = py;
    }
  const bool table_pay = T.has_payload && bpay == nullptr;

  // Payload sums over the matches (gathers issued before the scan so their
  // latency overlaps it).
  __int128 psum = 0, bsum = 0;
  if constexpr (UNIQUE) {
#pragma unroll
    for (int j = 0; j < kPairs; ++j)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        if (cnt[j][e]) {
          psum += pv[j][e];
          if (bpay) bsum += bpay[val[j][e]];
          else if (table_pay) bsum += tpay[j][e];
        }
      }
  }

  // Rank matches inside the tile: row order is (chunk, thread, element).
  // UNIQUE tiles hold at most kTileRows matches, so positions fit in 32 bits.
  using Pos = typename std::conditional<UNIQUE, uint32_t, uint64_t>::type;
  Pos local[kPairs];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    if constexpr (UNIQUE) {
      uint64_t b0 = ballot64(cnt[j][0] != 0), b1 = ballot64(cnt[j][1] != 0);
      local[j] = lanes_below(b0) + lanes_below(b1);
      if (lane == 0) s_wave_tot[j][wave] = __popcll(b0) + __popcll(b1);
    } else {
      uint64_t c = static_cast<uint64_t>(cnt[j][0]) + cnt[j][1];
      uint64_t x = c;
      for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
      }
      local[j] = x - c;
      if (lane == 63) s_wave_tot[j][wave] = x;
    }
  }
  __syncthreads();
  uint64_t tile_total = 0;
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    Pos before = static_cast<Pos>(tile_total);
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      uint64_t v = s_wave_tot[j][w];
      before += (w < wave) ? static_cast<Pos>(v) : 0;
      tile_total += v;
    }
    local[j] += before;
  }
  const uint64_t gbase = lookback<KIND, UNIQUE>(status, tile, tile_total, T, keys, n, s_lb);

  const bool want_pairs = out_p != nullptr;
#pragma unroll
  for (int j = 0; j < kPairs; ++j)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      if (cnt[j][e] == 0) continue;
      const uint64_t row = base + static_cast<uint64_t>(j) * kChunk + 2 * t + e;
      uint64_t pos = gbase + local[j] + (e ? cnt[j][0] : 0);
      if constexpr (UNIQUE) {
        if (want_pairs && pos < capacity) {
          __builtin_nontemporal_store(static_cast<uint32_t>(row) + probe_base, out_p + pos);
          __builtin_nontemporal_store(val[j][e], out_b + pos);
        }
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
          else if (table_pay) bsum += T.row_payload[start + r];
        }
        psum += static_cast<__int128>(pv[j][e]) * static_cast<__int128>(c);
      }
    }

  if (ppay || bpay || table_pay) {  // per-tile partials, summed by k_probe_sums (no contended atomics)
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
      tile_sums[2 * tile] = p;
      tile_sums[2 * tile + 1] = b;
    }
  }
  if (tile == n_tiles - 1 && t == 0) res->matches = gbase + tile_total;
}

// Sum of the per-tile payload partials (one workgroup).
__global__ void k_probe_sums(const __int128* __restrict__ tile_sums, uint64_t n_tiles, prism_probe_result* res) {
  __shared__ __int128 s_red[2][16];
  __int128 p = 0, b = 0;
  for (uint64_t i = threadIdx.x; i < n_tiles; i += blockDim.x) {
    p += tile_sums[2 * i];
    b += tile_sums[2 * i + 1];
  }
  wave_reduce_i128(p);
  wave_reduce_i128(b);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[0][w] = p;
    s_red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    p = 0;
    b = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) {
      p += s_red[0][i];
      b += s_red[1][i];
    }
    res->probe_sum_lo = static_cast<uint64_t>(p);
    res->probe_sum_hi = static_cast<int64_t>(p >> 64);
    res->build_sum_lo = static_cast<uint64_t>(b);
    res->build_sum_hi = static_cast<int64_t>(b >> 64);
  }
}

// Phase 2a: each workgroup takes a slice of kSliceTiles consecutive tiles,
// writes the exclusive scan of their match counts (offsets local to the
// slice) and the slice's match total and payload sums.
constexpr int kSliceThreads = 256, kSlicePer = 4, kSliceTiles = kSliceThreads * kSlicePer;

__global__ __launch_bounds__(kSliceThreads) void k_probe_slices(const uint32_t* __restrict__ tile_count,
                                                                uint64_t* __restrict__ tile_off,
                                                                const __int128* __restrict__ tile_sums,
                                                                uint64_t n_tiles, uint64_t* __restrict__ slice_total,
                                                                __int128* __restrict__ slice_sums) {
  __shared__ uint64_t s_wave[kSliceThreads / 64];
  __shared__ __int128 s_red[2][kSliceThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSliceTiles + static_cast<uint64_t>(t) * kSlicePer;
  uint32_t c[kSlicePer];
  uint64_t run = 0;
#pragma unroll
  for (int j = 0; j < kSlicePer; ++j) {
    c[j] = base + j < n_tiles ? tile_count[base + j] : 0;
    run += c[j];
  }
  uint64_t inc = run;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint64_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kSliceThreads / 64; ++w) {
    before += (w < wave) ? s_wave[w] : 0;
    total += s_wave[w];
  }
  if (tile_off) {
    uint64_t x = before + inc - run;
#pragma unroll
    for (int j = 0; j < kSlicePer; ++j) {
      if (base + j < n_tiles) tile_off[base + j] = x;
      x += c[j];
    }
  }
  if (t == 0) slice_total[blockIdx.x] = total;
  if (tile_sums) {
    __int128 p = 0, b = 0;
#pragma unroll
    for (int j = 0; j < kSlicePer; ++j)
      if (base + j < n_tiles) {
        p += tile_sums[2 * (base + j)];
        b += tile_sums[2 * (base + j) + 1];
      }
    wave_reduce_i128(p);
    wave_reduce_i128(b);
    if (lane == 0) {
      s_red[0][wave] = p;
      s_red[1][wave] = b;
    }
    __syncthreads();
    if (t == 0) {
      p = 0;
      b = 0;
      for (int w = 0; w < kSliceThreads / 64; ++w) {
        p += s_red[0][w];
        b += s_red[1][w];
      }
      slice_sums[2 * blockIdx.x] = p;
      slice_sums[2 * blockIdx.x + 1] = b;
    }
  }
}

// Phase 2b (one workgroup): exclusive scan of the slice totals in place,
// total matches and payload sums into the result.
constexpr int kFinishThreads = 1024;

__global__ __launch_bounds__(kFinishThreads) void k_probe_finish(uint64_t* __restrict__ slice_total,
                                                                 const __int128* __restrict__ slice_sums,
                                                                 uint64_t n_slices, prism_probe_result* res) {
  __shared__ uint64_t s_wave[kFinishThreads / 64];
  __shared__ __int128 s_red[2][kFinishThreads / 64];
  __shared__ uint64_t s_carry;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (uint64_t base = 0; base < n_slices; base += kFinishThreads) {
    const uint64_t i = base + t;
    const uint64_t v = i < n_slices ? slice_total[i] : 0;
    uint64_t inc = v;
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t o = __shfl_up(inc, d, 64);
      if (lane >= d) inc += o;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint64_t before = s_carry, total = 0;
    for (int w = 0; w < kFinishThreads / 64; ++w) {
      before += (w < wave) ? s_wave[w] : 0;
      total += s_wave[w];
    }
    if (i < n_slices) slice_total[i] = before + inc - v;
    __syncthreads();
    if (t == 0) s_carry += total;
    __syncthreads();
  }
  if (slice_sums) {
    __int128 p = 0, b = 0;
    for (uint64_t i = t; i < n_slices; i += kFinishThreads) {
      p += slice_sums[2 * i];
      b += slice_sums[2 * i + 1];
    }
    wave_reduce_i128(p);
    wave_reduce_i128(b);
    if (lane == 0) {
      s_red[0][wave] = p;
      s_red[1][wave] = b;
    }
    __syncthreads();
    if (t == 0) {
      p = 0;
      b = 0;
      for (int w = 0; w < kFinishThreads / 64; ++w) {
        p += s_red[0][w];
        b += s_red[1][w];
      }
      res->probe_sum_lo = static_cast<uint64_t>(p);
      res->probe_sum_hi = static_cast<int64_t>(p >> 64);
      res->build_sum_lo = static_cast<uint64_t>(b);
      res->build_sum_hi = static_cast<int64_t>(b >> 64);
    }
  }
  if (t == 0) res->matches = s_carry;
}

template <int KIND, bool UNIQUE>
static void launch_probe_vec(prism_ctx* ctx, bool vec, dim3 g, const TableView& T, const int64_t* keys,
                             uint64_t n, uint32_t n_tiles, uint32_t probe_base, const int64_t* ppay,
                             const int32_t* bpay, uint32_t* op, uint32_t* ob, uint64_t cap, uint64_t* status,
                             __int128* tile_sums, prism_probe_result* res) {
  if (vec)
    hipLaunchKernelGGL((k_probe<KIND, UNIQUE, true>), g, dim3(kPB), 0, ctx->stream, T, keys, n, n_tiles,
                       probe_base, ppay, bpay, op, ob, cap, status, tile_sums, res);
  else
    hipLaunchKernelGGL((k_probe<KIND, UNIQUE, false>), g, dim3(kPB), 0, ctx->stream, T, keys, n, n_tiles,
                       probe_base, ppay, bpay, op, ob, cap, status, tile_sums, res);
}


// ---------------------------------------------------------------------------
// Join -> GROUP BY build attribute, COUNT(*), SUM(probe payload), without
// materialising pairs (the fused join-aggregate of the prism-substrait
// executor, executor.rs:333-458).  Eager aggregation:
//   1. every probe row adds (1, payload, row) to its join key's slot
//      (count, exact int128 sum, first probe row); hot keys are merged in an
//      LDS cache first, so a Zipf head does not serialise on global atomics;
//   2. every build row adds its key's (count, sum) to its attribute's group,
//      and offers (first probe row, build row) as the group's first pair.
// Groups come out in first appearance over the join output (probe row
// ascending, build row ascending within a probe row), the order in which
// hash_aggregate meets them in the reference.
// ---------------------------------------------------------------------------
constexpr int kJaBlock = 256;
constexpr int kJaCache = 1024;   // LDS cache entries per workgroup
constexpr uint32_t kJaEmpty = 0xFFFFFFFFu;

// Slot of a key in the table (DENSE: key offset; HASHED: slot index, the
// INT64_MIN key at mask + 1), or -1 when no build row has the key.
template <int KIND>
__device__ __forceinline__ int64_t slot_of(const TableView& T, int64_t k, bool unique) {
  if constexpr (KIND == PRISM_TABLE_DENSE) {
    const uint64_t d = static_cast<uint64_t>(k) - static_cast<uint64_t>(T.min_key);
    if (d >= T.range) return -1;
    if (T.presence_only) return ((T.present_bits[d >> 5] >> (d & 31)) & 1u) ? static_cast<int64_t>(d) : -1;
    if (unique) return T.dense_u[d] != kNone ? static_cast<int64_t>(d) : -1;
    return T.dense_nu[d].y != 0 ? static_cast<int64_t>(d) : -1;
  } else {
    if (k == kEmpty) return T.sentinel_present ? static_cast<int64_t>(T.mask + 1) : -1;
    uint64_t s = hash_i64_fast(k) >> T.shift;
    while (true) {
      const int64_t key = T.slots[s].key;
      if (key == k) return static_cast<int64_t>(s);
      if (key == kEmpty) return -1;
      s = (s & ~T.pmask) | ((s + 1) & T.pmask);
    }
  }
}

// Probe-side aggregates per key slot in one 32-byte record (a slot's reads
// and writes touch one cache line): row count, int128 payload sum, first row.
struct JaAcc {
  unsigned long long* w;  // per key slot, 32 B: count, sum lo, sum hi, first probe row
};

// Probe key -> key slot.  A DENSE table tests presence in a bitmap of the
// key range (1 bit per key: small enough to stay in L2) when one is given.
template <int KIND>
__device__ __forceinline__ int64_t probe_slot(const TableView& T, int64_t k, bool unique, const uint32_t* present) {
  if constexpr (KIND == PRISM_TABLE_DENSE) {
    if (!present) return slot_of<KIND>(T, k, unique);
    const uint64_t d = static_cast<uint64_t>(k) - static_cast<uint64_t>(T.min_key);
    if (d >= T.range) return -1;
    return ((present[d >> 5] >> (d & 31)) & 1u) ? static_cast<int64_t>(d) : -1;
  } else {
    return slot_of<KIND>(T, k, unique);
  }
}

// LDS cache of hot key slots: per entry the slot, row count, split int128
// payload sum (low 32-bit halves unsigned, high halves signed) and first row.
// The cache starts with the slots placed at create time (the slots with the
// most build rows, the probe's likely Zipf head) and fills its free ways
// with the first other slots the workgroup meets.  Entries sit in buckets of
// four read with one 16-byte LDS load; a slot may sit in either of two
// buckets, so a lookup costs at most two loads.
constexpr int kJaWays = 4;
constexpr int kJaBuckets = kJaCache / kJaWays;
static_assert(kJaBuckets == 256, "two 8-bit bucket indices");

__host__ __device__ __forceinline__ uint32_t ja_bucket(uint32_t s, int t) {
  const uint32_t h = s * 0x9E3779B1u;  // one 32-bit multiply: the top two bytes pick the buckets
  return (h >> (t ? 16 : 24)) & (kJaBuckets - 1);
}


struct JaCache {
  uint32_t key[kJaCache];
  uint32_t cnt[kJaCache];  // rows of one workgroup in one pass (a pass has fewer than 2^32 rows)
  unsigned long long lo[kJaCache], first[kJaCache];
  long long hi[kJaCache];
};

__device__ __forceinline__ void ja_cache_init(JaCache& C, const uint32_t* __restrict__ keys) {
  for (int i = threadIdx.x; i < kJaCache; i += blockDim.x) {
    if (i < kJaCache) C.key[i] = keys[i];
    C.cnt[i] = 0;
    C.lo[i] = 0;
    C.hi[i] = 0;
    C.first[i] = ~0ULL;
  }
}

// One bucket's four ways in one 16-byte LDS load.  The pointer is cast to the
// LDS address space explicitly: a volatile access through a generic pointer
// would be a flat load, which counts as a global load too, so waiting for it
// would wait for every stream load in flight.  Volatile: ways claimed by other
// threads since the last look must be seen.
typedef __attribute__((address_space(3))) v4u32 lds_v4u32;
__device__ __forceinline__ v4u32 lds_load_ways(const uint32_t* p) {
  return *(const volatile lds_v4u32*)(p);
}

// Entry of slot s: a slot placed at create time, or a free way of either
// bucket claimed now (the first slots a workgroup meets, which covers a probe
// side skewed on keys the build side does not rank); -1 when both buckets
// are full.  A slot held by two entries would only split its totals.
__device__ __forceinline__ int ja_cache_entry(JaCache& C, uint32_t s) {
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const uint32_t b = ja_bucket(s, t);
    for (int attempt = 0; attempt < 2; ++attempt) {
      const v4u32 e = lds_load_ways(&C.key[kJaWays * b]);
      if (e.x == s) return kJaWays * b;
      if (e.y == s) return kJaWays * b + 1;
      if (e.z == s) return kJaWays * b + 2;
      if (e.w == s) return kJaWays * b + 3;
      const int f = e.x == kJaEmpty ? 0 : e.y == kJaEmpty ? 1 : e.z == kJaEmpty ? 2 : e.w == kJaEmpty ? 3 : -1;
      if (f < 0) break;
      const uint32_t old = atomicCAS(&C.key[kJaWays * b + f], kJaEmpty, s);
      if (old == kJaEmpty || old == s) return kJaWays * b + f;
    }
  }
  return -1;
}

// Adds one row to slot s's cache entry; false when the slot is not cached.
__device__ __forceinline__ bool ja_cache_add(JaCache& C, uint32_t s, int64_t x, unsigned long long row) {
  const int c = ja_cache_entry(C, s);
  if (c < 0) return false;
  atomicAdd(&C.cnt[c], 1u);
  const uint32_t lo = static_cast<uint32_t>(static_cast<uint64_t>(x));
  const long long hi = x >> 32;
  if (lo) atomicAdd(&C.lo[c], static_cast<unsigned long long>(lo));
  if (hi) atomicAdd(reinterpret_cast<unsigned long long*>(&C.hi[c]), static_cast<unsigned long long>(hi));
  if (row < __hip_atomic_load(&C.first[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMin(&C.first[c], row);
  return true;
}

// One row added to its slot with device atomics.
__device__ __forceinline__ void ja_add_direct(const JaAcc& A, uint64_t s, int64_t x, unsigned long long row) {
  atomicAdd(&A.w[4 * s], 1ULL);
  if (x) atomic_add_i128(&A.w[4 * s + 1], &A.w[4 * s + 2], static_cast<uint64_t>(x), x >> 63);
  if (row < __hip_atomic_load(&A.w[4 * s + 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&A.w[4 * s + 3], row);
}

// Cache entries merged into the slot arrays (after a barrier).
__device__ __forceinline__ void ja_cache_flush(JaCache& C, const JaAcc& A) {
  for (int c = threadIdx.x; c < kJaCache; c += blockDim.x) {
    if (C.key[c] == kJaEmpty || C.cnt[c] == 0) continue;
    const uint64_t s = C.key[c];
    atomicAdd(&A.w[4 * s], static_cast<unsigned long long>(C.cnt[c]));
    Split128 sp;
    sp.lo32 = C.lo[c];
    sp.hi32 = C.hi[c];
    uint64_t lo;
    int64_t hi;
    sp.to_i128(lo, hi);
    if (lo | static_cast<uint64_t>(hi)) atomic_add_i128(&A.w[4 * s + 1], &A.w[4 * s + 2], lo, hi);
    atomicMin(&A.w[4 * s + 3], C.first[c]);
  }
}

// Probe pass with every miss of the LDS cache added by device atomics (used
// when the slot space is too large for the partitioned tail below).
template <int KIND>
__global__ __launch_bounds__(kJaBlock) void k_ja_probe(TableView T, bool unique, const int64_t* __restrict__ keys,
                                                       const int64_t* __restrict__ pay, uint64_t n, uint64_t row_base,
                                                       JaAcc A, const uint32_t* __restrict__ present,
                                                       const uint32_t* __restrict__ cache_keys) {
  __shared__ JaCache C;
  ja_cache_init(C, cache_keys);
  __syncthreads();
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kJaBlock;
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kJaBlock) + threadIdx.x; r < n; r += stride) {
    const int64_t s = probe_slot<KIND>(T, keys[r], unique, present);
    if (s < 0) continue;
    const int64_t x = pay ? pay[r] : 0;
    const unsigned long long row = row_base + r;
    if (!ja_cache_add(C, static_cast<uint32_t>(s), x, row)) ja_add_direct(A, static_cast<uint64_t>(s), x, row);
  }
  __syncthreads();
  ja_cache_flush(C, A);
}

// ---------------------------------------------------------------------------
// Partitioned tail.  Device atomics on random slots run at about 24 G/s on
// MI355X whatever their scope or footprint (measured), so rows whose slot
// misses the LDS cache are not added to their slot directly.  They are
// radix-partitioned by slot in two passes - level 1 by the slot's high bits
// into at most 64 partitions, then into ranges of kJaSub slots - and each
// range is summed by one workgroup with LDS atomics and written back with
// plain read-modify-writes (no other workgroup touches that range).  Level-1
// regions have a fixed capacity; rows beyond it (a partition far above its
// share) fall back to device atomics, so any input stays exact.
// ---------------------------------------------------------------------------
constexpr int kJaSubShift = 12;
constexpr int kJaSub = 1 << kJaSubShift;  // slots per aggregation range (LDS: 24 B per slot)
constexpr int kJaP1Max = 64;              // level-1 partitions
constexpr int kJaFanMax = 64;             // aggregation ranges per level-1 partition
constexpr int kJaStage = 32;              // K1 LDS staging records per level-1 partition
constexpr int kJaFlushAt = 16;            // K1 flushes a staging buffer at this fill
constexpr int kJaBlock1 = 512;            // K1 threads (two workgroups per CU)
constexpr int kJaRows1 = 4;               // rows per thread per K1 round
constexpr int kJaWaves1 = kJaBlock1 / 64;
constexpr uint32_t kJaBlk = 256;          // level-1 records a K1 workgroup reserves at a time
constexpr uint32_t kJaHole = 0xFFFFFFFFu; // slot of an unused reserved record
constexpr int kJaPartBlock = 1024;
constexpr int kJaWavesPB = kJaPartBlock / 64;
constexpr uint32_t kJaPartRecs = 4096;    // level-1 records per block of the repartition
constexpr int kJaPartPerThread = kJaPartRecs / kJaPartBlock;

// Level 2 without a counting pass: a repartition block (kJaPartRecs level-1
// records of one partition) orders its records by range in LDS, takes room
// for all of them in its partition's l2 region with one device atomic, writes
// them there contiguously, and appends one run {offset, records} per range it
// holds to that range's run list.  A range's runs come from distinct blocks
// of its partition, so a list holds at most nb runs.
struct JaPart {
  unsigned long long* l1_sr;  // level-1 regions, partition p at p * cap1: slot | row << 32 (holes: slot kJaHole)
  long long* l1_pay;          //   and the row's payload
  unsigned long long* fill1;  // per level-1 partition: records reserved (may pass cap1)
  uint64_t cap1;
  int shift1;                 // slot >> shift1 = level-1 partition
  int fan;                    // aggregation ranges per level-1 partition (power of two)
  int n_p1;                   // level-1 partitions
  uint32_t nb;                // repartition blocks per level-1 partition, at most
  unsigned long long* off2;   // per partition: its l2 region (records K1 reserved there, scanned)
  unsigned long long* fill2;  // per partition: l2 records taken
  uint32_t* blk0;             // per partition: its first repartition block; [n_p1] = all blocks
  uint32_t* nruns;            // per range: runs
  unsigned long long* run_off;  // per (range, run): first record in l2
  uint32_t* run_cnt;            //   and its records
  uint4* l2;                  // {slot & (kJaSub - 1), row, payload lo, payload hi}, in runs of one range
};

// Barrier for LDS traffic only: global stores and atomics in flight are not
// waited for (the compiler still waits for loaded values before their use).
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

struct JaStage1 {
  unsigned long long sr[kJaP1Max][kJaStage];  // slot | row << 32
  long long pay[kJaP1Max][kJaStage];
  uint32_t n[kJaP1Max];
  unsigned long long cur[kJaP1Max];  // next free record of the reserved block
  unsigned long long end[kJaP1Max];  // end of the reserved block
  unsigned long long nxt[kJaP1Max];  // block reserved ahead (~0 = none)
};

// Staging buffers holding at least `at_least` records go to their level-1
// region, half a wave per buffer.  Space comes in blocks of kJaBlk records,
// reserved a round ahead (the reservation's latency hides behind that
// round's loads); records past the region's capacity are added to their
// slots with device atomics.
__device__ __forceinline__ void ja_stage1_flush(JaStage1& S, const JaPart& P, const JaAcc& A, uint32_t at_least) {
  static_assert(kJaStage == 32, "half a wave per staging buffer");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i = static_cast<uint32_t>(lane & 31);
  for (int p = 2 * wave + (lane >> 5); p < P.n_p1; p += 2 * kJaWaves1) {
    const uint32_t n0 = S.n[p];
    const uint32_t m = n0 < static_cast<uint32_t>(kJaStage) ? n0 : static_cast<uint32_t>(kJaStage);
    if (m == 0 || m < at_least) continue;
    const unsigned long long cur = S.cur[p], end = S.end[p];
    const uint32_t k = end - cur < m ? static_cast<uint32_t>(end - cur) : m;  // records into the current block
    unsigned long long nb = S.nxt[p];
    if (k < m && nb == ~0ULL) {  // no block reserved ahead (the last flush): reserve now
      if (i == 0) nb = atomicAdd(&P.fill1[p], static_cast<unsigned long long>(kJaBlk));
      nb = __shfl(nb, lane & 32, 64);
    }
    if (i < m) {
      const unsigned long long sr = S.sr[p][i];
      const long long x = S.pay[p][i];
      const uint64_t at = i < k ? cur + i : nb + (i - k);
      if (at < P.cap1) {
        const uint64_t d = static_cast<uint64_t>(p) * P.cap1 + at;
        P.l1_sr[d] = sr;
        P.l1_pay[d] = x;
      } else {
        ja_add_direct(A, static_cast<uint32_t>(sr), x, sr >> 32);
      }
    }
    if (i == 0) {
      S.n[p] = 0;
      if (k < m) {
        S.cur[p] = nb + (m - k);
        S.end[p] = nb + kJaBlk;
        S.nxt[p] = ~0ULL;
      } else {
        S.cur[p] = cur + m;
      }
    }
  }
}

// K1: probe rows -> LDS cache, misses staged by level-1 partition.  The
// hottest slots go through the cache like the others: LDS atomics of a Zipf
// head on one address cost less than summing the top slots in registers,
// which every row paid for (probe of 256M rows: 3.15 ms against 3.36 ms with
// four register slots).
template <int KIND>
__global__ __launch_bounds__(kJaBlock1) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_ja_probe_part(TableView T, bool unique, const int64_t* __restrict__ keys,
                                                             const int64_t* __restrict__ pay, uint64_t n,
                                                             uint64_t row_base, JaAcc A,
                                                             const uint32_t* __restrict__ present, JaPart P,
                                                             const uint32_t* __restrict__ cache_keys,
                                                             unsigned long long* __restrict__ cache_part) {
  __shared__ JaCache C;
  __shared__ JaStage1 S;
  ja_cache_init(C, cache_keys);
  for (int p = threadIdx.x; p < kJaP1Max; p += blockDim.x) {
    S.n[p] = 0;
    S.cur[p] = 0;
    S.end[p] = 0;
    S.nxt[p] = ~0ULL;
  }
  __syncthreads();
  constexpr uint64_t kRound = static_cast<uint64_t>(kJaBlock1) * kJaRows1;
  const uint64_t step = static_cast<uint64_t>(gridDim.x) * kRound;
  // A round's keys and payloads are loaded during the round before it, so the
  // stream overlaps the LDS work.  Every load is unconditional (rows past the
  // end read row n - 1 and are masked where used), so the compiler can count
  // the loads in flight and wait for exactly the ones a use needs.  Streamed
  // once: non-temporal, so the presence bits and slot records stay in L2.
  const int64_t* pay_src = pay ? pay : keys;
  auto load_round = [&](uint64_t b, int64_t (&kk)[kJaRows1], int64_t (&xx)[kJaRows1]) {
#pragma unroll
    for (int j = 0; j < kJaRows1; ++j) {
      const uint64_t r = b + static_cast<uint64_t>(j) * kJaBlock1 + threadIdx.x;
      const uint64_t rc = r < n ? r : n - 1;
      kk[j] = __builtin_nontemporal_load(keys + rc);
      xx[j] = __builtin_nontemporal_load(pay_src + rc);
    }
  };
  const bool hoist = KIND == PRISM_TABLE_DENSE && present != nullptr;
  int64_t k[kJaRows1], x[kJaRows1];
  load_round(blockIdx.x * kRound, k, x);
  for (uint64_t base = blockIdx.x * kRound; base < n; base += step) {
    // DENSE: the rows' presence words are requested before the next round's
    // loads, so waiting for them does not wait for those (loads return in order)
    uint32_t pw[kJaRows1];
    if (hoist) {
#pragma unroll
      for (int j = 0; j < kJaRows1; ++j) {
        const uint64_t r = base + static_cast<uint64_t>(j) * kJaBlock1 + threadIdx.x;
        const uint64_t d = static_cast<uint64_t>(k[j]) - static_cast<uint64_t>(T.min_key);
        const bool ok = r < n && d < T.range;
        const uint32_t w = present[(ok ? d : 0) >> 5];
        pw[j] = ok ? w : 0u;
      }
    }
    int64_t kn[kJaRows1], xn[kJaRows1];
    load_round(base + step, kn, xn);
    // a partition whose block may run out in this round's flush gets its
    // next block now; the reply is read at the end of the round
    const bool want = static_cast<int>(threadIdx.x) < P.n_p1 && S.nxt[threadIdx.x] == ~0ULL &&
                      S.end[threadIdx.x] - S.cur[threadIdx.x] < static_cast<unsigned long long>(2 * kJaStage);
    unsigned long long ahead = 0;
    if (want) ahead = atomicAdd(&P.fill1[threadIdx.x], static_cast<unsigned long long>(kJaBlk));
#pragma unroll
    for (int j = 0; j < kJaRows1; ++j) {
      const uint64_t r = base + static_cast<uint64_t>(j) * kJaBlock1 + threadIdx.x;
      if (r >= n) continue;
      if (!pay) x[j] = 0;
      int64_t s;
      if (hoist) {
        const uint64_t d = static_cast<uint64_t>(k[j]) - static_cast<uint64_t>(T.min_key);
        s = ((pw[j] >> (d & 31)) & 1u) ? static_cast<int64_t>(d) : -1;
      } else {
        s = probe_slot<KIND>(T, k[j], unique, present);
      }
      if (s < 0) continue;
      const unsigned long long row = row_base + r;
      const uint32_t su = static_cast<uint32_t>(s);
      if (ja_cache_add(C, static_cast<uint32_t>(s), x[j], row)) continue;
      const uint32_t p = static_cast<uint32_t>(static_cast<uint64_t>(s) >> P.shift1);
      const uint32_t i = atomicAdd(&S.n[p], 1u);
      if (i < static_cast<uint32_t>(kJaStage)) {
        S.sr[p][i] = static_cast<unsigned long long>(su) | (row << 32);
        S.pay[p][i] = x[j];
      } else {
        ja_add_direct(A, static_cast<uint64_t>(s), x[j], row);
      }
    }
    if (want) S.nxt[threadIdx.x] = ahead;
    lds_barrier();
    ja_stage1_flush(S, P, A, kJaFlushAt);
    lds_barrier();
#pragma unroll
    for (int j = 0; j < kJaRows1; ++j) {
      k[j] = kn[j];
      x[j] = xn[j];
    }
  }
  ja_stage1_flush(S, P, A, 1);
  lds_barrier();
  // unused reserved records (the current block's rest, the block reserved
  // ahead) are marked as holes
  for (int p = threadIdx.x >> 6; p < P.n_p1; p += kJaWaves1) {
    const unsigned long long cur = S.cur[p], end = S.end[p] < P.cap1 ? S.end[p] : P.cap1;
    for (unsigned long long at = cur + (threadIdx.x & 63); at < end; at += 64)
      P.l1_sr[static_cast<uint64_t>(p) * P.cap1 + at] = kJaHole;
    const unsigned long long nx = S.nxt[p];
    if (nx != ~0ULL) {
      const unsigned long long ne = nx + kJaBlk < P.cap1 ? nx + kJaBlk : P.cap1;
      for (unsigned long long at = nx + (threadIdx.x & 63); at < ne; at += 64)
        P.l1_sr[static_cast<uint64_t>(p) * P.cap1 + at] = kJaHole;
    }
  }
  // this workgroup's totals of the slots placed at create time are added to
  // its own row of per-workgroup partials (the same row in every pass of the
  // probe call; one reduction after the last pass adds them to the slot
  // records); entries claimed at run time hold slots of this workgroup only
  // and go to the slot records directly
  __syncthreads();
  unsigned long long* row = cache_part + static_cast<uint64_t>(blockIdx.x) * kJaCache * 4;
  for (int c = threadIdx.x; c < kJaCache; c += blockDim.x) {
    const unsigned long long cnt = C.cnt[c];
    if (!cnt) continue;
    if (C.key[c] != cache_keys[c]) {  // claimed by this workgroup: its own slot
      const uint64_t s = C.key[c];
      atomicAdd(&A.w[4 * s], cnt);
      Split128 sp;
      sp.lo32 = C.lo[c];
      sp.hi32 = C.hi[c];
      uint64_t lo;
      int64_t hi;
      sp.to_i128(lo, hi);
      if (lo | static_cast<uint64_t>(hi)) atomic_add_i128(&A.w[4 * s + 1], &A.w[4 * s + 2], lo, hi);
      atomicMin(&A.w[4 * s + 3], C.first[c]);
      continue;
    }
    unsigned long long* e = row + 4 * c;
    e[0] += cnt;
    e[1] += C.lo[c];
    e[2] += static_cast<unsigned long long>(C.hi[c]);
    if (C.first[c] < e[3]) e[3] = C.first[c];
  }
}

// Per-workgroup cache partials summed per entry and added to the slot records:
// a block takes 32 entries, its 8 thread groups split the workgroup rows.
constexpr int kJaRedEntries = 32;
__global__ __launch_bounds__(256) void k_ja_cache_reduce(const unsigned long long* __restrict__ cache_part,
                                                         uint32_t n_wg, const uint32_t* __restrict__ hot, JaAcc A) {
  __shared__ unsigned long long r_cnt[8][kJaRedEntries], r_lo[8][kJaRedEntries], r_f[8][kJaRedEntries];
  __shared__ long long r_hi[8][kJaRedEntries];
  const uint32_t ce = threadIdx.x % kJaRedEntries, g = threadIdx.x / kJaRedEntries;
  const uint32_t c = blockIdx.x * kJaRedEntries + ce;
  unsigned long long cnt = 0, lo = 0, f = ~0ULL;
  long long hi = 0;
  if (c < static_cast<uint32_t>(kJaCache))
    for (uint32_t w = g; w < n_wg; w += 8) {
      const unsigned long long* e = cache_part + (static_cast<uint64_t>(w) * kJaCache + c) * 4;
      cnt += e[0];
      lo += e[1];
      hi += static_cast<long long>(e[2]);
      f = e[3] < f ? e[3] : f;
    }
  r_cnt[g][ce] = cnt;
  r_lo[g][ce] = lo;
  r_hi[g][ce] = hi;
  r_f[g][ce] = f;
  __syncthreads();
  if (g != 0 || c >= static_cast<uint32_t>(kJaCache)) return;
  for (int k = 1; k < 8; ++k) {
    cnt += r_cnt[k][ce];
    lo += r_lo[k][ce];
    hi += r_hi[k][ce];
    f = r_f[k][ce] < f ? r_f[k][ce] : f;
  }
  const uint32_t s = hot[c];
  if (s == kJaEmpty || !cnt) return;
  atomicAdd(&A.w[4 * static_cast<uint64_t>(s)], cnt);
  Split128 sp;
  sp.lo32 = lo;
  sp.hi32 = hi;
  uint64_t l;
  int64_t h;
  sp.to_i128(l, h);
  if (l | static_cast<uint64_t>(h))
    atomic_add_i128(&A.w[4 * static_cast<uint64_t>(s) + 1], &A.w[4 * static_cast<uint64_t>(s) + 2], l, h);
  atomicMin(&A.w[4 * static_cast<uint64_t>(s) + 3], f);
}

// Slot's build-side row count (the table keeps key multiplicities).
template <int KIND>
__device__ __forceinline__ uint32_t ja_slot_count(const TableView& T, uint64_t s, bool unique) {
  if constexpr (KIND == PRISM_TABLE_DENSE) {
    if (T.presence_only) return T.dense_u[s];
    if (unique) return T.has_payload ? (T.dense_up[s].x != kNone ? 1u : 0u) : (T.dense_u[s] != kNone ? 1u : 0u);
    return T.dense_nu[s].y;
  } else {
    const Slot e = T.slots[s];
    return e.key == kEmpty ? 0u : (unique ? 1u : e.cnt);
  }
}

// Hot-slot selection: a histogram of build-side row counts on a log scale
// (8 bins per doubling) gives the host a count threshold; the slots at or
// above it are listed, and the host ranks them.
constexpr int kJaHotBins = 256;
constexpr uint32_t kJaHotListMax = 4096;

__device__ __forceinline__ int ja_count_bin(uint32_t c) {
  if (c == 0) return -1;
  const int b = static_cast<int>(8.0f * __log2f(static_cast<float>(c)));
  return b < kJaHotBins - 1 ? b : kJaHotBins - 1;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_ja_count_hist(TableView T, bool unique, uint64_t n_slots,
                                                       unsigned int* hist) {
  __shared__ unsigned int h[kJaHotBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t i = blockIdx.x * 256ULL + threadIdx.x; i < n_slots; i += static_cast<uint64_t>(gridDim.x) * 256) {
    const int b = ja_count_bin(ja_slot_count<KIND>(T, i, unique));
    if (b >= 0) atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

template <int KIND>
__global__ __launch_bounds__(256) void k_ja_hot_list(TableView T, bool unique, uint64_t n_slots, int min_bin,
                                                     unsigned long long* list, unsigned int* n_list) {
  __shared__ unsigned long long l[256];
  __shared__ unsigned int ln, lbase;
  for (uint64_t i0 = blockIdx.x * 256ULL; i0 < n_slots; i0 += static_cast<uint64_t>(gridDim.x) * 256) {
    if (threadIdx.x == 0) ln = 0;
    __syncthreads();
    const uint64_t i = i0 + threadIdx.x;
    if (i < n_slots) {
      const uint32_t c = ja_slot_count<KIND>(T, i, unique);
      if (ja_count_bin(c) >= min_bin) l[atomicAdd(&ln, 1u)] = (static_cast<unsigned long long>(c) << 32) | i;
    }
    __syncthreads();
    if (threadIdx.x == 0 && ln) lbase = atomicAdd(n_list, ln);
    __syncthreads();
    if (threadIdx.x < ln && lbase + threadIdx.x < kJaHotListMax) list[lbase + threadIdx.x] = l[threadIdx.x];
    __syncthreads();
  }
}

// The hot-slot candidates in placement order - most build rows first, then
// the lower slot - as keys (~rows << 32 | slot) sorted ascending by a bitonic
// sort in LDS (one workgroup).
constexpr int kJaHotSortThreads = 1024;
__global__ __launch_bounds__(kJaHotSortThreads) void k_ja_hot_sort(unsigned long long* list, uint32_t n) {
  __shared__ unsigned long long v[kJaHotListMax];
  for (uint32_t i = threadIdx.x; i < kJaHotListMax; i += kJaHotSortThreads) {
    const unsigned long long e = i < n ? list[i] : 0ULL;
    v[i] = i < n ? (static_cast<unsigned long long>(~static_cast<uint32_t>(e >> 32)) << 32) | static_cast<uint32_t>(e)
                 : ~0ULL;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= kJaHotListMax; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < kJaHotListMax; i += kJaHotSortThreads) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const unsigned long long a = v[i], b = v[l];
          if ((a > b) == ((i & k) == 0)) {
            v[i] = b;
            v[l] = a;
          }
        }
      }
      __syncthreads();
    }
  for (uint32_t i = threadIdx.x; i < n; i += kJaHotSortThreads) list[i] = v[i];
}

// Per level-1 partition (one workgroup): its l2 region and first repartition
// block (exclusive scans of the records K1 reserved, holes included, and of
// their blocks), the l2 fill and the range run counts cleared.
__global__ __launch_bounds__(kJaPartBlock) void k_ja_regions(JaPart P) {
  const int lane = threadIdx.x & 63;
  const uint32_t n_ranges = static_cast<uint32_t>(P.n_p1) * static_cast<uint32_t>(P.fan);
  for (uint32_t r = threadIdx.x; r < n_ranges; r += kJaPartBlock) P.nruns[r] = 0;
  if (threadIdx.x >= 64) return;
  static_assert(kJaP1Max <= 64, "one wave scans the partitions");
  const uint64_t f0 = lane < P.n_p1 ? P.fill1[lane] : 0;
  const uint64_t f = f0 < P.cap1 ? f0 : P.cap1;
  const uint64_t nbk = ceil_div(f, kJaPartRecs);
  uint64_t inc = f, incb = nbk;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = __shfl_up(inc, d, 64), ob = __shfl_up(incb, d, 64);
    if (lane >= d) {
      inc += o;
      incb += ob;
    }
  }
  if (lane < P.n_p1) {
    P.off2[lane] = inc - f;
    P.fill2[lane] = 0;
    P.blk0[lane] = static_cast<uint32_t>(incb - nbk);
  }
  if (lane == 63) P.blk0[P.n_p1] = static_cast<uint32_t>(incb);
}

// K1b: repartition blocks (persistent workgroups, blocks taken by stride).
// Holes are skipped; a block's records are ranked by range with LDS counters,
// placed range by range in LDS and written out in one coalesced run.
__global__ __launch_bounds__(kJaPartBlock) void k_ja_scatter2(JaPart P) {
  __shared__ uint4 st[kJaPartRecs];
  __shared__ uint32_t s_cnt[kJaFanMax], s_start[kJaFanMax], s_blk0[kJaP1Max + 1];
  __shared__ unsigned long long s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t fan = static_cast<uint32_t>(P.fan);
  for (int i = threadIdx.x; i <= P.n_p1; i += kJaPartBlock) s_blk0[i] = P.blk0[i];
  __syncthreads();
  const uint32_t n_blocks = s_blk0[P.n_p1];
  for (uint32_t b = blockIdx.x; b < n_blocks; b += gridDim.x) {
    uint32_t p = 0;  // the partition holding block b (uniform)
    while (p + 1 < static_cast<uint32_t>(P.n_p1) && s_blk0[p + 1] <= b) ++p;
    const uint64_t f0 = P.fill1[p];
    const uint64_t fill = f0 < P.cap1 ? f0 : P.cap1;
    const uint64_t lo = static_cast<uint64_t>(b - s_blk0[p]) * kJaPartRecs;
    const uint64_t hi = lo + kJaPartRecs < fill ? lo + kJaPartRecs : fill;
    if (threadIdx.x < fan) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t r0 = static_cast<uint64_t>(p) * P.cap1;
    uint4 rec[kJaPartPerThread];
    uint32_t grp[kJaPartPerThread], pos[kJaPartPerThread];
#pragma unroll
    for (int j = 0; j < kJaPartPerThread; ++j) {
      const uint64_t i = lo + static_cast<uint64_t>(j) * kJaPartBlock + threadIdx.x;
      grp[j] = 0xFFFFFFFFu;
      if (i < hi) {
        const unsigned long long sr = __builtin_nontemporal_load(P.l1_sr + r0 + i);  // last read of the record
        const uint32_t sl = static_cast<uint32_t>(sr);
        if (sl != kJaHole) {
          const uint64_t x = static_cast<uint64_t>(__builtin_nontemporal_load(P.l1_pay + r0 + i));
          rec[j] = make_uint4(sl & (kJaSub - 1), static_cast<uint32_t>(sr >> 32), static_cast<uint32_t>(x),
                              static_cast<uint32_t>(x >> 32));
          grp[j] = (sl >> kJaSubShift) & (fan - 1);
          pos[j] = atomicAdd(&s_cnt[grp[j]], 1u);
        }
      }
    }
    __syncthreads();
    if (wave == 0) {
      const uint32_t c = static_cast<uint32_t>(lane) < fan ? s_cnt[lane] : 0u;
      uint32_t inc = c;
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
      }
      const uint32_t total = __shfl(inc, 63, 64);
      unsigned long long base = 0;
      if (lane == 0 && total) base = P.off2[p] + atomicAdd(&P.fill2[p], static_cast<unsigned long long>(total));
      base = __shfl(base, 0, 64);
      if (static_cast<uint32_t>(lane) < fan) s_start[lane] = inc - c;
      if (lane == 0) s_base = base;
      if (c) {
        const uint32_t g = p * fan + static_cast<uint32_t>(lane);
        const uint32_t k = atomicAdd(&P.nruns[g], 1u);
        P.run_off[static_cast<uint64_t>(g) * P.nb + k] = base + (inc - c);
        P.run_cnt[static_cast<uint64_t>(g) * P.nb + k] = c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kJaPartPerThread; ++j)
      if (grp[j] != 0xFFFFFFFFu) st[s_start[grp[j]] + pos[j]] = rec[j];
    __syncthreads();
    const uint32_t total = s_start[fan - 1] + s_cnt[fan - 1];
    const unsigned long long base = s_base;
    for (uint32_t i = threadIdx.x; i < total; i += kJaPartBlock) P.l2[base + i] = st[i];
    __syncthreads();  // st, s_cnt and s_start are rewritten by the next block
  }
}

// K2: one workgroup per range sums its records in LDS and adds the totals to
// the slot arrays.  The run list is read into LDS in batches; a wave takes
// kJaRunsAtOnce runs at a time, their first 64 records loaded together.
constexpr uint32_t kJaRunBatch = 2048;
constexpr int kJaRunsAtOnce = 4;
__device__ __forceinline__ void ja_agg_rec(const v4u32& r, uint32_t* a_cnt, uint32_t* a_first, unsigned long long* a_lo,
                                           long long* a_hi) {
  atomicAdd(&a_cnt[r.x], 1u);
  if (r.z) atomicAdd(&a_lo[r.x], static_cast<unsigned long long>(r.z));
  if (r.w) atomicAdd(reinterpret_cast<unsigned long long*>(&a_hi[r.x]),
                     static_cast<unsigned long long>(static_cast<long long>(static_cast<int32_t>(r.w))));
  atomicMin(&a_first[r.x], r.y);  // no return: nothing waits on it
}

__global__ __launch_bounds__(kJaPartBlock) void k_ja_agg2(JaPart P, JaAcc A) {
  __shared__ uint32_t a_cnt[kJaSub], a_first[kJaSub];
  __shared__ unsigned long long a_lo[kJaSub];
  __shared__ long long a_hi[kJaSub];
  __shared__ unsigned long long s_off[kJaRunBatch];
  __shared__ uint32_t s_m[kJaRunBatch];
  const uint32_t g = blockIdx.x;
  const uint32_t nr = P.nruns[g];
  if (nr == 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kJaSub; i += blockDim.x) {
    a_cnt[i] = 0;
    a_first[i] = 0xFFFFFFFFu;
    a_lo[i] = 0;
    a_hi[i] = 0;
  }
  const v4u32* l2 = reinterpret_cast<const v4u32*>(P.l2);
  for (uint32_t b0 = 0; b0 < nr; b0 += kJaRunBatch) {
    const uint32_t nb = nr - b0 < kJaRunBatch ? nr - b0 : kJaRunBatch;
    __syncthreads();  // the previous batch is done with s_off / s_m
    for (uint32_t k = threadIdx.x; k < nb; k += kJaPartBlock) {
      s_off[k] = P.run_off[static_cast<uint64_t>(g) * P.nb + b0 + k];
      s_m[k] = P.run_cnt[static_cast<uint64_t>(g) * P.nb + b0 + k];
    }
    __syncthreads();
    for (uint32_t k = static_cast<uint32_t>(wave) * kJaRunsAtOnce; k < nb; k += kJaWavesPB * kJaRunsAtOnce) {
      uint32_t m[kJaRunsAtOnce];
      unsigned long long off[kJaRunsAtOnce];
      v4u32 r[kJaRunsAtOnce];
#pragma unroll
      for (int u = 0; u < kJaRunsAtOnce; ++u) {
        m[u] = k + u < nb ? s_m[k + u] : 0u;
        off[u] = k + u < nb ? s_off[k + u] : 0ULL;
      }
#pragma unroll
      for (int u = 0; u < kJaRunsAtOnce; ++u)
        if (static_cast<uint32_t>(lane) < m[u]) r[u] = __builtin_nontemporal_load(l2 + off[u] + lane);  // only read
#pragma unroll
      for (int u = 0; u < kJaRunsAtOnce; ++u)
        if (static_cast<uint32_t>(lane) < m[u]) ja_agg_rec(r[u], a_cnt, a_first, a_lo, a_hi);
#pragma unroll
      for (int u = 0; u < kJaRunsAtOnce; ++u)
        for (uint32_t i = 64 + static_cast<uint32_t>(lane); i < m[u]; i += 64)
          ja_agg_rec(__builtin_nontemporal_load(l2 + off[u] + i), a_cnt, a_first, a_lo, a_hi);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kJaSub; i += blockDim.x) {
    const uint32_t c = a_cnt[i];
    if (!c) continue;
    const uint64_t s = static_cast<uint64_t>(g) * kJaSub + i;
    A.w[4 * s] += c;
    Split128 sp;
    sp.lo32 = a_lo[i];
    sp.hi32 = a_hi[i];
    uint64_t lo;
    int64_t hi;
    sp.to_i128(lo, hi);
    const uint64_t olo = A.w[4 * s + 1];
    const uint64_t nlo = olo + lo;
    A.w[4 * s + 1] = nlo;
    A.w[4 * s + 2] += static_cast<uint64_t>(hi) + (nlo < olo ? 1u : 0u);
    const unsigned long long f = a_first[i];
    if (f < A.w[4 * s + 3]) A.w[4 * s + 3] = f;
  }
}

struct JaGroups {
  unsigned long long* cnt;    // per group
  unsigned long long* sum;    // per group: lo, hi
  unsigned long long* first;  // per group: (probe row << 32) | build row
  unsigned long long* key;    // HASHED groups: attribute + 2^32 (0 = empty); null for DENSE
  int64_t attr_min;
  uint64_t n;                 // DENSE: attribute range; HASHED: slots (power of two)
  uint32_t* flags;
};

__device__ __forceinline__ int64_t ja_group(const JaGroups& G, int32_t attr) {
  if (!G.key) return static_cast<int64_t>(static_cast<int64_t>(attr) - G.attr_min);
  const unsigned long long want = static_cast<unsigned long long>(static_cast<uint32_t>(attr)) + (1ULL << 32);
  uint64_t s = (static_cast<uint64_t>(want) * 0x9E3779B97F4A7C15ULL) & (G.n - 1);
  for (uint64_t p = 0; p < G.n; ++p) {
    const unsigned long long cur = __hip_atomic_load(&G.key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == want) return static_cast<int64_t>(s);
    if (cur == 0) {
      const unsigned long long old = atomicCAS(&G.key[s], 0ULL, want);
      if (old == 0 || old == want) return static_cast<int64_t>(s);
    }
    s = (s + 1) & (G.n - 1);
  }
  atomicOr(G.flags, 1u);
  return -1;
}

template <int KIND>
__global__ void k_ja_build(TableView T, bool unique, const int64_t* __restrict__ bkeys,
                           const int32_t* __restrict__ battr, uint64_t n_build, JaAcc A, JaGroups G) {
  for (uint64_t b = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; b < n_build;
       b += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int64_t s = slot_of<KIND>(T, bkeys[b], unique);
    if (s < 0) continue;
    const unsigned long long c = A.w[4 * s];
    if (c == 0) continue;
    const int64_t g = ja_group(G, battr[b]);
    if (g < 0) continue;
    atomicAdd(&G.cnt[g], c);
    const uint64_t lo = A.w[4 * s + 1];
    const int64_t hi = static_cast<int64_t>(A.w[4 * s + 2]);
    if (lo | static_cast<uint64_t>(hi)) atomic_add_i128(&G.sum[2 * g], &G.sum[2 * g + 1], lo, hi);
    atomicMin(&G.first[g], (A.w[4 * s + 3] << 32) | b);
  }
}

// Dense groups: build rows sorted by attribute at create time (stable, so
// rows ascend within an attribute).  Sort key = attribute offset, value =
// (slot << 32) | build row.
template <int KIND>
__global__ void k_ja_attr_key(TableView T, bool unique, const int64_t* __restrict__ bkeys,
                              const int32_t* __restrict__ attr, uint64_t n, int64_t attr_min, uint32_t* __restrict__ key,
                              unsigned long long* __restrict__ val) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    key[i] = static_cast<uint32_t>(static_cast<int64_t>(attr[i]) - attr_min);
    const int64_t s = slot_of<KIND>(T, bkeys[i], unique);
    const unsigned long long sl = s < 0 ? 0xFFFFFFFFULL : static_cast<unsigned long long>(s);
    val[i] = (sl << 32) | i;
  }
}

// Fold of the per-slot probe totals into the attribute groups over the
// attribute-sorted build rows: each thread sums the runs of equal attribute
// in its kFoldRows rows; a run wholly inside the thread's rows is the whole
// group and is stored, a run cut by a chunk edge is added with atomics.
constexpr int kFoldRows = 32;

__device__ __forceinline__ void ja_group_put(const JaGroups& G, uint32_t g, bool partial, unsigned long long c,
                                             uint64_t lo, int64_t hi, unsigned long long f) {
  if (partial) {
    atomicAdd(&G.cnt[g], c);
    if (lo | static_cast<uint64_t>(hi)) atomic_add_i128(&G.sum[2 * g], &G.sum[2 * g + 1], lo, hi);
    atomicMin(&G.first[g], f);
  } else {
    G.cnt[g] = c;
    G.sum[2 * g] = lo;
    G.sum[2 * g + 1] = static_cast<unsigned long long>(hi);
    G.first[g] = f;
  }
}

__global__ void k_ja_fold(const uint32_t* __restrict__ sattr, const unsigned long long* __restrict__ sval, uint64_t n,
                          JaAcc A, JaGroups G) {
  constexpr int kAhead = 8;  // slot records in flight per thread
  const uint64_t lo = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) * kFoldRows;
  if (lo >= n) return;
  const uint64_t hi = lo + kFoldRows < n ? lo + kFoldRows : n;
  uint32_t a = sattr[lo];
  bool partial = lo > 0 && sattr[lo - 1] == a;
  unsigned long long c = 0, f = ~0ULL;
  uint64_t slo = 0;
  int64_t shi = 0;
  const uint4* rec = reinterpret_cast<const uint4*>(A.w);
  for (uint64_t i0 = lo; i0 < hi; i0 += kAhead) {
    uint32_t at[kAhead];
    unsigned long long v[kAhead];
    uint4 r0[kAhead], r1[kAhead];
#pragma unroll
    for (int j = 0; j < kAhead; ++j) {
      const uint64_t i = i0 + j < hi ? i0 + j : hi - 1;
      at[j] = sattr[i];
      v[j] = sval[i];
    }
#pragma unroll
    for (int j = 0; j < kAhead; ++j) {
      const uint32_t s = static_cast<uint32_t>(v[j] >> 32);
      const uint64_t q = s == 0xFFFFFFFFu ? 0 : 2 * static_cast<uint64_t>(s);
      r0[j] = rec[q];
      r1[j] = rec[q + 1];
    }
#pragma unroll
    for (int j = 0; j < kAhead; ++j) {
      if (i0 + j >= hi) break;
      if (at[j] != a) {
        if (c) ja_group_put(G, a, partial, c, slo, shi, f);
        a = at[j];
        partial = false;
        c = 0;
        f = ~0ULL;
        slo = 0;
        shi = 0;
      }
      if (static_cast<uint32_t>(v[j] >> 32) == 0xFFFFFFFFu) continue;
      const unsigned long long cc = (static_cast<unsigned long long>(r0[j].y) << 32) | r0[j].x;
      if (!cc) continue;
      c += cc;
      const uint64_t l = (static_cast<uint64_t>(r0[j].w) << 32) | r0[j].z;
      const uint64_t nl = slo + l;
      shi += static_cast<int64_t>((static_cast<uint64_t>(r1[j].y) << 32) | r1[j].x) + (nl < slo ? 1 : 0);
      slo = nl;
      const unsigned long long fi = (static_cast<unsigned long long>(r1[j].z) << 32) | (v[j] & 0xFFFFFFFFULL);
      f = fi < f ? fi : f;
    }
  }
  // the last run goes on in the next thread's rows when the next attribute is the same
  if (hi < n && sattr[hi] == a) partial = true;
  if (c) ja_group_put(G, a, partial, c, slo, shi, f);
}

// Groups holding rows, compacted without atomics: each block counts the
// groups of its span, then writes them after the earlier blocks' groups.
// Dense attribute groups without sorting the build rows: the rows are
// partitioned by attribute bucket (kFoldGroups consecutive attribute values;
// a counting pass per block of rows, one scan, a scatter with LDS cursors),
// then one workgroup per bucket adds each row's key totals to its group's
// accumulator in LDS and writes the bucket's groups once.  Sums are exact and
// the first pair is a minimum, so the order of the rows inside a bucket does
// not matter.
constexpr int kFoldGroups = 2048;           // groups per bucket (LDS: 32 B each)
constexpr int kFoldBlockRows = 4096;        // build rows per partition block (ordered by bucket in LDS)
constexpr int kFoldScatterThreads = 1024;
constexpr uint32_t kFoldMaxBuckets = 4096;  // a block's bucket counters in LDS (16 KB)
constexpr int kFoldThreads = 512;
constexpr int kFoldAhead = 4;               // rows (and slot records) in flight per thread

__global__ __launch_bounds__(1024) void k_fold_hist(const int32_t* __restrict__ attr, uint64_t n, int64_t attr_min,
                                                    uint32_t n_buckets, uint32_t n_blocks, uint32_t* __restrict__ hist) {
  extern __shared__ uint32_t s_h[];
  for (uint32_t i = threadIdx.x; i < n_buckets; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kFoldBlockRows;
  for (uint32_t i = threadIdx.x; i < static_cast<uint32_t>(kFoldBlockRows); i += blockDim.x) {
    const uint64_t r = base + i;
    if (r < n) atomicAdd(&s_h[static_cast<uint32_t>(static_cast<int64_t>(attr[r]) - attr_min) / kFoldGroups], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < n_buckets; b += blockDim.x)
    hist[static_cast<uint64_t>(b) * n_blocks + blockIdx.x] = s_h[b];
}

// Each row's (attribute offset, slot << 32 | row) to its bucket's place
// (slot ~0: the key has no slot; the fold skips it).  The block's rows are
// first ordered by bucket in LDS, so that consecutive threads then write
// consecutive places of each bucket's run.
template <int KIND>
__global__ __launch_bounds__(kFoldScatterThreads) void k_fold_scatter(
    TableView T, bool unique, const int64_t* __restrict__ bkeys, const int32_t* __restrict__ attr, uint64_t n,
    int64_t attr_min, uint32_t n_buckets, uint32_t n_blocks, const uint32_t* __restrict__ offs,
    uint32_t* __restrict__ sattr, unsigned long long* __restrict__ sval) {
  extern __shared__ uint32_t s_dyn[];
  uint32_t* s_cnt = s_dyn;                 // rows per bucket in this block, then their start in the block
  uint32_t* s_a = s_dyn + ((n_buckets + 1) & ~1u);  // block rows ordered by bucket: attribute offset
  uint32_t* s_b = s_a + kFoldBlockRows;    //   its bucket
  unsigned long long* s_v = reinterpret_cast<unsigned long long*>(s_b + kFoldBlockRows);  // slot << 32 | row
  __shared__ uint32_t s_wave[kFoldScatterThreads / 64];
  for (uint32_t i = threadIdx.x; i < n_buckets; i += kFoldScatterThreads) s_cnt[i] = 0;
  __syncthreads();
  constexpr int kPer = kFoldBlockRows / kFoldScatterThreads;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kFoldBlockRows;
  uint32_t a[kPer], bk[kPer], rank[kPer];
  unsigned long long v[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint64_t r = base + static_cast<uint64_t>(j) * kFoldScatterThreads + threadIdx.x;
    bk[j] = ~0u;
    if (r < n) {
      a[j] = static_cast<uint32_t>(static_cast<int64_t>(attr[r]) - attr_min);
      const int64_t sl = slot_of<KIND>(T, bkeys[r], unique);
      v[j] = ((sl < 0 ? 0xFFFFFFFFULL : static_cast<unsigned long long>(sl)) << 32) | r;
      bk[j] = a[j] / kFoldGroups;
      rank[j] = atomicAdd(&s_cnt[bk[j]], 1u);
    }
  }
  __syncthreads();
  // exclusive scan of the bucket counts in place (thread t: buckets t*q ..)
  const uint32_t q = (n_buckets + kFoldScatterThreads - 1) / kFoldScatterThreads;
  uint32_t run = 0;
  for (uint32_t i = 0; i < q; ++i) {
    const uint32_t b = threadIdx.x * q + i;
    if (b < n_buckets) run += s_cnt[b];
  }
  uint32_t inc = run;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint32_t before = inc - run;
  for (int w = 0; w < wave; ++w) before += s_wave[w];
  for (uint32_t i = 0; i < q; ++i) {
    const uint32_t b = threadIdx.x * q + i;
    if (b < n_buckets) {
      const uint32_t c = s_cnt[b];
      s_cnt[b] = before;
      before += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (bk[j] == ~0u) continue;
    const uint32_t at = s_cnt[bk[j]] + rank[j];
    s_a[at] = a[j];
    s_b[at] = bk[j];
    s_v[at] = v[j];
  }
  __syncthreads();
  const uint32_t rows = n - base < static_cast<uint64_t>(kFoldBlockRows) ? static_cast<uint32_t>(n - base)
                                                                         : static_cast<uint32_t>(kFoldBlockRows);
  for (uint32_t i = threadIdx.x; i < rows; i += kFoldScatterThreads) {
    const uint32_t b = s_b[i];
    const uint32_t at = offs[static_cast<uint64_t>(b) * n_blocks + blockIdx.x] + (i - s_cnt[b]);
    sattr[at] = s_a[i];
    sval[at] = s_v[i];
  }
}

// One workgroup per bucket: every row adds its key slot's totals (count,
// int128 sum, first probe row) to its group in LDS; the bucket's groups are
// then written with plain stores (no other workgroup has them).
__global__ __launch_bounds__(kFoldThreads) void k_fold_lds(const uint32_t* __restrict__ sattr,
                                                           const unsigned long long* __restrict__ sval,
                                                           const uint32_t* __restrict__ offs, uint32_t n_buckets,
                                                           uint32_t n_blocks, uint64_t n, JaAcc A, JaGroups G) {
  __shared__ unsigned long long s_cnt[kFoldGroups], s_lo[kFoldGroups], s_hi[kFoldGroups], s_f[kFoldGroups];
  const uint32_t b = blockIdx.x;
  for (int g = threadIdx.x; g < kFoldGroups; g += kFoldThreads) {
    s_cnt[g] = 0;
    s_lo[g] = 0;
    s_hi[g] = 0;
    s_f[g] = ~0ULL;
  }
  __syncthreads();
  const uint64_t beg = offs[static_cast<uint64_t>(b) * n_blocks];
  const uint64_t end = b + 1 < n_buckets ? offs[static_cast<uint64_t>(b + 1) * n_blocks] : n;
  const uint32_t g0 = b * static_cast<uint32_t>(kFoldGroups);
  const uint4* rec = reinterpret_cast<const uint4*>(A.w);
  for (uint64_t i0 = beg + threadIdx.x; i0 < end; i0 += static_cast<uint64_t>(kFoldThreads) * kFoldAhead) {
    uint32_t at[kFoldAhead];
    unsigned long long v[kFoldAhead];
    uint4 r0[kFoldAhead], r1[kFoldAhead];
#pragma unroll
    for (int j = 0; j < kFoldAhead; ++j) {
      const uint64_t i = i0 + static_cast<uint64_t>(j) * kFoldThreads;
      const uint64_t ic = i < end ? i : end - 1;
      at[j] = sattr[ic];
      v[j] = sval[ic];
    }
#pragma unroll
    for (int j = 0; j < kFoldAhead; ++j) {
      const uint32_t s = static_cast<uint32_t>(v[j] >> 32);
      const uint64_t q = s == 0xFFFFFFFFu ? 0 : 2 * static_cast<uint64_t>(s);
      r0[j] = rec[q];
      r1[j] = rec[q + 1];
    }
#pragma unroll
    for (int j = 0; j < kFoldAhead; ++j) {
      const uint64_t i = i0 + static_cast<uint64_t>(j) * kFoldThreads;
      if (i >= end || static_cast<uint32_t>(v[j] >> 32) == 0xFFFFFFFFu) continue;
      const unsigned long long cc = (static_cast<unsigned long long>(r0[j].y) << 32) | r0[j].x;
      if (!cc) continue;
      const uint32_t g = at[j] - g0;
      atomicAdd(&s_cnt[g], cc);
      const uint64_t l = (static_cast<uint64_t>(r0[j].w) << 32) | r0[j].z;
      const uint64_t h = (static_cast<uint64_t>(r1[j].y) << 32) | r1[j].x;
      const unsigned long long old = atomicAdd(&s_lo[g], static_cast<unsigned long long>(l));
      const uint64_t hh = h + (old + l < old ? 1u : 0u);
      if (hh) atomicAdd(&s_hi[g], static_cast<unsigned long long>(hh));
      atomicMin(&s_f[g], (static_cast<unsigned long long>(r1[j].z) << 32) | (v[j] & 0xFFFFFFFFULL));
    }
  }
  __syncthreads();
  for (int g = threadIdx.x; g < kFoldGroups; g += kFoldThreads) {
    const uint64_t gg = static_cast<uint64_t>(g0) + g;
    if (gg >= G.n || !s_cnt[g]) continue;
    G.cnt[gg] = s_cnt[g];
    G.sum[2 * gg] = s_lo[g];
    G.sum[2 * gg + 1] = s_hi[g];
    G.first[gg] = s_f[g];
  }
}

constexpr int kCollectBlocks = 1024;

__device__ __forceinline__ void ja_span(const JaGroups& G, uint64_t& lo, uint64_t& hi) {
  const uint64_t span = ceil_div(G.n, static_cast<uint64_t>(gridDim.x));
  lo = blockIdx.x * span;
  hi = lo + span < G.n ? lo + span : G.n;
}

__global__ __launch_bounds__(256) void k_ja_collect_count(JaGroups G, unsigned int* block_count) {
  __shared__ unsigned int tot;
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  uint64_t lo, hi;
  ja_span(G, lo, hi);
  unsigned int c = 0;
  for (uint64_t g = lo + threadIdx.x; g < hi; g += blockDim.x) c += G.cnt[g] != 0;
  for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
  if (lane_id() == 0 && c) atomicAdd(&tot, c);
  __syncthreads();
  if (threadIdx.x == 0) block_count[blockIdx.x] = tot;
}

// first_out: the group's first pair as a sort key, probe row << bb | build row
// (bb = 32: the pair as stored).
__global__ __launch_bounds__(256) void k_ja_collect(JaGroups G, const unsigned int* __restrict__ block_count,
                                                    unsigned long long* n_found, uint64_t* first_out, uint32_t* idx_out,
                                                    int bb) {
  __shared__ unsigned long long base;
  __shared__ unsigned int wave_n[4];
  if (threadIdx.x < 64) {  // this block's offset: the groups of the blocks before it
    unsigned long long s = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 64) s += block_count[b];
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
    if (threadIdx.x == 0) base = s;
  }
  __syncthreads();
  uint64_t lo, hi;
  ja_span(G, lo, hi);
  unsigned long long at = base;
  const int wave = threadIdx.x >> 6;
  for (uint64_t g0 = lo; g0 < hi; g0 += blockDim.x) {
    const uint64_t g = g0 + threadIdx.x;
    const bool found = g < hi && G.cnt[g] != 0;
    const uint64_t m = ballot64(found);
    if (lane_id() == 0) wave_n[wave] = __popcll(m);
    __syncthreads();
    unsigned long long before = at;
    for (int w = 0; w < wave; ++w) before += wave_n[w];
    if (found) {
      const uint64_t i = before + lanes_below(m);
      const uint64_t f = G.first[g];
      first_out[i] = ((f >> 32) << bb) | (f & 0xFFFFFFFFULL);
      idx_out[i] = static_cast<uint32_t>(g);
    }
    for (int w = 0; w < 4; ++w) at += wave_n[w];
    __syncthreads();
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *n_found = at;
}

// Output row i: attribute, count, sum lo, sum hi, first pair.
// The first n groups of `order` (n_dev: at most the count there).
__global__ void k_ja_gather(JaGroups G, const uint32_t* __restrict__ order, uint64_t n,
                            const unsigned long long* __restrict__ n_dev, int32_t* attr, unsigned long long* cnt,
                            unsigned long long* sum_lo, long long* sum_hi, unsigned long long* first) {
  if (n_dev && *n_dev < n) n = *n_dev;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t g = order[i];
    attr[i] = G.key ? static_cast<int32_t>(static_cast<uint32_t>(G.key[g])) : static_cast<int32_t>(G.attr_min + g);
    cnt[i] = G.cnt[g];
    sum_lo[i] = G.sum[2 * g];
    sum_hi[i] = static_cast<long long>(G.sum[2 * g + 1]);
    first[i] = G.first[g];
  }
}

// Slot records cleared: count 0, sum 0, first probe row "none".
__global__ void k_fill_acc(unsigned long long* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
  {
    reinterpret_cast<uint4*>(p)[2 * i] = make_uint4(0, 0, 0, 0);
    reinterpret_cast<uint4*>(p)[2 * i + 1] = make_uint4(0, 0, ~0u, ~0u);
  }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}


// min / max of an int32 column (two stages, no contended atomics).
__global__ void k_minmax_i32_part(const int32_t* __restrict__ v, uint64_t n, int64_t* __restrict__ part) {
  __shared__ int64_t lds[16];
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int64_t x = v[i];
    mn = x < mn ? x : mn;
    mx = x > mx ? x : mx;
  }
  mn = block_reduce(mn, MinOp(), lds);
  mx = block_reduce(mx, MaxOp(), lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = mn;
    part[2 * blockIdx.x + 1] = mx;
  }
}

// Final min / max of k_minmax_i32_part's partials (one workgroup of 256).
__global__ __launch_bounds__(256) void k_minmax_pairs(const int64_t* __restrict__ part, int n, int64_t* __restrict__ out) {
  __shared__ int64_t lds[16];
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (int i = threadIdx.x; i < n; i += 256) {
    mn = part[2 * i] < mn ? part[2 * i] : mn;
    mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
  }
  mn = block_reduce(mn, MinOp(), lds);
  mx = block_reduce(mx, MaxOp(), lds);
  if (threadIdx.x == 0) {
    out[0] = mn;
    out[1] = mx;
  }
}

}  // namespace join
}  // namespace prism

struct prism_join_agg {
  prism::Ctx* ctx = nullptr;
  prism_join_table* table = nullptr;
  const int64_t* build_keys = nullptr;
  const int32_t* build_attr = nullptr;
  uint64_t n_build = 0;
  uint64_t n_slots = 0;
  prism::DeviceBuffer acc;     // cnt | sum (2 words) | first, per key slot
  prism::DeviceBuffer groups;  // cnt | sum | first | key, per group
  prism::DeviceBuffer counters;
  int64_t attr_min = 0;
  uint64_t n_groups_cap = 0;
  bool dense_groups = true;
  prism::DeviceBuffer result;  // finished rows: attr(i32) cnt sum_lo sum_hi first
  uint64_t n_result = 0;
  uint64_t result_rows = 0;    // rows the result's columns are laid out for (>= n_result)
  uint64_t probe_end = 0;      // largest probe row + 1 since the last clear
  bool finished = false;
  // partitioned tail: level-1 regions (slot, row, payload arrays), level-2
  // records, per-partition and per-range counters
  prism::DeviceBuffer part_l1, part_l2, part_ctl;
  prism::DeviceBuffer hot;  // the slots with the most build rows: the LDS cache keys (u32)
  prism::DeviceBuffer cache_part;  // per K1 workgroup: cache and register totals (4 words per entry)
  // dense groups: build rows sorted by attribute: attribute offsets (u32), then
  // (slot << 32 | row) values (u64, 16-byte aligned)
  prism::DeviceBuffer fold;
  unsigned long long* fold_vals() const {
    return reinterpret_cast<unsigned long long*>(fold.as<char>() + ((n_build * 4 + 15) & ~15ULL));
  }
  uint64_t part_chunk = 0;  // probe rows per partitioned pass the scratch is sized for
  // dense groups folded by bucket: the scan of the (bucket, block) row counts,
  // which also gives each bucket's first row in `fold`; 0 buckets = `fold`
  // holds the rows sorted by attribute instead
  prism::DeviceBuffer fold_offs;
  uint32_t fold_buckets = 0, fold_blocks = 0;

  prism::join::JaAcc accs() {
    prism::join::JaAcc A;
    A.w = acc.as<unsigned long long>();
    return A;
  }
  prism::join::JaGroups groups_view() {
    prism::join::JaGroups G;
    G.cnt = groups.as<unsigned long long>();
    G.sum = G.cnt + n_groups_cap;
    G.first = G.sum + 2 * n_groups_cap;
    G.key = dense_groups ? nullptr : G.first + n_groups_cap;
    G.attr_min = attr_min;
    G.n = n_groups_cap;
    G.flags = counters.as<uint32_t>() + 2;
    return G;
  }
};

namespace prism {
namespace join {

static void ja_clear(prism_join_agg* J) {
  hipStream_t st = J->ctx->stream;
  const uint64_t S = J->n_slots;
  unsigned long long* a = J->acc.as<unsigned long long>();
  hipLaunchKernelGGL(k_fill_acc, grid_for(J->ctx, S), dim3(kBuildBlock), 0, st, a, S);
  PRISM_CHECK_LAUNCH();
  J->finished = false;
  J->n_result = 0;
  J->probe_end = 0;
}

}  // namespace join
}  // namespace prism

using namespace prism;
using namespace prism::join;

extern "C" {

static prism_status build_impl(prism_ctx* ctx, const int64_t* d_keys, const int32_t* d_payload, uint64_t n_rows,
                               int32_t kind, prism_join_table** out, bool keys_only = false,
                               const prism_join_build_opts* opts = nullptr) {
  return guarded([&] {
    if (!ctx || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or output");
    *out = nullptr;
    uint64_t part_slots = 0;
    double load = 0.5;
    int32_t probe_form = PRISM_PROBE_DIRECT;
    if (opts) {
      part_slots = opts->partition_slots;
      if (opts->load_factor != 0.0) load = opts->load_factor;
      probe_form = opts->probe;
      if (!(load > 0.0 && load <= 0.9))
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: load factor must lie in (0, 0.9]");
      if (part_slots && (part_slots & (part_slots - 1)))
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: slots per partition must be a power of two");
      if (probe_form != PRISM_PROBE_DIRECT && probe_form != PRISM_PROBE_PARTITIONED)
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown probe form");
      if (probe_form == PRISM_PROBE_PARTITIONED && kind != PRISM_TABLE_HASHED)
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: the partitioned probe needs a HASHED table");
      if (probe_form == PRISM_PROBE_PARTITIONED && (part_slots < 256 || part_slots > kPtMaxSlots))
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: the partitioned probe needs 256.." +
                                             std::to_string(kPtMaxSlots) + " slots per partition");
    }
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

    DeviceBuffer sbuf(c, sizeof(Stats) + 16 * kRedBlocks);
    Stats* dst = sbuf.as<Stats>();
    void* part = sbuf.as<char>() + sizeof(Stats);
    Stats* pinned = static_cast<Stats*>(readback_block(ctx));
    // Unique keys whose range fits n + 1024 entries (a permutation, dense ids):
    // the dense table is built from the device-side range without waiting for
    // it, and one read at the end confirms range and uniqueness.
    const bool speculate = kind != PRISM_TABLE_HASHED && !keys_only && n_rows;
    const uint64_t cap = n_rows + 1024;
    uint32_t* d = nullptr;
    uint2* up = nullptr;
    if (speculate) {
      if (d_payload) {
        T->payload.alloc(c, cap * 8);
        up = T->payload.as<uint2>();
      } else {
        T->dense.alloc(c, cap * 4);
        d = T->dense.as<uint32_t>();
      }
    }
    if (n_rows) launch_minmax(ctx, d_keys, n_rows, static_cast<int64_t*>(part), dst, d, up, speculate ? cap : 0);
    else hipLaunchKernelGGL(k_init_stats, dim3(1), dim3(1), 0, st, dst);
    PRISM_CHECK_LAUNCH();
    if (speculate) {
      hipLaunchKernelGGL(k_dense_scatter_dev, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows, cap,
                         static_cast<const Stats*>(dst), d, up, d_payload);
      hipLaunchKernelGGL(k_dense_verify_dev, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows, cap, dst,
                         static_cast<const uint32_t*>(d), static_cast<const uint2*>(up));
      PRISM_CHECK_LAUNCH();
      PRISM_HIP(hipMemcpyAsync(pinned, dst, sizeof(Stats), hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      const Stats h1 = *pinned;
      const uint64_t range = static_cast<uint64_t>(h1.max_key) - static_cast<uint64_t>(h1.min_key) + 1;
      if (h1.max_key >= h1.min_key && range != 0 && range <= cap && h1.dups == 0) {  // range 0: 2^64 wrapped
        T->info.kind = PRISM_TABLE_DENSE;
        T->info.slots = range;
        T->info.min_key = h1.min_key;
        T->info.max_key = h1.max_key;
        T->info.n_distinct = n_rows;
        T->info.max_dup = 1;
        T->info.unique = true;
        T->view.dense_u = d;
        T->view.min_key = h1.min_key;
        T->view.range = range;
        if (d_payload) {  // the packed {row, payload} entries replace the row array
          T->view.has_payload = 1;
          T->view.dense_up = up;
          T->dense.release();
          T->view.dense_u = nullptr;
        }
        T->info.device_bytes = T->slots.bytes + T->dense.bytes + T->rows.bytes + T->payload.bytes;
        *out = guard.release();
        return;
      }
      T->dense.release();
      T->payload.release();
    }
    PRISM_HIP(hipMemcpyAsync(pinned, dst, sizeof(Stats), hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    Stats hs = *pinned;
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
    bool fast = false;  // unique DENSE table built by the speculative claim (payload included)
    if (dense && !keys_only && n_rows) {
      T->dense.alloc(c, range * 4);
      uint32_t* d = T->dense.as<uint32_t>();
      uint2* up = nullptr;
      if (d_payload) {
        T->payload.alloc(c, range * 8);
        up = T->payload.as<uint2>();
      }
      hipLaunchKernelGGL(k_dense_fill2, grid_for(ctx, range), dim3(kBuildBlock), 0, st, d, up, range);
      hipLaunchKernelGGL(k_dense_scatter, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows, hs.min_key,
                         d, up, d_payload);
      hipLaunchKernelGGL(k_dense_verify, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows, hs.min_key,
                         static_cast<const uint32_t*>(d), static_cast<const uint2*>(up), &dst->dups);
      PRISM_CHECK_LAUNCH();
      Stats h2;
      PRISM_HIP(hipMemcpyAsync(&h2, dst, sizeof(Stats), hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      if (h2.dups == 0) {
        fast = true;
        T->info.kind = PRISM_TABLE_DENSE;
        T->info.slots = range;
        T->info.n_distinct = n_rows;
        T->info.max_dup = 1;
        T->info.unique = true;
        T->view.dense_u = d;
        T->view.min_key = hs.min_key;
        T->view.range = range;
        if (d_payload) {  // the packed {row, payload} entries replace the row array
          T->view.has_payload = 1;
          T->view.dense_up = up;
          T->dense.release();
          T->view.dense_u = nullptr;
        }
      } else {
        T->dense.release();
        T->payload.release();
      }
    }
    if (fast) {
      // built above
    } else if (dense) {
      T->info.kind = PRISM_TABLE_DENSE;
      T->info.slots = range;
      DeviceBuffer cnt(c, range * 4);
      hipLaunchKernelGGL(k_fill_u32, grid_for(ctx, range), dim3(kBuildBlock), 0, st, cnt.as<uint32_t>(), range, 0u);
      if (keys_only) {
        // Key presence only (the join aggregate): a presence bitmap (exact)
        // and dense_u[d] = the rows of key d among every 2^shift-th run of
        // build rows (an estimate, used to pick hot keys).  Exact counts of a
        // skewed key set cost device atomics for every key outside a
        // workgroup's LDS table (~350 us for 16M Zipf keys).
        const int shift = n_rows >= (4ULL << 20) ? 3 : 0;
        const uint64_t words = ceil_div(range, 32);
        DeviceBuffer map(c, words * 32);
        T->bits.alloc(c, words * 4);
        hipLaunchKernelGGL(k_fill_u32, grid_for(ctx, words * 8), dim3(kBuildBlock), 0, st, map.as<uint32_t>(), words * 8,
                           0u);
        if (n_rows) {
          hipLaunchKernelGGL(k_dense_mark, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows, hs.min_key,
                             map.as<uint8_t>());
          // one workgroup per CU: its LDS table sees enough sampled rows to
          // merge the repeated keys before their counts go to HBM
          hipLaunchKernelGGL(k_dense_count, dim3(static_cast<unsigned>(ctx->num_cus)), dim3(kBuildBlock), 0, st, d_keys,
                             n_rows, hs.min_key, cnt.as<uint32_t>(), nullptr, shift);
        }
        hipLaunchKernelGGL(k_pack_marks, grid_for(ctx, words), dim3(kBuildBlock), 0, st, map.as<uint4>(), words,
                           T->bits.as<uint32_t>());
        PRISM_CHECK_LAUNCH();
        T->view.present_bits = T->bits.as<uint32_t>();
        T->info.n_distinct = 0;  // not counted
        T->info.max_dup = 0;
        T->info.unique = false;
        T->dense = std::move(cnt);
        T->view.dense_u = T->dense.as<uint32_t>();
        T->view.presence_only = 1;
        T->view.min_key = T->info.min_key;
        T->view.range = range;
        T->info.device_bytes = T->slots.bytes + T->dense.bytes + T->rows.bytes + T->payload.bytes + T->bits.bytes;
        PRISM_HIP(hipStreamSynchronize(st));  // the byte map goes back to the pool
        *out = guard.release();
        return;
      }
      if (n_rows)
        hipLaunchKernelGGL(k_dense_count, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows,
                           hs.min_key, cnt.as<uint32_t>(), slot_of_row.as<uint32_t>());
      launch_count_stats(ctx, cnt.as<uint32_t>(), range, static_cast<uint64_t*>(part), dst);
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
      T->view.min_key = T->info.min_key;  // hs was reused for the count statistics
      T->view.range = range;
    } else {
      T->info.kind = PRISM_TABLE_HASHED;
      const uint64_t want = static_cast<uint64_t>(static_cast<double>(n_rows) / load) + 1;
      uint64_t S = next_pow2(want > 64 ? want : 64);
      if (part_slots && S < part_slots) S = part_slots;
      const uint64_t PS = part_slots ? part_slots : S;  // slots per partition
      T->info.slots = S;
      T->info.partition_slots = PS;
      T->info.partitions = static_cast<uint32_t>(S / PS);
      const int shift = 64 - log2_pow2(S);
      if (S / PS > 1 && n_rows) {
        // every partition must keep an empty slot, or an insert run there
        // never ends: rows per partition (an upper bound of its keys) < PS
        const uint64_t P = S / PS;
        DeviceBuffer pc(c, P * 4);
        hipLaunchKernelGGL(k_fill_u32, grid_for(ctx, P), dim3(kBuildBlock), 0, st, pc.as<uint32_t>(), P, 0u);
        hipLaunchKernelGGL(k_pt_count, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows,
                           64 - log2_pow2(P), pc.as<uint32_t>());
        PRISM_CHECK_LAUNCH();
        std::vector<uint32_t> h(P);
        PRISM_HIP(hipMemcpyAsync(h.data(), pc.ptr, P * 4, hipMemcpyDeviceToHost, st));
        PRISM_HIP(hipStreamSynchronize(st));
        const uint32_t mx = *std::max_element(h.begin(), h.end());
        if (mx >= PS)
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: a partition of " + std::to_string(PS) + " slots would hold " +
                                               std::to_string(mx) + " keys; use fewer, larger partitions");
      }
      T->slots.alloc(c, (S + 1) * sizeof(Slot));
      Slot* slots = T->slots.as<Slot>();
      hipLaunchKernelGGL(k_init_slots, grid_for(ctx, S + 1), dim3(kBuildBlock), 0, st, slots, S + 1);
      if (n_rows)
        hipLaunchKernelGGL(k_hashed_insert, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, d_keys, n_rows,
                           slots, S - 1, PS - 1, shift, slot_of_row.as<uint32_t>());
      DeviceBuffer cnt(c, (S + 1) * 4);
      hipLaunchKernelGGL(k_slot_counts, grid_for(ctx, S + 1), dim3(kBuildBlock), 0, st, slots, S + 1,
                         cnt.as<uint32_t>());
      launch_count_stats(ctx, cnt.as<uint32_t>(), S + 1, static_cast<uint64_t*>(part), dst);
      hipLaunchKernelGGL(k_copy_sentinel, dim3(1), dim3(1), 0, st, cnt.as<uint32_t>() + S, dst);
      PRISM_CHECK_LAUNCH();
      PRISM_HIP(hipMemcpyAsync(&hs, dst, sizeof(Stats), hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      T->info.n_distinct = hs.n_distinct;
      T->info.max_dup = hs.max_dup;
      T->info.unique = hs.max_dup <= 1;
      if (keys_only) {
        // slots keep key and count; no row lists
      } else if (T->info.unique) {
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
      T->view.pmask = PS - 1;
      T->view.shift = shift;
      T->view.sentinel_present = hs.sentinel != 0;
      if (probe_form == PRISM_PROBE_PARTITIONED) {
        if (!T->info.unique || keys_only)
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: the partitioned probe needs unique build keys");
        const uint64_t P = S / PS;
        // one partition is looked up in LDS beside a 32 KB tile buffer; more
        // are joined one partition per workgroup
        const uint64_t lds_need = PS * sizeof(Slot) + (P == 1 ? kTileRows * sizeof(uint2) + 4096 : 0);
        if (lds_need > ctx->lds_per_block)
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: a partition of " + std::to_string(PS) +
                                               " slots does not fit the " + std::to_string(ctx->lds_per_block) +
                                               " bytes of LDS a workgroup may use" +
                                               (P == 1 ? " beside the tile buffer" : ""));
        if (P > kPtMaxParts)
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: the partitioned probe takes at most " +
                                               std::to_string(kPtMaxParts) + " partitions (" + std::to_string(P) +
                                               " asked); use larger partitions");
        T->info.probe = PRISM_PROBE_PARTITIONED;
      }
    }
    if (d_payload && n_rows && !fast) {  // carry the build payload in the table entries
      T->view.has_payload = 1;
      if (!T->info.unique) {
        T->payload.alloc(c, n_rows * 4);
        hipLaunchKernelGGL(k_gather_i32, grid_for(ctx, n_rows), dim3(kBuildBlock), 0, st, T->payload.as<int32_t>(),
                           T->rows.as<uint32_t>(), d_payload, n_rows);
        T->view.row_payload = T->payload.as<int32_t>();
      } else if (T->info.kind == PRISM_TABLE_DENSE) {
        T->payload.alloc(c, T->info.slots * 8);
        hipLaunchKernelGGL(k_dense_pack_payload, grid_for(ctx, T->info.slots), dim3(kBuildBlock), 0, st,
                           T->payload.as<uint2>(), T->dense.as<uint32_t>(), d_payload, T->info.slots);
        T->view.dense_up = T->payload.as<uint2>();
        T->dense.release();  // the packed entries replace the row array
        T->view.dense_u = nullptr;
      } else {
        hipLaunchKernelGGL(k_slot_payload, grid_for(ctx, T->info.slots + 1), dim3(kBuildBlock), 0, st,
                           T->slots.as<Slot>(), T->info.slots + 1, d_payload);
      }
      PRISM_CHECK_LAUNCH();
    }
    PRISM_HIP(hipStreamSynchronize(st));
    T->info.device_bytes = T->slots.bytes + T->dense.bytes + T->rows.bytes + T->payload.bytes;
    *out = guard.release();
  });
}

prism_status prism_join_build_i64(prism_ctx* ctx, const int64_t* d_keys, uint64_t n_rows, int32_t kind,
                                  prism_join_table** out) {
  return build_impl(ctx, d_keys, nullptr, n_rows, kind, out);
}

prism_status prism_join_build_i64_payload(prism_ctx* ctx, const int64_t* d_keys, const int32_t* d_payload,
                                          uint64_t n_rows, int32_t kind, prism_join_table** out) {
  return build_impl(ctx, d_keys, d_payload, n_rows, kind, out);
}

prism_status prism_join_build_i64_opts(prism_ctx* ctx, const int64_t* d_keys, const int32_t* d_payload,
                                       uint64_t n_rows, const prism_join_build_opts* opts, prism_join_table** out) {
  if (!opts) return build_impl(ctx, d_keys, d_payload, n_rows, PRISM_TABLE_AUTO, out);
  return build_impl(ctx, d_keys, d_payload, n_rows, opts->kind, out, false, opts);
}

prism_status prism_join_tune(prism_ctx* ctx, const int64_t* d_keys, const int32_t* d_payload, uint64_t n_rows,
                             const int64_t* d_probe_sample, const int64_t* d_probe_payload, uint64_t n_sample,
                             prism_join_build_opts* best, prism_join_tune_candidate* cands, uint32_t cap,
                             uint32_t* n_cands) {
  return guarded([&] {
    if (!ctx || !best) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or output");
    if (n_sample && !d_probe_sample) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL probe sample");
    use_device(ctx);
    Ctx* c = as_ctx(ctx);
    std::vector<prism_join_build_opts> list;
    list.push_back({PRISM_TABLE_AUTO, PRISM_PROBE_DIRECT, 0, 0.0});
    for (double lf : {0.5, 0.7}) list.push_back({PRISM_TABLE_HASHED, PRISM_PROBE_DIRECT, 0, lf});
    for (double lf : {0.5, 0.7})
      for (uint64_t ps : {1024ULL, 2048ULL, 4096ULL, 8192ULL}) {
        const uint64_t want = static_cast<uint64_t>(static_cast<double>(n_rows) / lf) + 1;
        const uint64_t S = std::max<uint64_t>(next_pow2(want > 64 ? want : 64), ps);
        const uint64_t P = S / ps;
        if (P > kPtMaxParts || (P == 1 && ps > 4096)) continue;
        list.push_back({PRISM_TABLE_HASHED, PRISM_PROBE_PARTITIONED, ps, lf});
      }
    const uint64_t room = n_sample ? n_sample : 1;
    DeviceBuffer op(c, room * 4), ob(c, room * 4), res(c, sizeof(prism_probe_result));
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    PRISM_HIP(hipEventCreate(&e0));
    PRISM_HIP(hipEventCreate(&e1));
    PRISM_HIP(hipEventCreate(&e2));
    struct Events {
      hipEvent_t* e[3];
      ~Events() {
        for (hipEvent_t* x : e) (void)hipEventDestroy(*x);
      }
    } guard_ev{{&e0, &e1, &e2}};
    double best_ms = -1.0;
    uint32_t k = 0;
    for (const prism_join_build_opts& o : list) {
      prism_join_tune_candidate cd{};
      cd.opts = o;
      prism_join_table* t = nullptr;
      PRISM_HIP(hipEventRecord(e0, ctx->stream));
      cd.status = build_impl(ctx, d_keys, d_payload, n_rows, o.kind, &t, false, &o);
      if (cd.status == PRISM_OK) {
        PRISM_HIP(hipEventRecord(e1, ctx->stream));
        prism_status s = PRISM_OK;
        for (int rep = 0; rep < 2 && s == PRISM_OK; ++rep) {  // a warm-up, then the timed probe
          if (rep == 1) PRISM_HIP(hipEventRecord(e1, ctx->stream));
          s = prism_join_probe_i64(ctx, t, d_probe_sample, n_sample, 0, d_probe_payload, nullptr,
                                   op.as<uint32_t>(), ob.as<uint32_t>(), room, res.as<prism_probe_result>());
        }
        PRISM_HIP(hipEventRecord(e2, ctx->stream));
        PRISM_HIP(hipEventSynchronize(e2));
        float b = 0, pr = 0;
        PRISM_HIP(hipEventElapsedTime(&b, e0, e1));
        PRISM_HIP(hipEventElapsedTime(&pr, e1, e2));
        cd.build_ms = b;
        cd.probe_ms = pr;
        cd.status = s;
        if (s == PRISM_OK && (best_ms < 0 || pr < best_ms)) {
          best_ms = pr;
          *best = o;
        }
        prism_join_table_free(t);
      }
      if (cands && k < cap) cands[k] = cd;
      ++k;
    }
    clear_last_error();
    if (n_cands) *n_cands = k < cap || !cands ? k : cap;
    if (best_ms < 0) fail(PRISM_ERR_EXECUTION, "Execution error: no table layout could be built for these keys");
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

}  // extern "C"

namespace prism {
namespace join {

static uint64_t align256(uint64_t b) { return (b + 255) & ~255ULL; }
static uint64_t stage_rows_bytes(uint64_t n_tiles) { return align256(n_tiles * kTileRows * 4); }
static uint64_t stage_bits_bytes(uint64_t n_tiles) { return align256(n_tiles * kTileBallots * 8); }

static uint64_t env_u64(const char* name, uint64_t dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  const unsigned long long x = std::strtoull(v, nullptr, 10);
  return x ? x : dflt;
}

// Scratch of one probe of n_probe rows: per-tile int128 payload partials,
// then either the unique path's tile counts / offsets / slice totals / staged
// build rows and match ballots (PairStage) or the look-back status words of
// the duplicate path and the one-pass probe.  Every word the
// kernels read is written by them first (k_probe_clear clears the status
// words), so the scratch needs no clearing.
// Scratch of the partitioned probe's passes A-D (after the direct probe's):
// segment counts and offsets (partitions x tiles + 1), the scan's temporary
// storage, the records' keys and rows in tile, the per-record results.
struct PtPlan {
  bool on = false;   // the table is probed partitioned
  bool direct = false;  // one partition: looked up in LDS without passes A-D
  uint32_t P = 1, n_tiles = 0, M = 1;
  uint64_t hist = 0, offs = 0, tmp = 0, tmp_bytes = 0, rkey = 0, rrow = 0, res = 0, bytes = 0;
};

static PtPlan pt_plan(const prism_join_table* t, uint64_t n_probe) {
  PtPlan q;
  if (t->info.probe != PRISM_PROBE_PARTITIONED || !t->info.unique || t->info.kind != PRISM_TABLE_HASHED) return q;
  q.on = true;
  q.P = t->info.partitions;
  q.direct = q.P == 1;
  q.n_tiles = static_cast<uint32_t>(ceil_div(n_probe, kTileRows));
  if (q.direct || n_probe == 0) return q;
  const uint64_t cells = static_cast<uint64_t>(q.P) * q.n_tiles;
  if (cells >= (1ULL << 31)) {  // the segment table is scanned with int counts: probe directly
    q.on = false;
    return q;
  }
  q.M = static_cast<uint32_t>(ceil_div(ceil_div(n_probe, q.P), kPtChunk));
  if (q.M < 1) q.M = 1;
  size_t tb = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb, static_cast<const uint32_t*>(nullptr),
                                         static_cast<uint32_t*>(nullptr), static_cast<int>(cells));
  q.tmp_bytes = tb;
  q.hist = 0;
  q.offs = q.hist + align256(cells * 4);
  q.tmp = q.offs + align256((cells + 1) * 4);
  q.rkey = q.tmp + align256(tb);
  q.rrow = q.rkey + align256(n_probe * 8);
  q.res = q.rrow + align256(n_probe * 2);
  q.bytes = q.res + align256(n_probe * 8);
  return q;
}

// Probe form of a unique table probed directly: staged (lookup pass, tile
// scan, move pass; 8 B of staging per probe row) by default, the faster form
// on MI355X; one-pass (k_probe_unique_lb: pairs written once at offsets found
// by looking back over the tiles' counts; 40 B of scratch per 4096-row tile)
// for the allocation-free probe, whose fixed scratch the staging would bound,
// and when PRISM_PROBE_ONE_PASS=1.  The partitioned probe always stages.
static bool probe_one_pass(const prism_join_table* t, uint64_t n_probe, bool noalloc) {
  if (!t->info.unique || pt_plan(t, n_probe).on) return false;
  if (noalloc) return true;
  const char* v = std::getenv("PRISM_PROBE_ONE_PASS");
  return v && v[0] == '1';
}

static uint64_t probe_scratch_bytes(const prism_join_table* t, uint64_t n_probe, bool pairs, bool one_pass) {
  const uint64_t n_tiles = ceil_div(n_probe, kTileRows);
  const uint64_t sums_bytes = align256(n_tiles * 32 + 32);
  if (t->info.unique && !one_pass) {
    const uint64_t n_slices_max = ceil_div(n_tiles, kSliceTiles) + 1;
    const uint64_t cnt_bytes = align256(n_tiles * 4 + 4), off_bytes = align256(n_tiles * 8 + 8);
    const uint64_t slice_bytes = align256(n_slices_max * 8) + align256(n_slices_max * 32);
    const uint64_t stage_bytes = pairs ? stage_rows_bytes(n_tiles) + stage_bits_bytes(n_tiles) : 0;
    return sums_bytes + cnt_bytes + off_bytes + slice_bytes + stage_bytes + pt_plan(t, n_probe).bytes;
  }
  return sums_bytes + align256((n_tiles + 1) * 8);
}

// Passes A-E of the partitioned probe (see k_pt_restore), writing what
// k_probe_unique writes: staged build rows and ballots, tile counts, tile
// payload partials.
static void pt_probe(prism_ctx* ctx, const prism_join_table* t, const PtPlan& q, const int64_t* keys, uint64_t n,
                     const int64_t* ppay, const int32_t* bpay, PairStage stage, uint32_t* tile_count,
                     __int128* tile_sums, char* pt) {
  hipStream_t st = ctx->stream;
  const TableView& T = t->view;
  const uint64_t PS = T.pmask + 1;
  if (q.direct) {
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(q.n_tiles, static_cast<uint64_t>(ctx->num_cus) * 2));
    hipLaunchKernelGGL(k_pt_restore<true>, dim3(grid), dim3(kPB), PS * sizeof(Slot), st, T, keys, n, q.n_tiles,
                       nullptr, 1u, nullptr, nullptr, ppay, bpay, stage, tile_count, tile_sums);
    PRISM_CHECK_LAUNCH();
    return;
  }
  uint32_t* hist = reinterpret_cast<uint32_t*>(pt + q.hist);
  uint32_t* offs = reinterpret_cast<uint32_t*>(pt + q.offs);
  int64_t* rkey = reinterpret_cast<int64_t*>(pt + q.rkey);
  uint16_t* rrow = reinterpret_cast<uint16_t*>(pt + q.rrow);
  uint2* res = reinterpret_cast<uint2*>(pt + q.res);
  const int pshift = 64 - log2_pow2(q.P);
  const uint64_t cells = static_cast<uint64_t>(q.P) * q.n_tiles;
  hipLaunchKernelGGL(k_pt_hist, dim3(q.n_tiles), dim3(kPB), q.P * 4, st, keys, n, pshift, q.P, q.n_tiles, hist);
  PRISM_CHECK_LAUNCH();
  size_t tb = q.tmp_bytes;
  PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(pt + q.tmp, tb, static_cast<const uint32_t*>(hist), offs,
                                             static_cast<int>(cells), st));
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, offs + cells, static_cast<uint32_t>(n));
  hipLaunchKernelGGL(k_pt_scatter, dim3(q.n_tiles), dim3(kPB), q.P * 4, st, keys, n, pshift, q.P, q.n_tiles,
                     static_cast<const uint32_t*>(offs), rkey, rrow);
  hipLaunchKernelGGL(k_pt_join, dim3(q.P * q.M), dim3(kPB), PS * sizeof(Slot), st, T,
                     static_cast<const int64_t*>(rkey), static_cast<const uint32_t*>(offs), q.n_tiles, q.M, res);
  hipLaunchKernelGGL(k_pt_restore<false>, dim3(q.n_tiles), dim3(kPB), 0, st, T, keys, n, q.n_tiles,
                     static_cast<const uint32_t*>(offs), q.P, static_cast<const uint16_t*>(rrow),
                     static_cast<const uint2*>(res), ppay, bpay, stage, tile_count, tile_sums);
  PRISM_CHECK_LAUNCH();
}

static void probe_checks(prism_ctx* ctx, const prism_join_table* t, const int64_t* d_probe_keys, uint64_t n_probe,
                         uint32_t probe_row_base, uint32_t* d_probe_idx, uint32_t* d_build_idx,
                         prism_probe_result* d_result) {
  if (!ctx || !t || !d_result) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/table/result");
  if ((d_probe_idx == nullptr) != (d_build_idx == nullptr))
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: give both index outputs or neither");
  if (n_probe && !d_probe_keys) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL probe keys");
  if (n_probe + probe_row_base > 0xFFFFFFFFULL)
    fail(PRISM_ERR_JOIN, "Join error: probe rows exceed u32 row indices (UInt32Array, hash_join.rs:68-69)");
}

// The probe's launches over `base` (probe_scratch_bytes of scratch); the
// caller holds ctx->mu.  `timed`: bracket the main kernels with the context's
// per-kernel timers.
static void probe_launch(prism_ctx* ctx, const prism_join_table* t, const int64_t* d_probe_keys, uint64_t n_probe,
                         uint32_t probe_row_base, const int64_t* d_probe_payload, const int32_t* d_build_payload,
                         uint32_t* d_probe_idx, uint32_t* d_build_idx, uint64_t capacity, prism_probe_result* d_result,
                         char* base, bool timed, bool one_pass) {
  const uint64_t n_tiles = ceil_div(n_probe, kTileRows);
  const bool vec = (reinterpret_cast<uintptr_t>(d_probe_keys) & 15) == 0;
  const TableView& T = t->view;
  dim3 g(static_cast<unsigned>(n_tiles ? n_tiles : 1));
  const uint32_t nt = static_cast<uint32_t>(n_tiles);
  const bool pairs = d_probe_idx != nullptr;
  const uint64_t sums_bytes = align256(n_tiles * 32 + 32);
  __int128* tile_sums = reinterpret_cast<__int128*>(base);
  if (one_pass) {
    uint64_t* status = reinterpret_cast<uint64_t*>(base + sums_bytes);
    // PRISM_LOOKBACK_POLLS (tests): polls before a tile counts an unpublished
    // predecessor itself; 1 sends nearly every look-back down that path.
    const uint32_t stall_polls = static_cast<uint32_t>(std::max<uint64_t>(1, env_u64("PRISM_LOOKBACK_POLLS", kStallPolls)));
    hipLaunchKernelGGL(k_probe_clear, grid_for(ctx, n_tiles + 1), dim3(kBuildBlock), 0, ctx->stream, status,
                       n_tiles, d_result);
    PRISM_CHECK_LAUNCH();
    if (n_probe == 0) return;
    {
      std::unique_ptr<KernelTimer> kt(timed ? new KernelTimer(ctx, "join_probe") : nullptr);
#define PRISM_PROBE_LB(KIND_, VEC_)                                                                             \
  hipLaunchKernelGGL((k_probe_unique_lb<KIND_, VEC_>), g, dim3(kPB), 0, ctx->stream, T, d_probe_keys, n_probe, \
                     nt, probe_row_base, d_probe_payload, d_build_payload, d_probe_idx, d_build_idx, capacity,  \
                     status, tile_sums, d_result, stall_polls)
      if (t->info.kind == PRISM_TABLE_DENSE) {
        if (vec) PRISM_PROBE_LB(PRISM_TABLE_DENSE, true);
        else PRISM_PROBE_LB(PRISM_TABLE_DENSE, false);
      } else {
        if (vec) PRISM_PROBE_LB(PRISM_TABLE_HASHED, true);
        else PRISM_PROBE_LB(PRISM_TABLE_HASHED, false);
      }
#undef PRISM_PROBE_LB
    }
    PRISM_CHECK_LAUNCH();
    if (d_probe_payload || d_build_payload || t->view.has_payload) {
      hipLaunchKernelGGL(k_probe_sums, dim3(1), dim3(1024), 0, ctx->stream, tile_sums, n_tiles, d_result);
      PRISM_CHECK_LAUNCH();
    }
    return;
  }
  if (t->info.unique) {
    const uint64_t n_slices_max = ceil_div(n_tiles, kSliceTiles) + 1;
    const uint64_t cnt_bytes = align256(n_tiles * 4 + 4), off_bytes = align256(n_tiles * 8 + 8);
    const uint64_t slice_bytes = align256(n_slices_max * 8) + align256(n_slices_max * 32);
    char* p = base + sums_bytes;
    uint32_t* tile_count = reinterpret_cast<uint32_t*>(p);
    uint64_t* tile_off = reinterpret_cast<uint64_t*>(p + cnt_bytes);
    uint64_t* slice_total = reinterpret_cast<uint64_t*>(p + cnt_bytes + off_bytes);
    __int128* slice_sums = reinterpret_cast<__int128*>(p + cnt_bytes + off_bytes + align256(n_slices_max * 8));
    char* sp = p + cnt_bytes + off_bytes + slice_bytes;
    PairStage stage{nullptr, nullptr};
    if (pairs) {
      stage.brow = reinterpret_cast<uint32_t*>(sp);
      stage.bits = reinterpret_cast<unsigned long long*>(sp + stage_rows_bytes(n_tiles));
    }
    const uint64_t stage_bytes = pairs ? stage_rows_bytes(n_tiles) + stage_bits_bytes(n_tiles) : 0;
    hipLaunchKernelGGL(k_probe_clear, dim3(1), dim3(64), 0, ctx->stream, nullptr, 0, d_result);
    PRISM_CHECK_LAUNCH();
    if (n_probe == 0) return;
#define PRISM_PROBE1(KIND_, VEC_)                                                                          \
  hipLaunchKernelGGL((k_probe_unique<KIND_, VEC_>), g, dim3(kPB), 0, ctx->stream, T, d_probe_keys, n_probe, \
                     d_probe_payload, d_build_payload, stage, tile_count, tile_sums)
    const PtPlan q = pt_plan(t, n_probe);
    if (q.on) {
      std::unique_ptr<KernelTimer> kt(timed ? new KernelTimer(ctx, "join_probe") : nullptr);
      pt_probe(ctx, t, q, d_probe_keys, n_probe, d_probe_payload, d_build_payload, stage, tile_count, tile_sums,
               sp + stage_bytes);
    } else {
      std::unique_ptr<KernelTimer> kt(timed ? new KernelTimer(ctx, "join_probe") : nullptr);
      if (t->info.kind == PRISM_TABLE_DENSE) {
        if (vec) PRISM_PROBE1(PRISM_TABLE_DENSE, true);
        else PRISM_PROBE1(PRISM_TABLE_DENSE, false);
      } else {
        if (vec) PRISM_PROBE1(PRISM_TABLE_HASHED, true);
        else PRISM_PROBE1(PRISM_TABLE_HASHED, false);
      }
    }
#undef PRISM_PROBE1
    PRISM_CHECK_LAUNCH();
    const bool want_sums = d_probe_payload || d_build_payload || t->view.has_payload;
    const uint64_t n_slices = ceil_div(n_tiles, kSliceTiles);
    hipLaunchKernelGGL(k_probe_slices, dim3(static_cast<unsigned>(n_slices)), dim3(kSliceThreads), 0, ctx->stream,
                       tile_count, pairs ? tile_off : nullptr, want_sums ? tile_sums : nullptr, n_tiles, slice_total,
                       slice_sums);
    PRISM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_probe_finish, dim3(1), dim3(kFinishThreads), 0, ctx->stream, slice_total,
                       want_sums ? slice_sums : nullptr, n_slices, d_result);
    PRISM_CHECK_LAUNCH();
    if (pairs) {
      std::unique_ptr<KernelTimer> kt(timed ? new KernelTimer(ctx, "join_probe_move") : nullptr);
      hipLaunchKernelGGL(k_probe_move, g, dim3(kPB), 0, ctx->stream, stage, tile_count, tile_off, slice_total,
                         log2_pow2(kSliceTiles), probe_row_base, d_probe_idx, d_build_idx, capacity);
      PRISM_CHECK_LAUNCH();
    }
    return;
  }
  uint64_t* status = reinterpret_cast<uint64_t*>(base + sums_bytes);
  hipLaunchKernelGGL(k_probe_clear, grid_for(ctx, n_tiles + 1), dim3(kBuildBlock), 0, ctx->stream, status, n_tiles,
                     d_result);
  PRISM_CHECK_LAUNCH();
  if (n_probe == 0) return;
  {
    std::unique_ptr<KernelTimer> kt(timed ? new KernelTimer(ctx, "join_probe") : nullptr);
    if (t->info.kind == PRISM_TABLE_DENSE)
      launch_probe_vec<PRISM_TABLE_DENSE, false>(ctx, vec, g, T, d_probe_keys, n_probe, nt, probe_row_base,
                                                 d_probe_payload, d_build_payload, d_probe_idx, d_build_idx, capacity,
                                                 status, tile_sums, d_result);
    else
      launch_probe_vec<PRISM_TABLE_HASHED, false>(ctx, vec, g, T, d_probe_keys, n_probe, nt, probe_row_base,
                                                  d_probe_payload, d_build_payload, d_probe_idx, d_build_idx, capacity,
                                                  status, tile_sums, d_result);
  }
  PRISM_CHECK_LAUNCH();
  if (d_probe_payload || d_build_payload || t->view.has_payload) {
    hipLaunchKernelGGL(k_probe_sums, dim3(1), dim3(1024), 0, ctx->stream, tile_sums, n_tiles, d_result);
    PRISM_CHECK_LAUNCH();
  }
}

// Scratch of prism_join_probe_i64_noalloc: a fixed device array, one copy per
// GPU, held until the process ends.  Host code reaches it only through the
// address hipGetSymbolAddress returns on each device.
constexpr uint64_t kNoAllocScratchBytes = 128ULL << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int, char*> g_scratch_ptr;  // device -> address of `scratch` there

// Address of `scratch` on ctx's device (looked up once per device, with that
// device current).
static char* noalloc_scratch(prism_ctx* ctx) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(ctx->device);
  if (it != g_scratch_ptr.end()) return it->second;
  void* ptr = nullptr;
  PRISM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(PRISM_ERR_EXECUTION, "Execution error: no device address for the probe scratch array");
  g_scratch_ptr[ctx->device] = static_cast<char*>(ptr);
  return static_cast<char*>(ptr);
}

}  // namespace join
}  // namespace prism

extern "C" {

prism_status prism_join_probe_i64(prism_ctx* ctx, const prism_join_table* t, const int64_t* d_probe_keys,
                                  uint64_t n_probe, uint32_t probe_row_base, const int64_t* d_probe_payload,
                                  const int32_t* d_build_payload, uint32_t* d_probe_idx, uint32_t* d_build_idx,
                                  uint64_t capacity, prism_probe_result* d_result) {
  return guarded([&] {
    probe_checks(ctx, t, d_probe_keys, n_probe, probe_row_base, d_probe_idx, d_build_idx, d_result);
    use_device(ctx);
    Ctx* c = as_ctx(ctx);
    std::lock_guard<std::mutex> lock(ctx->mu);
    const bool one_pass = probe_one_pass(t, n_probe, false);
    c->scratch.reserve(c, probe_scratch_bytes(t, n_probe, d_probe_idx != nullptr, one_pass));
    probe_launch(ctx, t, d_probe_keys, n_probe, probe_row_base, d_probe_payload, d_build_payload, d_probe_idx,
                 d_build_idx, capacity, d_result, c->scratch.as<char>(), true, one_pass);
  });
}

prism_status prism_join_probe_i64_noalloc(prism_ctx* ctx, const prism_join_table* t, const int64_t* d_probe_keys,
                                          uint64_t n_probe, uint32_t probe_row_base, const int64_t* d_probe_payload,
                                          const int32_t* d_build_payload, uint32_t* d_probe_idx,
                                          uint32_t* d_build_idx, uint64_t capacity, prism_probe_result* d_result) {
  return guarded([&] {
    probe_checks(ctx, t, d_probe_keys, n_probe, probe_row_base, d_probe_idx, d_build_idx, d_result);
    const bool one_pass = probe_one_pass(t, n_probe, true);
    const uint64_t need = probe_scratch_bytes(t, n_probe, d_probe_idx != nullptr, one_pass);
    if (need > kNoAllocScratchBytes)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: probe of " + std::to_string(n_probe) + " rows needs " +
                                           std::to_string(need) + " bytes of scratch; the allocation-free probe has " +
                                           std::to_string(kNoAllocScratchBytes));
    use_device(ctx);
    char* base = noalloc_scratch(ctx);
    std::lock_guard<std::mutex> lock(ctx->mu);
    probe_launch(ctx, t, d_probe_keys, n_probe, probe_row_base, d_probe_payload, d_build_payload, d_probe_idx,
                 d_build_idx, capacity, d_result, base, false, one_pass);
  });
}

static bool t_kind_dense(const prism_join_table* t) { return t->info.kind == PRISM_TABLE_DENSE; }

// The slots with the most build rows go into the LDS cache's buckets (first
// free way of either bucket, in rank order).  hot buffer: the kJaCache cache
// keys (u32, kJaEmpty = free way).  Three steps, so that the create's other
// device work runs while the host waits and places: a histogram of the slots'
// row counts (read back with the create's first wait), the candidates at or
// above the count bin that keeps at most kJaHotListMax of them (listed and
// ordered on the device, read back behind an event), and their placement.
constexpr size_t kRbJaAttr = 4096, kRbJaHist = 5120, kRbJaHot = 8192, kRbJaList = 16384;
static_assert(kRbJaHist + kJaHotBins * 4 <= kRbJaHot && kRbJaHot + kJaCache * 4 <= kRbJaList &&
                  kRbJaList + kJaHotListMax * 8 <= kReadbackBytes,
              "join aggregate readbacks fit the readback block");

struct JaHotSel {
  DeviceBuffer scratch;
  uint64_t taken = 0;
  hipEvent_t listed = nullptr;
  ~JaHotSel() {
    if (listed) (void)hipEventDestroy(listed);
  }
};

static void ja_hot_count(prism_join_agg* J, hipStream_t st, JaHotSel& H) {
  prism_join_table* t = J->table;
  const uint64_t n_slots = t->info.slots;  // HASHED: the INT64_MIN sentinel slot is never hot
  H.scratch.alloc(J->ctx, kJaHotBins * 4 + 16 + kJaHotListMax * 8);
  unsigned int* hist = H.scratch.as<unsigned int>();
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, hist, static_cast<uint64_t>(kJaHotBins + 4), 0u);
  const dim3 g(n_slots ? static_cast<unsigned>(std::min<uint64_t>(ceil_div(n_slots, 256), 1024)) : 1);
  if (t->info.kind == PRISM_TABLE_DENSE)
    hipLaunchKernelGGL(k_ja_count_hist<PRISM_TABLE_DENSE>, g, dim3(256), 0, st, t->view, t->info.unique, n_slots, hist);
  else
    hipLaunchKernelGGL(k_ja_count_hist<PRISM_TABLE_HASHED>, g, dim3(256), 0, st, t->view, t->info.unique, n_slots, hist);
  PRISM_CHECK_LAUNCH();
  char* rb = static_cast<char*>(readback_block(J->ctx));
  PRISM_HIP(hipMemcpyAsync(rb + kRbJaHist, hist, kJaHotBins * 4, hipMemcpyDeviceToHost, st));
}

// After the wait that covers ja_hot_count's readback.
static void ja_hot_list(prism_join_agg* J, hipStream_t st, JaHotSel& H) {
  prism_join_table* t = J->table;
  const uint64_t n_slots = t->info.slots;
  char* rb = static_cast<char*>(readback_block(J->ctx));
  const unsigned int* h = reinterpret_cast<const unsigned int*>(rb + kRbJaHist);
  int min_bin = kJaHotBins;
  uint64_t taken = 0;
  while (min_bin > 0 && taken + h[min_bin - 1] <= kJaHotListMax) taken += h[--min_bin];
  H.taken = taken;
  if (!taken) return;
  unsigned int* n_list = H.scratch.as<unsigned int>() + kJaHotBins;
  unsigned long long* list = reinterpret_cast<unsigned long long*>(H.scratch.as<char>() + kJaHotBins * 4 + 16);
  const dim3 g(n_slots ? static_cast<unsigned>(std::min<uint64_t>(ceil_div(n_slots, 256), 1024)) : 1);
  if (t->info.kind == PRISM_TABLE_DENSE)
    hipLaunchKernelGGL(k_ja_hot_list<PRISM_TABLE_DENSE>, g, dim3(256), 0, st, t->view, t->info.unique, n_slots, min_bin,
                       list, n_list);
  else
    hipLaunchKernelGGL(k_ja_hot_list<PRISM_TABLE_HASHED>, g, dim3(256), 0, st, t->view, t->info.unique, n_slots, min_bin,
                       list, n_list);
  // most rows first, then the lower slot (the slot stays in the low word)
  hipLaunchKernelGGL(k_ja_hot_sort, dim3(1), dim3(kJaHotSortThreads), 0, st, list, static_cast<uint32_t>(taken));
  PRISM_CHECK_LAUNCH();
  PRISM_HIP(hipMemcpyAsync(rb + kRbJaList, list, taken * 8, hipMemcpyDeviceToHost, st));
  PRISM_HIP(hipEventCreateWithFlags(&H.listed, hipEventDisableTiming));
  PRISM_HIP(hipEventRecord(H.listed, st));
}

// Waits for the ordered candidates only (work queued after them keeps
// running), places them and queues the upload of the cache keys.
static void ja_hot_place(prism_join_agg* J, hipStream_t st, JaHotSel& H) {
  char* rb = static_cast<char*>(readback_block(J->ctx));
  uint32_t* keys = reinterpret_cast<uint32_t*>(rb + kRbJaHot);
  for (int i = 0; i < kJaCache; ++i) keys[i] = kJaEmpty;
  if (H.taken) {
    PRISM_HIP(hipEventSynchronize(H.listed));
    const unsigned long long* l = reinterpret_cast<const unsigned long long*>(rb + kRbJaList);
    for (uint64_t i = 0; i < H.taken; ++i) {
      const uint32_t sl = static_cast<uint32_t>(l[i]);
      bool placed = false;
      for (int tb = 0; tb < 2 && !placed; ++tb) {
        const uint32_t b = ja_bucket(sl, tb);
        for (int w = 0; w < kJaWays && !placed; ++w)
          if (keys[kJaWays * b + w] == kJaEmpty) {
            keys[kJaWays * b + w] = sl;
            placed = true;
          }
      }
    }
  }
  J->hot.alloc(J->ctx, kJaCache * 4);
  PRISM_HIP(hipMemcpyAsync(J->hot.ptr, keys, kJaCache * 4, hipMemcpyHostToDevice, st));
}

prism_status prism_join_agg_create(prism_ctx* ctx, const int64_t* d_build_keys, const int32_t* d_build_attr,
                                   uint64_t n_build, prism_join_agg** out) {
  return guarded([&] {
    if (!ctx || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or output");
    *out = nullptr;
    if (n_build && (!d_build_keys || !d_build_attr))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL build keys or attribute");
    auto J = std::make_unique<prism_join_agg>();
    prism_join_table* t = nullptr;
    prism_status st0 = build_impl(ctx, d_build_keys, nullptr, n_build, PRISM_TABLE_AUTO, &t, true);
    if (st0 != PRISM_OK) fail(st0, prism_last_error());
    J->table = t;
    use_device(ctx);
    Ctx* c = as_ctx(ctx);
    std::lock_guard<std::mutex> lock(ctx->mu);
    J->ctx = c;
    J->build_keys = d_build_keys;
    J->build_attr = d_build_attr;
    J->n_build = n_build;
    J->n_slots = t->info.kind == PRISM_TABLE_DENSE ? (t->info.slots ? t->info.slots : 1) : t->info.slots + 2;
    if (J->n_slots >= 0xFFFFFFFFULL) fail(PRISM_ERR_JOIN, "Join error: build side too large for the join aggregate");
    J->acc.alloc(c, J->n_slots * 4 * 8);
    J->counters.alloc(c, 64);
    hipStream_t st = ctx->stream;
    // attribute range (dense groups when it is small, else a hashed group
    // table), read back with the hot-slot histogram
    int64_t* h_attr = reinterpret_cast<int64_t*>(static_cast<char*>(readback_block(ctx)) + kRbJaAttr);
    DeviceBuffer apart;
    if (n_build) {
      apart.alloc(c, kRedBlocks * 16 + 16);
      int64_t* pa = apart.as<int64_t>();
      hipLaunchKernelGGL(k_minmax_i32_part, dim3(kRedBlocks), dim3(kBuildBlock), 0, st, d_build_attr, n_build, pa);
      hipLaunchKernelGGL(k_minmax_pairs, dim3(1), dim3(256), 0, st, pa, kRedBlocks, pa + 2 * kRedBlocks);
      PRISM_CHECK_LAUNCH();
      PRISM_HIP(hipMemcpyAsync(h_attr, pa + 2 * kRedBlocks, 16, hipMemcpyDeviceToHost, st));
    }
    JaHotSel hot;
    ja_hot_count(J.get(), st, hot);
    PRISM_HIP(hipStreamSynchronize(st));
    int64_t lo = 0, hi = -1;
    if (n_build) {
      lo = h_attr[0];
      hi = h_attr[1];
    }
    ja_hot_list(J.get(), st, hot);
    // fold set-up buffers, released after the final wait below
    DeviceBuffer fold_hist, fold_tmp;
    const uint64_t range = hi >= lo ? static_cast<uint64_t>(hi - lo) + 1 : 1;
    J->attr_min = lo;
    if (range <= (64ULL << 20) && range <= 4 * n_build + 1024) {
      J->dense_groups = true;
      J->n_groups_cap = range;
      J->groups.alloc(c, range * 4 * 8);
      const uint64_t n_buckets = ceil_div(range, kFoldGroups);
      if (n_build && n_buckets <= kFoldMaxBuckets) {
        // build rows partitioned by attribute bucket, for the LDS fold
        J->fold.alloc(c, n_build * 12 + 16);
        const uint32_t nbk = static_cast<uint32_t>(n_buckets);
        const uint32_t nbl = static_cast<uint32_t>(ceil_div(n_build, kFoldBlockRows));
        const uint64_t cells = static_cast<uint64_t>(nbk) * nbl;
        J->fold_offs.alloc(c, cells * 4 + 16);
        fold_hist.alloc(c, cells * 4 + 16);
        uint32_t* hist = fold_hist.as<uint32_t>();
        hipLaunchKernelGGL(k_fold_hist, dim3(nbl), dim3(1024), nbk * 4, st, d_build_attr, n_build, lo, nbk, nbl,
                           hist);
        PRISM_CHECK_LAUNCH();
        size_t tmp = 0;
        PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, hist, J->fold_offs.as<uint32_t>(),
                                                   static_cast<int>(cells), st));
        fold_tmp.alloc(c, tmp);
        PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(fold_tmp.ptr, tmp, hist, J->fold_offs.as<uint32_t>(),
                                                   static_cast<int>(cells), st));
        const size_t scatter_lds = ((nbk + 1) & ~1u) * 4 + static_cast<size_t>(kFoldBlockRows) * 16;
        if (t_kind_dense(J->table))
          hipLaunchKernelGGL(k_fold_scatter<PRISM_TABLE_DENSE>, dim3(nbl), dim3(kFoldScatterThreads), scatter_lds, st,
                             J->table->view,
                             J->table->info.unique, d_build_keys, d_build_attr, n_build, lo, nbk, nbl,
                             J->fold_offs.as<uint32_t>(), J->fold.as<uint32_t>(), J->fold_vals());
        else
          hipLaunchKernelGGL(k_fold_scatter<PRISM_TABLE_HASHED>, dim3(nbl), dim3(kFoldScatterThreads), scatter_lds, st,
                             J->table->view,
                             J->table->info.unique, d_build_keys, d_build_attr, n_build, lo, nbk, nbl,
                             J->fold_offs.as<uint32_t>(), J->fold.as<uint32_t>(), J->fold_vals());
        PRISM_CHECK_LAUNCH();
        J->fold_buckets = nbk;
        J->fold_blocks = nbl;
      } else if (n_build) {
        // build rows sorted by attribute once, for the atomic-free fold
        J->fold.alloc(c, n_build * 12 + 16);
        uint32_t* sattr = J->fold.as<uint32_t>();
        unsigned long long* sval = J->fold_vals();
        DeviceBuffer tmp_kv(c, n_build * 12 + 16);
        uint32_t* k_in = tmp_kv.as<uint32_t>();
        unsigned long long* v_in = reinterpret_cast<unsigned long long*>(tmp_kv.as<char>() + ((n_build * 4 + 15) & ~15ULL));
        if (t_kind_dense(J->table))
          hipLaunchKernelGGL(k_ja_attr_key<PRISM_TABLE_DENSE>, grid_for(c, n_build), dim3(kBuildBlock), 0, st,
                             J->table->view, J->table->info.unique, d_build_keys, d_build_attr, n_build, lo, k_in, v_in);
        else
          hipLaunchKernelGGL(k_ja_attr_key<PRISM_TABLE_HASHED>, grid_for(c, n_build), dim3(kBuildBlock), 0, st,
                             J->table->view, J->table->info.unique, d_build_keys, d_build_attr, n_build, lo, k_in, v_in);
        PRISM_CHECK_LAUNCH();
        int bits = 1;
        while (bits < 32 && (1ULL << bits) < range) ++bits;
        size_t tmp = 0;
        PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k_in, sattr, v_in, sval, static_cast<int>(n_build), 0,
                                                     bits, st));
        DeviceBuffer t(c, tmp);
        PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(t.ptr, tmp, k_in, sattr, v_in, sval, static_cast<int>(n_build), 0,
                                                     bits, st));
        PRISM_HIP(hipStreamSynchronize(st));
      }
    } else {
      J->dense_groups = false;
      J->n_groups_cap = next_pow2(2 * n_build > 1024 ? 2 * n_build : 1024);
      J->groups.alloc(c, J->n_groups_cap * 5 * 8);
    }
    ja_clear(J.get());
    ja_hot_place(J.get(), st, hot);
    PRISM_HIP(hipStreamSynchronize(st));
    *out = J.release();
  });
}

namespace prism {
namespace join {

// Shape of the partitioned tail for a slot space: ranges of kJaSub slots, at
// most kJaP1Max level-1 partitions of `fan` ranges each.  False when the slot
// space needs more ranges than two levels of staging cover.
struct JaShape {
  int n_p1 = 0, fan = 1, shift1 = kJaSubShift;
};
static bool ja_shape(uint64_t n_slots, JaShape& sh) {
  const uint64_t ranges = ceil_div(n_slots, kJaSub);
  uint64_t fan = next_pow2(ceil_div(ranges, kJaP1Max));
  if (fan < 1) fan = 1;
  if (fan > static_cast<uint64_t>(kJaFanMax)) return false;
  sh.fan = static_cast<int>(fan);
  sh.shift1 = kJaSubShift + log2_pow2(fan);
  sh.n_p1 = static_cast<int>(ceil_div(ranges, fan));
  return true;
}


}  // namespace join
}  // namespace prism

prism_status prism_join_agg_probe(prism_join_agg* J, const int64_t* d_probe_keys, const int64_t* d_probe_payload,
                                  uint64_t n_probe, uint64_t probe_row_base) {
  return guarded([&] {
    if (!J) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL join aggregate");
    if (n_probe && !d_probe_keys) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL probe keys");
    if (n_probe + probe_row_base > 0xFFFFFFFFULL)
      fail(PRISM_ERR_JOIN, "Join error: probe rows exceed u32 row indices (UInt32Array, hash_join.rs:68-69)");
    use_device(J->ctx);
    std::lock_guard<std::mutex> lock(J->ctx->mu);
    J->finished = false;
    if (n_probe == 0 || J->n_build == 0) return;
    J->probe_end = std::max<uint64_t>(J->probe_end, probe_row_base + n_probe);
    Ctx* c = J->ctx;
    hipStream_t st = c->stream;
    const TableView& T = J->table->view;
    const bool unique = J->table->info.unique;
    const bool dense = J->table->info.kind == PRISM_TABLE_DENSE;
    const uint32_t* present = dense ? J->table->view.present_bits : nullptr;
    JaShape sh;
    // PRISM_JA_PART=0 keeps every cache miss on device atomics (the
    // unpartitioned pass); PRISM_JA_CHUNK sets the probe rows per partitioned
    // pass (scratch: about 52 B per row of a pass).
    const bool part = env_u64("PRISM_JA_PART", 1) != 0 && ja_shape(J->n_slots, sh);
    KernelTimer kt(c, "join_agg_probe");
    if (!part) {
      uint64_t blocks = ceil_div(n_probe, kJaBlock * 8);
      const uint64_t cap = static_cast<uint64_t>(c->num_cus) * 4;
      if (blocks > cap) blocks = cap;
      if (dense)
        hipLaunchKernelGGL(k_ja_probe<PRISM_TABLE_DENSE>, dim3(static_cast<unsigned>(blocks)), dim3(kJaBlock), 0, st,
                           T, unique, d_probe_keys, d_probe_payload, n_probe, probe_row_base, J->accs(), present,
                           J->hot.as<uint32_t>());
      else
        hipLaunchKernelGGL(k_ja_probe<PRISM_TABLE_HASHED>, dim3(static_cast<unsigned>(blocks)), dim3(kJaBlock), 0, st,
                           T, unique, d_probe_keys, d_probe_payload, n_probe, probe_row_base, J->accs(), present,
                           J->hot.as<uint32_t>());
      PRISM_CHECK_LAUNCH();
      return;
    }
    // Larger passes amortise each pass's fixed work (the staging tails, the
    // region scan): for 256M probe rows, one pass of 256M rows (~13 GB of
    // scratch) measured 3.36 ms, passes of 128M 3.46 ms, of 64M 3.60 ms, of
    // 32M 4.43 ms (with the counted repartition before the run lists).  A
    // device with less free memory gets passes whose scratch fits a quarter
    // of it.
    uint64_t chunk = env_u64("PRISM_JA_CHUNK", 0);
    if (chunk == 0) {
      chunk = 256ULL << 20;
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        const uint64_t fit = static_cast<uint64_t>(free_b) / 4 / 52;
        if (fit < chunk) chunk = fit > (1ULL << 20) ? fit : (1ULL << 20);
      }
    }
    if (chunk > n_probe) chunk = n_probe;
    const uint64_t n_ranges = static_cast<uint64_t>(sh.n_p1) * sh.fan;
    // twice a partition's share of the pass, plus the unused tails of the
    // blocks K1 workgroups reserve
    uint64_t g1max = ceil_div(chunk, static_cast<uint64_t>(kJaBlock1) * kJaRows1);
    if (g1max > 2 * static_cast<uint64_t>(c->num_cus)) g1max = 2 * static_cast<uint64_t>(c->num_cus);
    uint64_t cap1 = 2 * chunk / sh.n_p1 + 64;
    if (cap1 > chunk) cap1 = chunk;
    cap1 += g1max * kJaBlk;
    // l2 holds what K1 reserved: the pass's rows plus the unused tails of the
    // current and next blocks of every K1 workgroup and partition
    const uint64_t l2_recs = chunk + g1max * sh.n_p1 * 2 * kJaBlk;
    if (chunk > J->part_chunk || J->part_l1.bytes < sh.n_p1 * cap1 * 16 || J->part_l2.bytes < l2_recs * 16) {
      J->part_l1.reserve(c, sh.n_p1 * cap1 * 16);
      J->part_l2.reserve(c, l2_recs * 16);
      J->part_chunk = chunk;
    }
    const uint32_t nb = static_cast<uint32_t>(ceil_div(cap1, kJaPartRecs));
    const uint64_t n_runs = n_ranges * nb;
    J->part_ctl.reserve(c, sh.n_p1 * 24 + (sh.n_p1 + 1) * 4 + n_ranges * 4 + n_runs * 12 + 256);
    JaPart P;
    char* l1 = J->part_l1.as<char>();
    P.l1_sr = reinterpret_cast<unsigned long long*>(l1);
    P.l1_pay = reinterpret_cast<long long*>(P.l1_sr + sh.n_p1 * cap1);
    P.cap1 = cap1;
    P.shift1 = sh.shift1;
    P.fan = sh.fan;
    P.n_p1 = sh.n_p1;
    char* ctl = J->part_ctl.as<char>();
    P.fill1 = reinterpret_cast<unsigned long long*>(ctl);
    P.off2 = P.fill1 + sh.n_p1;
    P.fill2 = P.off2 + sh.n_p1;
    P.run_off = P.fill2 + sh.n_p1;
    P.run_cnt = reinterpret_cast<uint32_t*>(P.run_off + n_runs);
    P.nruns = P.run_cnt + n_runs;
    P.blk0 = P.nruns + n_ranges;
    P.nb = nb;
    P.l2 = J->part_l2.as<uint4>();
    const JaAcc A = J->accs();
    J->cache_part.reserve(c, g1max * kJaCache * 32);
    unsigned long long* cpart = J->cache_part.as<unsigned long long>();
    hipLaunchKernelGGL(k_fill_acc, grid_for(c, g1max * kJaCache), dim3(kBuildBlock), 0, st, cpart,
                       g1max * kJaCache);
    for (uint64_t at = 0; at < n_probe; at += chunk) {
      const uint64_t m = n_probe - at < chunk ? n_probe - at : chunk;
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, st, P.fill1, static_cast<uint64_t>(sh.n_p1), 0ULL);
      uint64_t g1 = ceil_div(m, static_cast<uint64_t>(kJaBlock1) * kJaRows1);
      if (g1 > 2 * static_cast<uint64_t>(c->num_cus)) g1 = 2 * static_cast<uint64_t>(c->num_cus);
      const int64_t* kp = d_probe_keys + at;
      const int64_t* pp = d_probe_payload ? d_probe_payload + at : nullptr;
      if (dense)
        hipLaunchKernelGGL(k_ja_probe_part<PRISM_TABLE_DENSE>, dim3(static_cast<unsigned>(g1)), dim3(kJaBlock1), 0, st,
                           T, unique, kp, pp, m, probe_row_base + at, A, present, P, J->hot.as<uint32_t>(), cpart);
      else
        hipLaunchKernelGGL(k_ja_probe_part<PRISM_TABLE_HASHED>, dim3(static_cast<unsigned>(g1)), dim3(kJaBlock1), 0,
                           st, T, unique, kp, pp, m, probe_row_base + at, A, present, P, J->hot.as<uint32_t>(), cpart);
      hipLaunchKernelGGL(k_ja_regions, dim3(1), dim3(kJaPartBlock), 0, st, P);
      hipLaunchKernelGGL(k_ja_scatter2, dim3(static_cast<unsigned>(c->num_cus) * 2), dim3(kJaPartBlock), 0, st, P);
      hipLaunchKernelGGL(k_ja_agg2, dim3(static_cast<unsigned>(n_ranges)), dim3(kJaPartBlock), 0, st, P, A);
      PRISM_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ja_cache_reduce, dim3(static_cast<unsigned>(ceil_div(kJaCache, kJaRedEntries))), dim3(256), 0, st,
                       cpart, static_cast<uint32_t>(g1max), J->hot.as<uint32_t>(), A);
    PRISM_CHECK_LAUNCH();
  });
}

prism_status prism_join_agg_finish(prism_join_agg* J, uint64_t* n_groups) {
  return guarded([&] {
    if (!J || !n_groups) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL join aggregate/output");
    use_device(J->ctx);
    std::lock_guard<std::mutex> lock(J->ctx->mu);
    Ctx* c = J->ctx;
    hipStream_t st = c->stream;
    const uint64_t G = J->n_groups_cap;
    unsigned long long* g = J->groups.as<unsigned long long>();
    hipLaunchKernelGGL(k_fill_u64, grid_for(c, 3 * G), dim3(kBuildBlock), 0, st, g, 3 * G, 0ULL);
    hipLaunchKernelGGL(k_fill_u64, grid_for(c, G), dim3(kBuildBlock), 0, st, g + 3 * G, G, ~0ULL);
    if (!J->dense_groups) hipLaunchKernelGGL(k_fill_u64, grid_for(c, G), dim3(kBuildBlock), 0, st, g + 4 * G, G, 0ULL);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, J->counters.as<unsigned long long>(), 8, 0ULL);
    PRISM_CHECK_LAUNCH();
    const TableView& T = J->table->view;
    const bool unique = J->table->info.unique;
    JaGroups GV = J->groups_view();
    if (J->n_build && J->dense_groups && J->fold_buckets) {
      KernelTimer kt(c, "join_agg_build");
      hipLaunchKernelGGL(k_fold_lds, dim3(J->fold_buckets), dim3(kFoldThreads), 0, st, J->fold.as<uint32_t>(),
                         J->fold_vals(), J->fold_offs.as<uint32_t>(), J->fold_buckets, J->fold_blocks, J->n_build,
                         J->accs(), GV);
      PRISM_CHECK_LAUNCH();
    } else if (J->n_build && J->dense_groups) {
      KernelTimer kt(c, "join_agg_build");
      const uint64_t threads = ceil_div(J->n_build, kFoldRows);
      hipLaunchKernelGGL(k_ja_fold, dim3(static_cast<unsigned>(ceil_div(threads, kBuildBlock))), dim3(kBuildBlock), 0,
                         st, J->fold.as<uint32_t>(), J->fold_vals(), J->n_build, J->accs(), GV);
      PRISM_CHECK_LAUNCH();
    } else if (J->n_build) {
      KernelTimer kt(c, "join_agg_build");
      if (J->table->info.kind == PRISM_TABLE_DENSE)
        hipLaunchKernelGGL(k_ja_build<PRISM_TABLE_DENSE>, grid_for(c, J->n_build), dim3(kBuildBlock), 0, st, T, unique,
                           J->build_keys, J->build_attr, J->n_build, J->accs(), GV);
      else
        hipLaunchKernelGGL(k_ja_build<PRISM_TABLE_HASHED>, grid_for(c, J->n_build), dim3(kBuildBlock), 0, st, T,
                           unique, J->build_keys, J->build_attr, J->n_build, J->accs(), GV);
      PRISM_CHECK_LAUNCH();
    }
    DeviceBuffer first(c, G * 8), first_sorted(c, G * 8), idx(c, G * 4), idx_sorted(c, G * 4);
    unsigned long long* found_p = J->counters.as<unsigned long long>();
    // Sort keys: the first pair as probe row << bb | build row, bb the bits of
    // the largest build row, when both fit 63 bits (fewer radix passes).
    auto bits_of = [](uint64_t v) { return v <= 1 ? 1 : 64 - __builtin_clzll(v - 1); };
    const int bb0 = bits_of(J->n_build), pb = bits_of(J->probe_end);
    const bool compact = bb0 < 32 && bb0 + pb <= 63;
    const int bb = compact ? bb0 : 32, end_bit = compact ? bb0 + pb : 64;
    // Dense groups of at most 4M: every group slot is sorted (the missing
    // ones carry an all-ones key and sort last) and the gather reads the
    // group count on the device, so the finish waits only once, at its end.
    const bool one_wait = J->dense_groups && G <= (4ULL << 20);
    if (one_wait) hipLaunchKernelGGL(k_fill_u64, grid_for(c, G), dim3(kBuildBlock), 0, st, first.as<unsigned long long>(), G, ~0ULL);
    DeviceBuffer bcount(c, kCollectBlocks * 4);
    hipLaunchKernelGGL(k_ja_collect_count, dim3(kCollectBlocks), dim3(256), 0, st, GV, bcount.as<unsigned int>());
    hipLaunchKernelGGL(k_ja_collect, dim3(kCollectBlocks), dim3(256), 0, st, GV, bcount.as<unsigned int>(), found_p,
                       first.as<uint64_t>(), idx.as<uint32_t>(), bb);
    PRISM_CHECK_LAUNCH();
    unsigned long long* h = static_cast<unsigned long long*>(readback_block(c));
    uint64_t found = 0, n_sort = G;
    if (!one_wait) {
      PRISM_HIP(hipMemcpyAsync(h, J->counters.ptr, 16, hipMemcpyDeviceToHost, st));
      PRISM_HIP(hipStreamSynchronize(st));
      found = h[0];
      if (static_cast<uint32_t>(h[1]) & 1u) fail(PRISM_ERR_INTERNAL, "Internal error: join aggregate group table full");
      n_sort = found;
    }
    if (n_sort > 1) {
      size_t tmp = 0;
      PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, first.as<uint64_t>(), first_sorted.as<uint64_t>(),
                                                   idx.as<uint32_t>(), idx_sorted.as<uint32_t>(), static_cast<int>(n_sort),
                                                   0, end_bit, st));
      DeviceBuffer t(c, tmp);
      PRISM_HIP(hipcub::DeviceRadixSort::SortPairs(t.ptr, tmp, first.as<uint64_t>(), first_sorted.as<uint64_t>(),
                                                   idx.as<uint32_t>(), idx_sorted.as<uint32_t>(), static_cast<int>(n_sort),
                                                   0, end_bit, st));
    } else if (n_sort == 1) {
      PRISM_HIP(hipMemcpyAsync(idx_sorted.ptr, idx.ptr, 4, hipMemcpyDeviceToDevice, st));
    }
    const uint64_t nr = n_sort ? n_sort : 1;
    J->result.reserve(c, nr * (4 + 8 * 4) + 64);
    char* r = J->result.as<char>();
    int32_t* attr = reinterpret_cast<int32_t*>(r);
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(r + ((nr * 4 + 7) & ~7ULL));
    if (n_sort)
      hipLaunchKernelGGL(k_ja_gather, grid_for(c, n_sort), dim3(kBuildBlock), 0, st, GV, idx_sorted.as<uint32_t>(),
                         n_sort, one_wait ? found_p : nullptr, attr, cnt, cnt + nr,
                         reinterpret_cast<long long*>(cnt + 2 * nr), cnt + 3 * nr);
    PRISM_CHECK_LAUNCH();
    if (one_wait) PRISM_HIP(hipMemcpyAsync(h, J->counters.ptr, 16, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    if (one_wait) {
      found = h[0];
      if (static_cast<uint32_t>(h[1]) & 1u) fail(PRISM_ERR_INTERNAL, "Internal error: join aggregate group table full");
    }
    J->result_rows = nr;
    J->n_result = found;
    J->finished = true;
    *n_groups = found;
  });
}

prism_status prism_join_agg_read(prism_join_agg* J, int32_t* attr, uint64_t* count, uint64_t* sum_lo, int64_t* sum_hi,
                                 uint64_t* first_pair) {
  return guarded([&] {
    if (!J) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL join aggregate");
    if (!J->finished) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: call prism_join_agg_finish first");
    use_device(J->ctx);
    const uint64_t n = J->n_result;
    if (!n) return;
    const uint64_t nr = J->result_rows;
    hipStream_t st = J->ctx->stream;
    const char* r = J->result.as<char>();
    const unsigned long long* cnt = reinterpret_cast<const unsigned long long*>(r + ((nr * 4 + 7) & ~7ULL));
    if (attr) PRISM_HIP(hipMemcpyAsync(attr, r, n * 4, hipMemcpyDeviceToHost, st));
    if (count) PRISM_HIP(hipMemcpyAsync(count, cnt, n * 8, hipMemcpyDeviceToHost, st));
    if (sum_lo) PRISM_HIP(hipMemcpyAsync(sum_lo, cnt + nr, n * 8, hipMemcpyDeviceToHost, st));
    if (sum_hi) PRISM_HIP(hipMemcpyAsync(sum_hi, cnt + 2 * nr, n * 8, hipMemcpyDeviceToHost, st));
    if (first_pair) PRISM_HIP(hipMemcpyAsync(first_pair, cnt + 3 * nr, n * 8, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
  });
}

prism_status prism_join_agg_reset(prism_join_agg* J) {
  return guarded([&] {
    if (!J) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL join aggregate");
    use_device(J->ctx);
    std::lock_guard<std::mutex> lock(J->ctx->mu);
    ja_clear(J);
  });
}

void prism_join_agg_destroy(prism_join_agg* J) {
  if (!J) return;
  (void)hipSetDevice(J->ctx ? J->ctx->device : 0);
  if (J->ctx) (void)hipStreamSynchronize(J->ctx->stream);
  if (J->table) prism_join_table_free(J->table);
  delete J;
}

}  // extern "C"
