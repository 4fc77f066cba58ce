// Group-by device pieces shared by the prebuilt kernels and the kernels
// generated at run time (hipRTC): plan layout, LDS / HBM group tables,
// accumulator lanes, the LDS set-up and the merge into HBM.
#pragma once

#include "dev_program.hpp"

namespace prism {
namespace gb {

constexpr uint64_t kEmptyKey = ~0ULL;
constexpr uint64_t kNoRow = ~0ULL;
constexpr uint64_t kMix = 0x9E3779B97F4A7C15ULL;
constexpr int kW_KEY = 0, kW_FIRST = 1, kW_COUNT = 2, kW_LANES = 3;
constexpr int kBlock = 256;
// Rows per lane per iteration: 4 for programs of stack depth <= 4, else 2.
constexpr int kLdsProbes = 8;
constexpr size_t kLdsBudget = 64 * 1024;
constexpr size_t kPrivBudget = 56 * 1024;
constexpr uint64_t kSampleRows = 1 << 20;

enum LaneOp : int32_t { L_SUM_I = 1, L_SUM_F = 2, L_MIN_F = 3, L_MAX_F = 4, L_CNT_VALID = 5 };

struct Lane {
  int32_t op;
  int32_t value;
  int32_t word;
  int32_t is_int;  // value program yields int64 (converted for MIN/MAX)
};

struct DevPlan {
  int32_t n_keys;
  int32_t key_cols[kMaxKeys];
  int32_t key_pref[kMaxKeys];  // prefetch slot of a fixed-width key, -1 for Utf8
  int32_t n_pref;              // fixed-width columns loaded once per row batch
  int32_t pref_col[kMaxPref];
  int32_t filter_start, filter_len;
  int32_t n_values;
  int32_t value_start[kMaxValues];
  int32_t value_len[kMaxValues];
  int32_t n_lanes;
  Lane lanes[kMaxLanes];
  int32_t stride;
  int32_t skip_nulls;
  uint64_t tmpl[kW_LANES + 3 * kMaxLanes];
  prism_instr code[kMaxCode];
  StrTable strs;
};

// Flags written by the kernels.
constexpr uint32_t kFlagDropped = 1;   // a group could not be placed: redo with a larger table
constexpr uint32_t kFlagOverflow = 2;  // exact int64 arithmetic overflowed
constexpr uint32_t kFlagGrow = 4;      // global table above 3/4 load

struct GTable {
  unsigned long long* w;
  uint64_t cap;  // power of two; slot `cap` is reserved for the hash value kEmptyKey
  int shift;
  int stride;
  unsigned long long* n_groups;
};

__device__ __forceinline__ int lds_find(unsigned long long* lds, int stride, int slots, int shift, uint64_t h) {
  uint32_t s = static_cast<uint32_t>((h * kMix) >> shift);
  for (int p = 0; p < kLdsProbes; ++p) {
    unsigned long long* kp = &lds[static_cast<uint32_t>(s) * stride];
    unsigned long long cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == h) return static_cast<int>(s);
    if (cur == kEmptyKey) {
      unsigned long long old = atomicCAS(kp, kEmptyKey, h);
      if (old == kEmptyKey || old == h) return static_cast<int>(s);
    }
    s = (s + 1) & (slots - 1);
  }
  return -1;
}

__device__ __forceinline__ int64_t g_find(const GTable& G, uint64_t h, uint32_t* flags) {
  if (h == kEmptyKey) return static_cast<int64_t>(G.cap);
  uint64_t s = (h * kMix) >> G.shift;
  for (uint64_t p = 0; p < G.cap; ++p) {
    unsigned long long* kp = &G.w[s * G.stride];
    unsigned long long cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == h) return static_cast<int64_t>(s);
    if (cur == kEmptyKey) {
      unsigned long long old = atomicCAS(kp, kEmptyKey, h);
      if (old == kEmptyKey) {
        unsigned long long g = atomicAdd(G.n_groups, 1ULL) + 1;
        if (g * 4 > G.cap * 3) atomicOr(flags, kFlagGrow);
        return static_cast<int64_t>(s);
      }
      if (old == h) return static_cast<int64_t>(s);
    }
    s = (s + 1) & (G.cap - 1);
  }
  atomicOr(flags, kFlagDropped);
  return -1;
}

__device__ __forceinline__ bool integral_small(double d) {
  return fabs(d) < 4611686018427387904.0 && d == trunc(d);  // |d| < 2^62
}

__device__ __forceinline__ void lane_lds(unsigned long long* L, const Lane& ln, uint64_t v) {
  switch (ln.op) {
    case L_SUM_I: {
      const int64_t x = static_cast<int64_t>(v);
      const uint32_t lo = static_cast<uint32_t>(x);
      const int64_t hi = x >> 32;
      if (lo) atomicAdd(&L[ln.word], static_cast<unsigned long long>(lo));
      if (hi) atomicAdd(&L[ln.word + 1], static_cast<unsigned long long>(hi));
      break;
    }
    case L_SUM_F: {
      double d = as_f(v);
      if (integral_small(d)) {
        const int64_t x = static_cast<int64_t>(d);
        const uint32_t lo = static_cast<uint32_t>(x);
        const int64_t hi = x >> 32;
        if (lo) atomicAdd(&L[ln.word], static_cast<unsigned long long>(lo));
        if (hi) atomicAdd(&L[ln.word + 1], static_cast<unsigned long long>(hi));
      } else {
        atomicAdd(reinterpret_cast<double*>(&L[ln.word + 2]), d);
      }
      break;
    }
    case L_MIN_F:
    case L_MAX_F: {
      double d = ln.is_int ? static_cast<double>(static_cast<int64_t>(v)) : as_f(v);
      if (!isnan(d)) {
        if (ln.op == L_MIN_F) atomicMin(&L[ln.word], static_cast<unsigned long long>(f64_order(d)));
        else atomicMax(&L[ln.word], static_cast<unsigned long long>(f64_order(d)));
      }
      break;
    }
    default:  // L_CNT_VALID
      atomicAdd(&L[ln.word], 1ULL);
  }
}

__device__ __forceinline__ void lane_global(unsigned long long* G, const Lane& ln, uint64_t v) {
  switch (ln.op) {
    case L_SUM_I: {
      int64_t x = static_cast<int64_t>(v);
      atomic_add_i128(&G[ln.word], &G[ln.word + 1], static_cast<uint64_t>(x), x >> 63);
      break;
    }
    case L_SUM_F: {
      double d = as_f(v);
      if (integral_small(d)) {
        int64_t x = static_cast<int64_t>(d);
        atomic_add_i128(&G[ln.word], &G[ln.word + 1], static_cast<uint64_t>(x), x >> 63);
      } else {
        unsafeAtomicAdd(reinterpret_cast<double*>(&G[ln.word + 2]), d);
      }
      break;
    }
    case L_MIN_F:
    case L_MAX_F: {
      double d = ln.is_int ? static_cast<double>(static_cast<int64_t>(v)) : as_f(v);
      if (!isnan(d)) {
        if (ln.op == L_MIN_F) atomicMin(&G[ln.word], static_cast<unsigned long long>(f64_order(d)));
        else atomicMax(&G[ln.word], static_cast<unsigned long long>(f64_order(d)));
      }
      break;
    }
    default:
      atomicAdd(&G[ln.word], 1ULL);
  }
}

// Thread-private accumulators: word w of this thread's slot lives at
// base[(w - kW_COUNT) * kBlock].  Only the owning thread touches it.
__device__ __forceinline__ void add_i128_priv(unsigned long long* lo_p, unsigned long long* hi_p, int64_t x) {
  const unsigned long long lo = *lo_p, nlo = lo + static_cast<unsigned long long>(x);
  *lo_p = nlo;
  *hi_p = *hi_p + static_cast<unsigned long long>(x >> 63) + (nlo < lo ? 1ULL : 0ULL);
}

__device__ __forceinline__ void lane_priv(unsigned long long* base, const Lane& ln, uint64_t v) {
  unsigned long long* w = base + (ln.word - kW_COUNT) * kBlock;
  switch (ln.op) {
    case L_SUM_I:
      add_i128_priv(w, w + kBlock, static_cast<int64_t>(v));
      break;
    case L_SUM_F: {
      const double d = as_f(v);
      if (integral_small(d)) add_i128_priv(w, w + kBlock, static_cast<int64_t>(d));
      else w[2 * kBlock] = f_bits(as_f(w[2 * kBlock]) + d);
      break;
    }
    case L_MIN_F:
    case L_MAX_F: {
      const double d = ln.is_int ? static_cast<double>(static_cast<int64_t>(v)) : as_f(v);
      if (!isnan(d)) {
        const unsigned long long u = f64_order(d);
        if (ln.op == L_MIN_F ? u < *w : u > *w) *w = u;
      }
      break;
    }
    default:
      *w += 1;
  }
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ void wave_sum_i128(uint64_t& lo, int64_t& hi) {
  __int128 v = static_cast<__int128>((static_cast<unsigned __int128>(static_cast<uint64_t>(hi)) << 64) | lo);
  for (int d = 32; d > 0; d >>= 1) {
    const uint64_t a = static_cast<uint64_t>(v), b = static_cast<uint64_t>(v >> 64);
    const uint64_t oa = __shfl_xor(a, d, 64), ob = __shfl_xor(b, d, 64);
    v += static_cast<__int128>((static_cast<unsigned __int128>(ob) << 64) | oa);
  }
  lo = static_cast<uint64_t>(v);
  hi = static_cast<int64_t>(v >> 64);
}

// Map of the private form: group hash -> slot, linear probing over all slots.
__device__ __forceinline__ int pmap_find(unsigned long long* keys, int slots, int shift, uint64_t h) {
  uint32_t s = static_cast<uint32_t>((h * kMix) >> shift);
  for (int p = 0; p < slots; ++p) {
    const unsigned long long cur = __hip_atomic_load(&keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == h) return static_cast<int>(s);
    if (cur == kEmptyKey) {
      const unsigned long long old = atomicCAS(&keys[s], kEmptyKey, h);
      if (old == kEmptyKey || old == h) return static_cast<int>(s);
    }
    s = (s + 1) & (slots - 1);
  }
  return -1;
}

// Merge one LDS slot (split int128 sums) into a global slot (true int128).
__device__ __forceinline__ void merge_slot(unsigned long long* G, const unsigned long long* L, const DevPlan* P) {
  if (L[kW_COUNT]) atomicAdd(&G[kW_COUNT], L[kW_COUNT]);
  atomicMin(&G[kW_FIRST], L[kW_FIRST]);
  for (int l = 0; l < P->n_lanes; ++l) {
    const Lane ln = P->lanes[l];
    const int w = ln.word;
    switch (ln.op) {
      case L_SUM_I:
      case L_SUM_F: {
        Split128 sp;
        sp.lo32 = L[w];
        sp.hi32 = static_cast<int64_t>(L[w + 1]);
        uint64_t lo;
        int64_t hi;
        sp.to_i128(lo, hi);
        if (lo | static_cast<uint64_t>(hi)) atomic_add_i128(&G[w], &G[w + 1], lo, hi);
        if (ln.op == L_SUM_F && L[w + 2]) unsafeAtomicAdd(reinterpret_cast<double*>(&G[w + 2]), as_f(L[w + 2]));
        break;
      }
      case L_MIN_F:
        if (L[w] != P->tmpl[w]) atomicMin(&G[w], L[w]);
        break;
      case L_MAX_F:
        if (L[w] != P->tmpl[w]) atomicMax(&G[w], L[w]);
        break;
      default:
        if (L[w]) atomicAdd(&G[w], L[w]);
    }
  }
}

// LDS partial groups of one workgroup: private form = [map keys][map first
// rows][accumulators: slot, word, thread]; shared form = lds_slots slots.
template <bool kPriv>
__device__ __forceinline__ void gb_init(const DevPlan* __restrict__ P, unsigned long long* lds, int lds_slots) {
  const int stride = P->stride;
  const int tid = threadIdx.x;
  const int PW = stride - kW_COUNT;
  unsigned long long* mapk = lds;
  unsigned long long* mapf = lds + lds_slots;
  unsigned long long* priv = lds + 2 * lds_slots;
  if (kPriv) {
    for (int i = tid; i < lds_slots; i += kBlock) {
      mapk[i] = kEmptyKey;
      mapf[i] = kNoRow;
    }
    for (int i = 0; i < lds_slots * PW; ++i) priv[i * kBlock + tid] = P->tmpl[kW_COUNT + i % PW];
  } else {
    for (int i = tid; i < lds_slots * stride; i += kBlock) lds[i] = P->tmpl[i % stride];
  }
  __syncthreads();
}

// Merge the workgroup's LDS partial groups into the global table.
template <bool kPriv>
__device__ __forceinline__ void gb_flush(const DevPlan* __restrict__ P, const GTable& G, unsigned long long* lds,
                                         int lds_slots, uint32_t* flags) {
  const int stride = P->stride;
  const int tid = threadIdx.x;
  const int PW = stride - kW_COUNT;
  unsigned long long* mapk = lds;
  unsigned long long* mapf = lds + lds_slots;
  unsigned long long* priv = lds + 2 * lds_slots;
  __syncthreads();
  if (!kPriv) {
    for (int s = tid; s < lds_slots; s += kBlock) {
      const unsigned long long* L = &lds[static_cast<uint32_t>(s) * stride];
      if (L[kW_KEY] == kEmptyKey) continue;
      int64_t gs = g_find(G, L[kW_KEY], flags);
      if (gs < 0) continue;
      merge_slot(&G.w[static_cast<uint64_t>(gs) * stride], L, P);
    }
    return;
  }
  // Private form: one wave per slot reduces the threads' accumulators and
  // merges the totals into HBM.
  const int lane = tid & 63, wave = tid >> 6;
  for (int s = wave; s < lds_slots; s += kBlock / 64) {
    const unsigned long long key = mapk[s];
    if (key == kEmptyKey) continue;
    const unsigned long long* A = &priv[(s * PW) * kBlock];
    unsigned long long cnt = 0;
    for (int t = lane; t < kBlock; t += 64) cnt += A[t];
    cnt = wave_sum_u64(cnt);
    if (cnt == 0) continue;
    long long gsl = -1;
    if (lane == 0) gsl = g_find(G, key, flags);
    gsl = __shfl(gsl, 0, 64);
    if (gsl < 0) continue;
    unsigned long long* gq = &G.w[static_cast<uint64_t>(gsl) * stride];
    if (lane == 0) {
      atomicAdd(&gq[kW_COUNT], cnt);
      atomicMin(&gq[kW_FIRST], mapf[s]);
    }
    for (int l = 0; l < P->n_lanes; ++l) {
      const Lane ln = P->lanes[l];
      const int w = ln.word;
      const unsigned long long* W = A + (w - kW_COUNT) * kBlock;
      switch (ln.op) {
        case L_SUM_I:
        case L_SUM_F: {
          __int128 acc = 0;
          for (int t = lane; t < kBlock; t += 64)
            acc += static_cast<__int128>((static_cast<unsigned __int128>(W[kBlock + t]) << 64) | W[t]);
          uint64_t lo = static_cast<uint64_t>(acc);
          int64_t hi = static_cast<int64_t>(acc >> 64);
          wave_sum_i128(lo, hi);
          if (lane == 0 && (lo | static_cast<uint64_t>(hi))) atomic_add_i128(&gq[w], &gq[w + 1], lo, hi);
          if (ln.op == L_SUM_F) {
            double f = 0;
            for (int t = lane; t < kBlock; t += 64) f += as_f(W[2 * kBlock + t]);
            for (int d = 32; d > 0; d >>= 1) f += __shfl_xor(f, d, 64);
            if (lane == 0 && f != 0) unsafeAtomicAdd(reinterpret_cast<double*>(&gq[w + 2]), f);
          }
          break;
        }
        case L_MIN_F:
        case L_MAX_F: {
          const bool mn = ln.op == L_MIN_F;
          unsigned long long m = P->tmpl[w];
          for (int t = lane; t < kBlock; t += 64) m = mn ? (W[t] < m ? W[t] : m) : (W[t] > m ? W[t] : m);
          for (int d = 32; d > 0; d >>= 1) {
            const unsigned long long o = __shfl_xor(m, d, 64);
            m = mn ? (o < m ? o : m) : (o > m ? o : m);
          }
          if (lane == 0 && m != P->tmpl[w]) {
            if (mn) atomicMin(&gq[w], m);
            else atomicMax(&gq[w], m);
          }
          break;
        }
        default: {
          unsigned long long c = 0;
          for (int t = lane; t < kBlock; t += 64) c += W[t];
          c = wave_sum_u64(c);
          if (lane == 0 && c) atomicAdd(&gq[w], c);
        }
      }
    }
  }
}


// ---------------------------------------------------------------------------
// Wave-wide reductions for wave-level pre-aggregation: row_shr DPP moves build
// an inclusive scan inside each 16-lane row (a lane whose source is outside
// its row keeps the identity), then the four row totals (lanes 15, 31, 47,
// 63) combine in scalar registers.  Every lane of the wave must be active.
// ---------------------------------------------------------------------------
template <int kCtrl>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x, uint64_t identity) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(static_cast<uint32_t>(identity)), static_cast<int>(static_cast<uint32_t>(x)), kCtrl, 0xf, 0xf,
      false));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(static_cast<uint32_t>(identity >> 32)), static_cast<int>(static_cast<uint32_t>(x >> 32)), kCtrl,
      0xf, 0xf, false));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, int lane) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x)), lane));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x >> 32)), lane));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;

__device__ __forceinline__ uint64_t wave_add_u64(uint64_t x) {
  x += dpp_u64<kDppRowShr1>(x, 0);
  x += dpp_u64<kDppRowShr2>(x, 0);
  x += dpp_u64<kDppRowShr4>(x, 0);
  x += dpp_u64<kDppRowShr8>(x, 0);
  return readlane_u64(x, 15) + readlane_u64(x, 31) + readlane_u64(x, 47) + readlane_u64(x, 63);
}
__device__ __forceinline__ double wave_add_f64(double d) {
  uint64_t x = f_bits(d);
  x = f_bits(as_f(x) + as_f(dpp_u64<kDppRowShr1>(x, 0)));
  x = f_bits(as_f(x) + as_f(dpp_u64<kDppRowShr2>(x, 0)));
  x = f_bits(as_f(x) + as_f(dpp_u64<kDppRowShr4>(x, 0)));
  x = f_bits(as_f(x) + as_f(dpp_u64<kDppRowShr8>(x, 0)));
  return (as_f(readlane_u64(x, 15)) + as_f(readlane_u64(x, 31))) + (as_f(readlane_u64(x, 47)) + as_f(readlane_u64(x, 63)));
}
__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x) {
  x = umin64(x, dpp_u64<kDppRowShr1>(x, ~0ULL));
  x = umin64(x, dpp_u64<kDppRowShr2>(x, ~0ULL));
  x = umin64(x, dpp_u64<kDppRowShr4>(x, ~0ULL));
  x = umin64(x, dpp_u64<kDppRowShr8>(x, ~0ULL));
  return umin64(umin64(readlane_u64(x, 15), readlane_u64(x, 31)), umin64(readlane_u64(x, 47), readlane_u64(x, 63)));
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
  x = umax64(x, dpp_u64<kDppRowShr1>(x, 0));
  x = umax64(x, dpp_u64<kDppRowShr2>(x, 0));
  x = umax64(x, dpp_u64<kDppRowShr4>(x, 0));
  x = umax64(x, dpp_u64<kDppRowShr8>(x, 0));
  return umax64(umax64(readlane_u64(x, 15), readlane_u64(x, 31)), umax64(readlane_u64(x, 47), readlane_u64(x, 63)));
}


// ---------------------------------------------------------------------------
// Persistent tile-queue form of the update (the operator service): a
// workgroup takes (request, row range) tiles from a queue until it is empty.
// ---------------------------------------------------------------------------
constexpr int kQueueLanes = 8;  // claim counters per queue: one per XCD

struct TileQueue {
  uint32_t n_tiles;
  uint32_t offset;  // first tile of this queue in the tile array
  uint32_t pad[14];
  uint32_t next[kQueueLanes * 16];  // counter of XCD x at next[16 * x], 64 bytes apart
};

struct RowTile {
  uint32_t req;
  uint32_t pad;
  uint64_t start, end;
};

// One group-by request of a batch.  A batch's requests sit kTileReqStride
// bytes apart (each request's record also holds its other descriptors and
// column descriptors), the group-by descriptor first.
struct TileReq {
  const DevPlan* P;
  const prism_column* cols;
  uint64_t n;
  GTable G;
  int32_t lds_slots, lds_shift;
  unsigned long long* out;  // n groups, flags, then rows (stride words each)
  uint32_t* flags;
  uint32_t pad;
};
constexpr uint32_t kTileReqStride = 1024;

__device__ __forceinline__ const TileReq& tile_req(const TileReq* base, uint32_t i) {
  return *reinterpret_cast<const TileReq*>(reinterpret_cast<const char*>(base) +
                                           static_cast<uint64_t>(i) * kTileReqStride);
}

// Next tile index of the workgroup.  Claims on one device-wide counter
// serialise (about 0.1 us each), so the first tile of a workgroup is its own
// index, and later ones come from the counter of its XCD (workgroup index mod
// kQueueLanes), which hands out that XCD's share of the remaining tiles:
// t = grid + x + kQueueLanes * k.  Every tile >= grid is handed out once as
// long as the grid holds every residue, i.e. grid >= kQueueLanes whenever
// tiles remain past the first round (grids are then a multiple of 8).
// `first` starts true; the result is uniform over the workgroup.
__device__ __forceinline__ uint32_t next_tile(TileQueue* q, uint32_t* s_t, bool& first) {
  if (first) {
    first = false;
    return blockIdx.x;
  }
  const uint32_t x = blockIdx.x % kQueueLanes;
  if (threadIdx.x == 0) *s_t = atomicAdd(&q->next[16 * x], 1u);
  __syncthreads();
  const uint32_t k = *s_t;
  __syncthreads();
  return gridDim.x + x + kQueueLanes * k;
}


// ---------------------------------------------------------------------------
// Small finish (one workgroup of kFinishThreads): the occupied slots of a
// group table with at most kFinishCap slots, ordered by first row (bitonic
// sort in LDS) and written stride words each to `rows`; returns the group
// count.  s_first / s_slot are LDS arrays of kFinishCap entries.
// ---------------------------------------------------------------------------
constexpr int kFinishCap = 4096;
constexpr int kFinishThreads = 256;

__device__ __forceinline__ uint32_t order_groups(const GTable& G, unsigned long long* rows, unsigned long long* s_first,
                                                 uint32_t* s_slot, uint32_t* s_n) {
  const int stride = G.stride;
  const uint64_t slots = G.cap + 1;
  if (threadIdx.x == 0) *s_n = 0;
  __syncthreads();
  // 4 slots' first rows in flight per thread (their loads do not wait for
  // each other's LDS appends)
  for (uint64_t s0 = threadIdx.x; s0 < slots; s0 += 4 * kFinishThreads) {
    unsigned long long f[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t s = s0 + static_cast<uint64_t>(u) * kFinishThreads;
      f[u] = s < slots ? G.w[s * stride + kW_FIRST] : kNoRow;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (f[u] == kNoRow) continue;
      const uint32_t i = atomicAdd(s_n, 1u);
      s_first[i] = f[u];
      s_slot[i] = static_cast<uint32_t>(s0 + static_cast<uint64_t>(u) * kFinishThreads);
    }
  }
  __syncthreads();
  const uint32_t n = *s_n;
  uint32_t p2 = 1;
  while (p2 < n) p2 <<= 1;
  for (uint32_t i = n + threadIdx.x; i < p2; i += kFinishThreads) {
    s_first[i] = ~0ULL;
    s_slot[i] = 0xFFFFFFFFu;
  }
  __syncthreads();
  for (uint32_t size = 2; size <= p2; size <<= 1)
    for (uint32_t half = size >> 1; half > 0; half >>= 1) {
      for (uint32_t i = threadIdx.x; i < p2; i += kFinishThreads) {
        const uint32_t j = i ^ half;
        if (j > i) {
          const bool up = (i & size) == 0;
          const unsigned long long a = s_first[i], b = s_first[j];
          if ((a > b) == up) {
            s_first[i] = b;
            s_first[j] = a;
            const uint32_t t = s_slot[i];
            s_slot[i] = s_slot[j];
            s_slot[j] = t;
          }
        }
      }
      __syncthreads();
    }
  // the ordered slots' words, 8 in flight per thread
  const uint64_t total = static_cast<uint64_t>(n) * stride;
  for (uint64_t e0 = threadIdx.x; e0 < total; e0 += 8 * kFinishThreads) {
    unsigned long long v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint64_t e = e0 + static_cast<uint64_t>(u) * kFinishThreads;
      if (e < total) {
        const uint64_t i = e / stride, j = e % stride;
        v[u] = G.w[static_cast<uint64_t>(s_slot[i]) * stride + j];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint64_t e = e0 + static_cast<uint64_t>(u) * kFinishThreads;
      if (e < total) rows[e] = v[u];
    }
  }
  return n;
}

}  // namespace gb
}  // namespace prism
