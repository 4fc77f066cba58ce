#pragma once
// Hash aggregation on gfx950 (native/prism-executor/src/hash_aggregate.rs:182-447),
// with the filter and projection stage fused into the load loop.
//
// Group identity is the reference's 64-bit FxHash of the key values
// (hash_aggregate.rs:483-504), without a key-equality check, exactly as the
// reference groups (hash_aggregate.rs:215,248-255).  Output order is first
// appearance (minimum global row per group), as hash_aggregate.rs:302-314
// emits it.
//
// Every fixed-width column the plan reads is loaded once per row batch, all
// loads in flight together, before the row programs run on registers.  Two
// workgroup-local forms then hold partial groups in LDS:
//   private - few groups: a small shared map (group hash -> slot) and, per
//             thread, its own accumulators for every slot ([slot][word][thread]
//             so a wave touches consecutive words); no atomics on hot groups.
//   shared  - many groups: one LDS table of slots updated with LDS atomics.
// Rows whose group does not fit go straight to the global table in HBM.  At
// the end of the launch the LDS partials merge into the global table.
//
// Accumulators:
//   SUM of integers   - exact int128.  LDS holds split sums (low 32-bit halves
//                       unsigned, high halves signed); HBM holds (lo, hi) words
//                       updated with carry-propagating 64-bit atomics.
//   SUM of f64        - integral values below 2^62 take the exact int128 path,
//                       the rest an f64 atomic sum; the two parts add at the end.
//   MIN / MAX         - f64 order-preserving u64 encoding, NaN never selected
//                       (GroupAccumulator::accumulate, hash_aggregate.rs:135-140).
//   COUNT             - row count per group (nulls included, hash_aggregate.rs:257-269);
//                       the global (no GROUP BY) form counts non-null values
//                       (hash_aggregate.rs:364-377).
#include <memory>
#include <vector>

#include "program.hpp"

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

// Rows of one workgroup: batches of kRows * kBlock rows starting at `begin`,
// `step` apart, below `n`; groups accumulate in LDS (see gb_init) or, when
// they do not fit there, straight into the global table.
template <int kRows, int kDepth, int kPref, bool kPriv>
__device__ __forceinline__ void gb_rows(const DevPlan* __restrict__ P, const prism_column* __restrict__ cols,
                                        uint64_t begin, uint64_t step, uint64_t n, uint64_t row_base,
                                        const GTable& G, unsigned long long* lds, int lds_slots, int lds_shift,
                                        uint32_t* flags) {
  const int stride = P->stride;
  const int tid = threadIdx.x;
  const int PW = stride - kW_COUNT;
  unsigned long long* mapk = lds;
  unsigned long long* mapf = lds + lds_slots;
  unsigned long long* priv = lds + 2 * lds_slots;
  const int n_pref = P->n_pref;
  for (uint64_t it = begin; it < n; it += step) {
    uint64_t row[kRows];
    bool in[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      row[r] = it + static_cast<uint64_t>(r) * kBlock + tid;
      in[r] = row[r] < n;
    }
    // One round trip for every fixed-width column of the plan.
    uint64_t pre[kPref][kRows];
    uint32_t pre_v = 0;
#pragma unroll
    for (int j = 0; j < kPref; ++j) {
      if (j < n_pref) {
        const prism_column& c = cols[P->pref_col[j]];
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          pre[j][r] = in[r] ? col_load_raw(c, row[r]) : 0;
          pre_v |= (in[r] && col_valid(c, row[r])) ? (1u << (j * kRows + r)) : 0u;
        }
      } else {
#pragma unroll
        for (int r = 0; r < kRows; ++r) pre[j][r] = 0;
      }
    }
    if (P->filter_len) {
      uint64_t fv[kRows];
      uint32_t fvalid;
      run_program<kRows, kDepth, kPref>(P->code + P->filter_start, P->filter_len, cols, &P->strs, row, in, pre, pre_v,
                                        fv, fvalid, flags);
#pragma unroll
      for (int r = 0; r < kRows; ++r) in[r] = in[r] && ((fvalid >> r) & 1) && (fv[r] & 1);
    }
    uint64_t h[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) h[r] = 0;
    for (int k = 0; k < P->n_keys; ++k) {
      const prism_column& c = cols[P->key_cols[k]];
      const int kp = P->key_pref[k];
      if (kp >= 0) {
#pragma unroll
        for (int j = 0; j < kPref; ++j)
          if (kp == j) {
#pragma unroll
            for (int r = 0; r < kRows; ++r)
              h[r] = fx_step(h[r], ((pre_v >> (j * kRows + r)) & 1) ? key_word(c.type, pre[j][r]) : 0);
          }
      } else {
#pragma unroll
        for (int r = 0; r < kRows; ++r)
          if (in[r]) h[r] = hash_key(c, row[r], h[r]);
      }
    }
    // Slot per row: LDS (private or shared form) if it fits, else HBM.
    int32_t lw[kRows];              // private: slot; shared: LDS word index; -1 = HBM
    unsigned long long* gp[kRows];  // HBM slot
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      lw[r] = -1;
      gp[r] = nullptr;
      if (!in[r]) continue;
      const unsigned long long fr = row_base + row[r];
      if (kPriv) {
        const int ps = (lds_slots && h[r] != kEmptyKey) ? pmap_find(mapk, lds_slots, lds_shift, h[r]) : -1;
        if (ps >= 0) {
          lw[r] = ps;
          unsigned long long* cnt = &priv[(ps * PW) * kBlock + tid];
          // this thread's rows only grow, so its first row of a slot is its minimum
          if (*cnt == 0) atomicMin(&mapf[ps], fr);
          *cnt += 1;
          continue;
        }
      } else {
        const int ls = (lds_slots && h[r] != kEmptyKey) ? lds_find(lds, stride, lds_slots, lds_shift, h[r]) : -1;
        if (ls >= 0) {
          lw[r] = ls * stride;
          atomicAdd(&lds[lw[r] + kW_COUNT], 1ULL);
          if (fr < __hip_atomic_load(&lds[lw[r] + kW_FIRST], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
            atomicMin(&lds[lw[r] + kW_FIRST], fr);
          continue;
        }
      }
      const int64_t gs = g_find(G, h[r], flags);
      if (gs < 0) {
        in[r] = false;
        continue;
      }
      gp[r] = &G.w[static_cast<uint64_t>(gs) * stride];
      atomicAdd(&gp[r][kW_COUNT], 1ULL);
      if (fr < __hip_atomic_load(&gp[r][kW_FIRST], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(&gp[r][kW_FIRST], fr);
    }
    for (int v = 0; v < P->n_values; ++v) {
      uint64_t val[kRows];
      uint32_t vvalid;
      run_program<kRows, kDepth, kPref>(P->code + P->value_start[v], P->value_len[v], cols, &P->strs, row, in, pre,
                                        pre_v, val, vvalid, flags);
      for (int l = 0; l < P->n_lanes; ++l) {
        const Lane ln = P->lanes[l];
        if (ln.value != v) continue;
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          if (!in[r]) continue;
          if (P->skip_nulls && !((vvalid >> r) & 1)) continue;
          if (lw[r] >= 0) {
            if (kPriv) lane_priv(&priv[(lw[r] * PW) * kBlock + tid], ln, val[r]);
            else lane_lds(&lds[lw[r]], ln, val[r]);
          } else {
            lane_global(gp[r], ln, val[r]);
          }
        }
      }
    }
  }
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


__global__ void k_fill_tmpl(unsigned long long* w, uint64_t n_words, const DevPlan* __restrict__ P);

}  // namespace gb
}  // namespace prism

struct prism_groupby {
  prism::Ctx* ctx = nullptr;
  prism::gb::DevPlan plan{};
  std::vector<int32_t> col_types;
  std::vector<prism_groupby_agg> aggs;
  std::vector<int32_t> value_types;     // ValType per value program
  std::vector<int32_t> agg_lane_sum;    // lane index of the SUM lanes per agg (-1)
  std::vector<int32_t> agg_lane_mm;     // MIN/MAX lane per agg (-1)
  std::vector<int32_t> agg_lane_cnt;    // CNT_VALID lane per agg (-1)
  prism::DeviceBuffer d_plan;
  prism::DeviceBuffer table, backup;
  uint64_t cap = 0;
  uint64_t known_groups = 0;
  uint64_t seen_groups = 0;  // largest group count of any earlier update, kept across resets
  prism::DeviceBuffer counters;  // n_groups, flags, n_found
  prism::DeviceBuffer d_cols;    // column descriptors of the current update
  prism::DeviceBuffer result;    // finished groups, stride words each
  uint64_t n_result = 0;
  uint64_t expected = 0;
  int max_depth = 1;
  bool finished = false;

  unsigned long long* n_groups_ptr() { return counters.as<unsigned long long>(); }
  uint32_t* flags_ptr() { return reinterpret_cast<uint32_t*>(counters.as<char>() + 8); }
  unsigned long long* n_found_ptr() { return reinterpret_cast<unsigned long long*>(counters.as<char>() + 16); }
  size_t slot_bytes() const { return static_cast<size_t>(plan.stride) * 8; }
  prism::gb::GTable view() {
    prism::gb::GTable G;
    G.w = table.as<unsigned long long>();
    G.cap = cap;
    G.shift = 64 - prism::log2_pow2(cap);
    G.stride = plan.stride;
    G.n_groups = n_groups_ptr();
    return G;
  }
};

namespace prism {
namespace gb {
// Convert finished group rows (stride words each, first-appearance order, on
// the host) into the caller's output arrays (prism_groupby_read).
void convert_rows(const prism_groupby* g, const unsigned long long* rows, uint64_t n, const prism_groupby_output* out);
}  // namespace gb
}  // namespace prism
