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
#include <map>
#include <memory>
#include <vector>

#include "dev_groupby.hpp"
#include "program.hpp"

namespace prism {
namespace gb {

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
      // Rows the filter drops take no further part (no overflow checks on them).
      uint32_t keep = 0;
#pragma unroll
      for (int r = 0; r < kRows; ++r) keep |= in[r] ? (1u << r) : 0u;
#pragma unroll
      for (int j = 0; j < kPref; ++j) pre_v &= ~(((1u << kRows) - 1) << (j * kRows)) | (keep << (j * kRows));
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
// LDS table of the shared form for `known` groups (0 = unknown): slots, hash
// shift and bytes; no LDS table (0 slots) when the groups cannot fit.
size_t shared_slots(const prism_groupby* g, uint64_t known, int& lds_slots, int& lds_shift);
}  // namespace gb
}  // namespace prism
