// LIKE '%literal%' over a Utf8 column as a byte scan (string_ops.rs:16-25
// string_like / string_ilike with the contains shape, the common text filter).
//
// Scanning each row with its own lane costs every lane the longest string of
// its wave and reads 64 strings per load instruction.  Here a workgroup takes
// 2,048 rows at a time and works on their bytes as one stream:
//   1. every thread takes aligned 16-byte blocks of the rows' byte range
//      (coalesced loads, the next block for the look-ahead) and marks the byte
//      positions where the literal's first 8 bytes start (byte-wise SWAR
//      equality of funnel-shifted words, ANDed over the literal's bytes; bytes
//      past the 8th are compared from memory for the few candidates), one bit
//      per byte position in LDS;
//   2. every row tests the bits of its own start positions [begin, end - m].
// A match that starts in one row and ends in the next is outside both rows'
// ranges, so rows never see their neighbours' bytes.  Sub-tiles whose bytes
// exceed the LDS bitmap (rows averaging over 64 bytes) are scanned per row.
#pragma once

#include "dev_program.hpp"

namespace prism {

constexpr int kLikeMaxLit = 64;

// The literal between the pattern's '%'s, prepared on the host.
struct LikeArg {
  int32_t m;             // literal bytes (0: every non-null row matches)
  int32_t ci;            // ILIKE: ASCII letters match either case
  uint32_t bcast[8];     // literal byte i (lower-cased letter for ILIKE) in every byte
  uint32_t fold[8];      // 0x20 in every byte when ILIKE and byte i is a letter
  uint8_t lit[kLikeMaxLit];
};

// One LIKE / ILIKE of a Utf8 column by a '%literal%' pattern with at most
// kLikeMaxLit literal bytes: the program this scan replaces.  Returns the
// column, or -1 for any other program.
inline int like_scan_prepare(const prism_instr* code, uint32_t n_code, const StrTable& strs, LikeArg& a) {
  if (n_code != 1 || (code[0].op != PRISM_OP_LIKE && code[0].op != PRISM_OP_ILIKE)) return -1;
  const int sidx = static_cast<int>(code[0].i64 & 0xff);
  if (sidx < 0 || sidx >= strs.n || strs.like_kind[sidx] != LK_CONTAINS) return -1;
  const int32_t np = strs.off[sidx + 1] - strs.off[sidx];
  const int32_t m = np >= 2 ? np - 2 : 0;
  if (m > kLikeMaxLit) return -1;
  memset(&a, 0, sizeof(a));
  a.m = m;
  a.ci = code[0].op == PRISM_OP_ILIKE;
  const uint8_t* p = strs.bytes + strs.off[sidx] + 1;
  for (int32_t i = 0; i < m; ++i) {
    uint8_t c = p[i];
    const bool letter = (c | 0x20) >= 'a' && (c | 0x20) <= 'z';
    if (a.ci && letter) c |= 0x20;
    a.lit[i] = c;
    if (i < 8) {
      a.bcast[i] = c * 0x01010101u;
      a.fold[i] = a.ci && letter ? 0x20202020u : 0u;
    }
  }
  return code[0].arg;
}

#if defined(__HIPCC__)

constexpr int kLsThreads = 256;
constexpr int kLsWaves = kLsThreads / 64;
constexpr int kLsRows = 2048;         // rows per sub-tile (a multiple of 64 * kLsWaves)
constexpr int kLsBitWords = 4096;     // LDS bitmap: 131,072 byte positions per sub-tile
static_assert(kLsRows % (64 * kLsWaves) == 0, "whole words per wave");

// 0x80 in exactly the bytes of y that are zero.
__device__ __forceinline__ uint32_t ls_zero_bytes(uint32_t y) {
  return ~(((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y | 0x7F7F7F7Fu);
}

// Match bits of the 16 byte positions of block j (blocks are the aligned
// 16-byte units from `base`; `last` = the last block holding a byte of the
// range, repeated past it).  Bit i = the literal's first min(m, 8) bytes start
// at byte i of the block.
__device__ __forceinline__ uint32_t ls_block_bits(const PRISM_GLOBAL v4u32* __restrict__ base, int64_t j, int64_t last,
                                                  const LikeArg& A, int mh) {
  const v4u32 v = base[j];
  const v4u32 u = base[j + 1 <= last ? j + 1 : last];
  const uint32_t W[8] = {v.x, v.y, v.z, v.w, u.x, u.y, u.z, u.w};
  uint32_t bits = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    uint32_t acc = 0x80808080u;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i >= mh) break;
      // bytes i .. i + 3 from byte 0 of word t
      const uint32_t x = i == 0 ? W[t]
                         : i < 4 ? __builtin_amdgcn_alignbyte(W[t + 1], W[t], i)
                         : i == 4 ? W[t + 1]
                                  : __builtin_amdgcn_alignbyte(W[t + 2], W[t + 1], i - 4);
      acc &= ls_zero_bytes((x | A.fold[i]) ^ A.bcast[i]);
    }
    // 0x80 flags of bytes 0..3 -> bits 0..3 (the products land on distinct bits)
    bits |= ((((acc >> 7) * 0x00204081u) >> 21) & 0xFu) << (4 * t);
  }
  return bits;
}

// Rows [r0, r1) (r0 a multiple of 64, r1 - r0 <= kLsRows) of column c: the
// LIKE result of every row as bitmap words out[row / 64] (value AND valid;
// when out is given) and the validity words (when out_valid is given); counts
// null and selected rows.  Called by every thread of the workgroup.
__device__ __forceinline__ void like_scan_tile(const prism_column& c, uint64_t r0, uint64_t r1, const LikeArg& A,
                                               uint32_t* s_bits, unsigned long long* out,
                                               unsigned long long* out_valid, unsigned long long& nulls,
                                               unsigned long long& selected) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // global-typed pointers: generic ones would make every load a flat load,
  // whose waits also drain the LDS bitmap traffic
  const PRISM_GLOBAL int32_t* off = (const PRISM_GLOBAL int32_t*)(c.offsets + c.offset);
  const PRISM_GLOBAL uint8_t* bytes = (const PRISM_GLOBAL uint8_t*)(c.data);
  const PRISM_GLOBAL uint8_t* vbits = (const PRISM_GLOBAL uint8_t*)(c.validity);
  const int32_t b0 = off[r0], b1 = off[r1];
  const uintptr_t a0 = (uintptr_t)(bytes + b0);
  const int32_t lead = static_cast<int32_t>(a0 & 15);
  const PRISM_GLOBAL v4u32* base = (const PRISM_GLOBAL v4u32*)(a0 - lead);
  const int64_t span = static_cast<int64_t>(b1 - b0) + lead;  // byte positions of the sub-tile
  const int m = A.m;
  const bool in_lds = span <= 32LL * kLsBitWords;
  if (in_lds && m > 0 && b1 > b0) {
    const int64_t last = (span - 1) >> 4;
    const int mh = m < 8 ? m : 8;
    uint16_t* s16 = reinterpret_cast<uint16_t*>(s_bits);
    for (int64_t j = threadIdx.x; j <= last; j += kLsThreads) {
      uint32_t bits = ls_block_bits(base, j, last, A, mh);
      if (m > 8) {  // candidates on the first 8 bytes: the rest from memory
        uint32_t cand = bits;
        while (cand) {
          const int i = __builtin_ctz(cand);
          cand &= cand - 1;
          const int64_t p = 16 * j + i;
          bool ok = p + m <= span;
          const PRISM_GLOBAL uint8_t* s = (const PRISM_GLOBAL uint8_t*)(base) + p;
          for (int k = 8; ok && k < m; ++k) ok = (A.ci ? fold_ascii(s[k]) : s[k]) == A.lit[k];
          if (!ok) bits &= ~(1u << i);
        }
      }
      s16[j] = static_cast<uint16_t>(bits);
    }
    __syncthreads();
  }
  for (int i = 0; i < kLsRows / (64 * kLsWaves); ++i) {
    const uint64_t wr = r0 + static_cast<uint64_t>(wave + i * kLsWaves) * 64;  // first row of this word
    if (wr >= r1) break;  // uniform per wave
    const uint64_t r = wr + lane;
    const bool in = r < r1;
    bool v = false, res = false;
    if (in) {
      v = vbits == nullptr || bit_at(vbits, static_cast<uint64_t>(c.offset) + r);
      const int32_t rb = off[r], re = off[r + 1];
      if (m == 0) {
        res = true;
      } else if (re - rb >= m) {
        if (in_lds) {
          const int32_t lo = rb - b0 + lead, hi = re - m - b0 + lead;  // candidate starts, inclusive
          const int32_t wl = lo >> 5, wh = hi >> 5;
          const uint32_t mlo = ~0u << (lo & 31), mhi = ~0u >> (31 - (hi & 31));
          uint32_t x;
          if (wl == wh) {
            x = s_bits[wl] & mlo & mhi;
          } else {
            x = (s_bits[wl] & mlo) | (s_bits[wh] & mhi);
            for (int32_t w = wl + 1; w < wh && !x; ++w) x = s_bits[w];
          }
          res = x != 0;
        } else {
          res = contains_lit((const uint8_t*)(bytes + rb), re - rb, A.lit, m, A.ci != 0);
        }
      }
    }
    const unsigned long long vw = __ballot(v);
    const unsigned long long bw = __ballot(v && res);
    const unsigned long long inw = __ballot(in);
    if (lane == 0) {
      if (out) out[wr >> 6] = bw;
      if (out_valid) out_valid[wr >> 6] = vw;
    }
    nulls += __popcll(inw & ~vw);
    selected += __popcll(bw);
  }
  __syncthreads();  // the bitmap is rewritten by the next sub-tile
}

#endif

}  // namespace prism
