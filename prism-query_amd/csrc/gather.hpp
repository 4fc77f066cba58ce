// Gather / selection helpers over device columns.
#pragma once

#include "arrow_io.hpp"

namespace prism {

// out[i] = src[idx[i]] for i < n; a null index (bit clear in idx_valid) yields
// a null output row.  The result has offset 0.  A valid index >= src.length
// reads nothing; without err_word the call then fails with PRISM_ERR_ARROW
// (arrow's take refuses it too) and synchronises, with err_word (a device
// word) it sets *err_word to 1 and the caller checks the word later.  Utf8
// sources synchronise either way (the output size is read back).
DevColumn take(prism_ctx* ctx, const DevColumn& src, const uint32_t* idx, const uint8_t* idx_valid, uint64_t n,
               uint32_t* err_word = nullptr);

// flags[r] = 1 for every r in idx[0..n) (flags has n_rows bytes, cleared first).
void mark_rows(prism_ctx* ctx, const uint32_t* idx, uint64_t n, uint8_t* flags, uint64_t n_rows);
// bits (ceil(n/64) words) = (flags[i] != 0) == want
void flags_to_bits(prism_ctx* ctx, const uint8_t* flags, uint64_t n, bool want, unsigned long long* bits);
// Ordered row indices of the set bits; returns the count (synchronises).
uint64_t select_bits(prism_ctx* ctx, const unsigned long long* bits, uint64_t n_rows, DeviceBuffer& out);
void iota(prism_ctx* ctx, uint32_t* out, uint64_t n, uint32_t base);
void fill_bytes(prism_ctx* ctx, uint8_t* p, uint64_t n, uint8_t v);
void copy_u32(prism_ctx* ctx, uint32_t* dst, const uint32_t* src, uint64_t n);

}  // namespace prism
