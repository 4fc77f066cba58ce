// Device-side view of a join hash table: shared by the join kernels, the
// join-aggregate and the operator service.
#pragma once

#include "common.hpp"
#include "dev_common.hpp"

namespace prism {
namespace join {

constexpr int64_t kEmpty = static_cast<int64_t>(0x8000000000000000ULL);  // INT64_MIN
constexpr uint32_t kNone = 0xFFFFFFFFu;
// Slots of one partition the partitioned probe loads into LDS (16 B each).
constexpr uint64_t kPtMaxSlots = 8192;

struct Slot {
  int64_t key;
  uint32_t val;  // unique: build row; duplicates: CSR start
  uint32_t cnt;  // rows of this key
};
static_assert(sizeof(Slot) == 16, "slot is 16 bytes");

struct TableView {
  const Slot* slots;        // HASHED: mask + 2 entries; entry mask+1 holds the INT64_MIN key
  uint64_t mask;
  int shift;
  // HASHED: slots of one partition - 1.  The slot array is cut into equal
  // partitions by the hash's top bits and a probe run wraps inside its own
  // partition, so each partition is an open-addressing table of its own (one
  // partition, pmask = mask, is the plain table).
  uint64_t pmask;
  const uint32_t* dense_u;  // DENSE unique: row per key offset, kNone = absent
  const uint2* dense_nu;    // DENSE duplicates: {start, cnt}
  int64_t min_key;
  uint64_t range;
  const uint32_t* rows;     // CSR rows (duplicates)
  const uint2* dense_up;    // DENSE unique with payload: {row, payload}
  const int32_t* row_payload;  // duplicates with payload: payload of rows[i]
  int has_payload;          // entries carry an int32 build payload
  int sentinel_present;     // HASHED: a build key equals INT64_MIN
  // DENSE key presence only (the join aggregate): bit d of present_bits = a
  // build row has key min_key + d; dense_u holds the rows of each key among
  // a sample of the build rows (an estimate, to pick hot keys); no rows
  int presence_only;
  const uint32_t* present_bits;
};

}  // namespace join
}  // namespace prism

struct prism_join_table {
  prism::Ctx* ctx = nullptr;
  prism_join_table_info info{};
  prism::DeviceBuffer slots;  // HASHED
  prism::DeviceBuffer dense;  // DENSE
  prism::DeviceBuffer rows;   // CSR rows
  prism::DeviceBuffer payload;  // DENSE unique {row, payload} entries / CSR row payloads
  prism::DeviceBuffer bits;     // DENSE presence-only: key presence bitmap
  prism::join::TableView view{};
};

namespace prism {
namespace join {

template <int KIND, bool UNIQUE>
__device__ __forceinline__ bool lookup(const TableView& T, int64_t k, uint32_t& val, uint32_t& cnt, int32_t& pay) {
  if constexpr (KIND == PRISM_TABLE_DENSE) {
    uint64_t d = static_cast<uint64_t>(k) - static_cast<uint64_t>(T.min_key);
    if (d >= T.range) return false;
    if constexpr (UNIQUE) {
      if (T.has_payload) {
        const uint2 e = gload<uint2>(T.dense_up, d);
        val = e.x;
        pay = static_cast<int32_t>(e.y);
        cnt = 1;
        return e.x != kNone;
      }
      uint32_t v = gload<uint32_t>(T.dense_u, d);
      val = v;
      cnt = 1;
      return v != kNone;
    } else {
      uint2 e = gload<uint2>(T.dense_nu, d);
      val = e.x;
      cnt = e.y;
      return e.y != 0;
    }
  } else {
    uint64_t s;
    if (k == kEmpty) {
      if (!T.sentinel_present) return false;
      const Slot e = gload<Slot>(T.slots, T.mask + 1);
      val = e.val;
      cnt = (UNIQUE && T.has_payload) ? 1u : e.cnt;
      pay = static_cast<int32_t>(e.cnt);
      return true;
    }
    s = hash_i64_fast(k) >> T.shift;
    while (true) {  // load factor <= 1/2: an empty slot always ends the run
      const Slot e = gload<Slot>(T.slots, s);
      if (e.key == k) {
        val = e.val;
        cnt = (UNIQUE && T.has_payload) ? 1u : e.cnt;
        pay = static_cast<int32_t>(e.cnt);
        return true;
      }
      if (e.key == kEmpty) return false;
      s = (s & ~T.pmask) | ((s + 1) & T.pmask);
    }
  }
}

template <int KIND, bool UNIQUE>
__device__ __forceinline__ bool lookup(const TableView& T, int64_t k, uint32_t& val, uint32_t& cnt) {
  int32_t pay;
  return lookup<KIND, UNIQUE>(T, k, val, cnt, pay);
}

}  // namespace join
}  // namespace prism
