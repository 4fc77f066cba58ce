// Shared host/device helpers for the gfx950 kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "prism_hip.h"

namespace prism {

// ---------------------------------------------------------------------------
// Errors.  PrismError (native/prism-executor/src/error.rs:5-27) becomes an
// exception inside the library and a status code at the C boundary.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  prism_status status;
  Error(prism_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void fail(prism_status s, const std::string& m) { throw Error(s, m); }

void set_last_error(const std::string& msg);
void clear_last_error();

template <class F>
prism_status guarded(F&& f) {
  try {
    clear_last_error();
    f();
    return PRISM_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("Execution error: host allocation failed");
    return PRISM_ERR_EXECUTION;
  } catch (const std::exception& e) {
    set_last_error(std::string("Internal error: ") + e.what());
    return PRISM_ERR_INTERNAL;
  } catch (...) {
    set_last_error("Internal error: unknown exception");
    return PRISM_ERR_INTERNAL;
  }
}

#define PRISM_HIP(expr)                                                                  \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      ::prism::fail(PRISM_ERR_EXECUTION, std::string("Execution error: HIP ") + #expr +  \
                                             " failed: " + hipGetErrorString(_e));       \
  } while (0)

#define PRISM_CHECK_LAUNCH() PRISM_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Device memory owned by a context.
// ---------------------------------------------------------------------------
struct Ctx;

struct DeviceBuffer {
  Ctx* ctx = nullptr;
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(Ctx* c, size_t n) { alloc(c, n); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept { *this = std::move(o); }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ctx = o.ctx; ptr = o.ptr; bytes = o.bytes;
      o.ctx = nullptr; o.ptr = nullptr; o.bytes = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }
  void alloc(Ctx* c, size_t n);
  // Grow-only: keeps the allocation when it is already large enough.
  void reserve(Ctx* c, size_t n) {
    if (n > bytes) { release(); alloc(c, n); }
  }
  void release();
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

}  // namespace prism

struct prism_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::atomic<uint64_t> device_bytes{0};
  int num_cus = 256;
  int num_xcds = 8;
  size_t lds_per_block = 64 * 1024;
  std::mutex mu;
  // Small per-context device scratch (tile status words, counters, results).
  prism::DeviceBuffer scratch;
  void* pinned = nullptr;   // small pinned host staging for results
  size_t pinned_bytes = 0;
  void* readback = nullptr;  // 4 KiB pinned block for small device -> host reads under `mu`
  hipEvent_t t0 = nullptr, t1 = nullptr;  // timer markers
  // Stream-ordered cache of device blocks released by library-owned buffers:
  // a block is reused only by work queued later on the context's stream, so
  // no synchronisation is needed; hipFree happens at trim/destroy.
  std::mutex pool_mu;
  std::multimap<size_t, void*> pool;
  size_t pooled_bytes = 0;
  size_t pool_limit = size_t(16) << 30;
  // Per-kernel timing (event pairs per kernel name) while enabled.
  bool ktime_on = false;
  std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> ktime;
  // The pinned staging buffer is shared by every upload on this context.
  std::mutex staging_mu;
  // Lanes: child contexts (own stream, scratch, staging, block cache) that the
  // Arrow-layer operator entries borrow for one call, so that calls from
  // several threads on one context run side by side on separate streams.
  prism_ctx* parent = nullptr;
  std::mutex lane_mu;
  std::vector<prism_ctx*> lanes, idle_lanes;
  // Held by a call that runs on a caller-chosen stream (prism_ctx_set_stream):
  // such calls run one at a time, in the caller's stream order.
  std::mutex serial_mu;
};

namespace prism {
struct Ctx : prism_ctx {};
inline Ctx* as_ctx(prism_ctx* c) { return static_cast<Ctx*>(c); }

// Make the context's device current on this thread.
void use_device(prism_ctx* ctx);

// Pinned host staging of at least n bytes owned by the context.
void* pinned_staging(prism_ctx* ctx, size_t n);

// Pinned 4 KiB block for small synchronous device -> host reads (statistics,
// counters); callers hold ctx->mu.
// Pinned block per context for small readbacks and uploads, used by one call
// at a time (under the context lock, drained before the call returns):
// [0, 4 KiB) statistics, counters and column descriptors; from 4 KiB the join
// aggregate's create (attribute range, hot-slot candidates and placement).
constexpr size_t kReadbackBytes = 64 << 10;
void* readback_block(prism_ctx* ctx);

// Return every cached block to the device (hipFree).
void pool_trim(prism_ctx* ctx);

// Borrows a lane of `ctx` for the lifetime of the guard and points `ctx` at
// it.  Work queued on the lane must be finished (or be ordered before any
// later use of what it writes) when the guard ends: the Arrow-layer entries
// synchronise before they return.  Already a lane: nothing to do.  A context
// whose stream the caller chose keeps that stream and runs calls one at a time.
class Lane {
 public:
  explicit Lane(prism_ctx*& ctx);
  ~Lane();
  Lane(const Lane&) = delete;
  Lane& operator=(const Lane&) = delete;

 private:
  prism_ctx* parent_ = nullptr;
  prism_ctx* lane_ = nullptr;
  std::unique_lock<std::mutex> serial_;
};

// Brackets the launches made during its lifetime with events on the context
// stream when per-kernel timing is enabled.
struct KernelTimer {
  prism_ctx* ctx;
  const char* name;
  hipEvent_t a = nullptr;
  KernelTimer(prism_ctx* c, const char* n);
  ~KernelTimer();
};

}  // namespace prism

// ---------------------------------------------------------------------------
// Device-side helpers (dev_common.hpp) and host integer helpers.
// ---------------------------------------------------------------------------
#include "dev_common.hpp"

namespace prism {

inline uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}
inline int log2_pow2(uint64_t p) {
  int l = 0;
  while ((1ULL << l) < p) ++l;
  return l;
}

}  // namespace prism
