// Context, error state and device-memory bookkeeping behind the C ABI.
#include "common.hpp"

namespace prism {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

void DeviceBuffer::alloc(Ctx* c, size_t n) {
  ctx = c;
  bytes = 0;
  ptr = nullptr;
  if (n == 0) return;
  hipError_t e = hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    fail(PRISM_ERR_EXECUTION, "Execution error: hipMalloc of " + std::to_string(n) +
                                  " bytes failed: " + hipGetErrorString(e));
  }
  bytes = n;
  if (ctx) ctx->device_bytes += n;
}

void DeviceBuffer::release() {
  if (ptr) {
    (void)hipFree(ptr);
    if (ctx) ctx->device_bytes -= bytes;
  }
  ptr = nullptr;
  bytes = 0;
}

void use_device(prism_ctx* ctx) { PRISM_HIP(hipSetDevice(ctx->device)); }

void* pinned_staging(prism_ctx* ctx, size_t n) {
  if (n > ctx->pinned_bytes) {
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
    size_t sz = next_pow2(n < 4096 ? 4096 : n);
    PRISM_HIP(hipHostMalloc(&ctx->pinned, sz, hipHostMallocDefault));
    ctx->pinned_bytes = sz;
  }
  return ctx->pinned;
}

}  // namespace prism

using namespace prism;

extern "C" {

const char* prism_last_error(void) { return g_last_error.c_str(); }

const char* prism_version(void) { return "prism-hip 0.1.0 (gfx950)"; }

prism_status prism_ctx_create(int device, prism_ctx** out) {
  return guarded([&] {
    if (!out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: out is NULL");
    *out = nullptr;
    int n = 0;
    PRISM_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: device " + std::to_string(device) +
                                           " out of range (" + std::to_string(n) + " devices)");
    PRISM_HIP(hipSetDevice(device));
    auto* c = new Ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
      c->num_cus = prop.multiProcessorCount;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      fail(PRISM_ERR_EXECUTION, std::string("Execution error: stream creation failed: ") +
                                    hipGetErrorString(e));
    }
    c->stream = c->own_stream;
    *out = c;
  });
}

void prism_ctx_destroy(prism_ctx* ctx) {
  if (!ctx) return;
  auto* c = as_ctx(ctx);
  (void)hipSetDevice(c->device);
  if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
  c->scratch.release();
  if (c->t0) (void)hipEventDestroy(c->t0);
  if (c->t1) (void)hipEventDestroy(c->t1);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

prism_status prism_ctx_set_stream(prism_ctx* ctx, void* hip_stream) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: ctx is NULL");
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
  });
}

void* prism_ctx_stream(prism_ctx* ctx) { return ctx ? ctx->stream : nullptr; }

prism_status prism_ctx_synchronize(prism_ctx* ctx) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: ctx is NULL");
    use_device(ctx);
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

uint64_t prism_ctx_device_bytes(prism_ctx* ctx) { return ctx ? ctx->device_bytes.load() : 0; }

prism_status prism_device_alloc(prism_ctx* ctx, uint64_t bytes, void** out) {
  return guarded([&] {
    if (!ctx || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context/output");
    use_device(ctx);
    *out = nullptr;
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *out = nullptr;
      fail(PRISM_ERR_EXECUTION, std::string("Execution error: hipMalloc failed: ") + hipGetErrorString(e));
    }
    ctx->device_bytes += bytes ? bytes : 1;
  });
}

prism_status prism_device_free(prism_ctx* ctx, void* p) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context");
    if (!p) return;
    use_device(ctx);
    size_t sz = 0;
    void* base = nullptr;
    if (hipMemGetAddressRange(&base, &sz, p) == hipSuccess && base == p) ctx->device_bytes -= sz;
    PRISM_HIP(hipFree(p));
  });
}

prism_status prism_copy_to_device(prism_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context");
    if (!bytes) return;
    use_device(ctx);
    PRISM_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

prism_status prism_copy_to_host(prism_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context");
    if (!bytes) return;
    use_device(ctx);
    PRISM_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

prism_status prism_copy_device(prism_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context");
    if (!bytes) return;
    use_device(ctx);
    PRISM_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  });
}

prism_status prism_ctx_timer_begin(prism_ctx* ctx) {
  return guarded([&] {
    if (!ctx) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context");
    use_device(ctx);
    if (!ctx->t0) PRISM_HIP(hipEventCreate(&ctx->t0));
    if (!ctx->t1) PRISM_HIP(hipEventCreate(&ctx->t1));
    PRISM_HIP(hipEventRecord(ctx->t0, ctx->stream));
  });
}

prism_status prism_ctx_timer_end(prism_ctx* ctx, double* ms) {
  return guarded([&] {
    if (!ctx || !ms || !ctx->t0) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: timer not started");
    use_device(ctx);
    PRISM_HIP(hipEventRecord(ctx->t1, ctx->stream));
    PRISM_HIP(hipEventSynchronize(ctx->t1));
    float f = 0.0f;
    PRISM_HIP(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
    *ms = f;
  });
}

}  // extern "C"
