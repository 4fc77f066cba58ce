// hipRTC compilation and module cache for generated kernels.
#include "jit.hpp"

#include <hip/hiprtc.h>

#include <chrono>
#include <cstdlib>
#include <map>
#include <unordered_map>

namespace prism {
namespace jit {
namespace {

// Device headers the generated sources include, embedded at build time.
struct Header {
  const char* name;
  const char* text;
};
const Header kHeaders[] = {
#include "jit_headers.inc"
    // hipRTC has no C library headers; prism_hip.h only needs size_t from here.
    {"stddef.h", "#pragma once\n"},
};
constexpr int kNumHeaders = sizeof(kHeaders) / sizeof(kHeaders[0]);

struct Entry {
  hipModule_t module = nullptr;  // nullptr: the compile or load failed
  std::map<std::string, hipFunction_t> fns;
};

std::mutex g_mu;
std::unordered_map<std::string, Entry> g_cache;
Stats g_stats;
std::string g_last_error;

std::string arch_of(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "gfx950";
  std::string a = prop.gcnArchName;
  const size_t colon = a.find(':');
  return colon == std::string::npos ? a : a.substr(0, colon);
}

}  // namespace

bool compile_code(const std::string& src, const std::string& arch, std::string& code, std::string& err) {
  const char* names[kNumHeaders];
  const char* texts[kNumHeaders];
  for (int i = 0; i < kNumHeaders; ++i) {
    names[i] = kHeaders[i].name;
    texts[i] = kHeaders[i].text;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "prism_generated.hip", kNumHeaders, texts, names) != HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  const std::string arch_opt = "--offload-arch=" + arch;
  // Same code-generation flags as the prebuilt kernels (no packed fp32).
  const char* opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-munsafe-fp-atomics"};
  const hiprtcResult r = hiprtcCompileProgram(prog, static_cast<int>(sizeof(opts) / sizeof(opts[0])), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    err = std::string("hipRTC: ") + hiprtcGetErrorString(r) + "\n" + log;
    hiprtcDestroyProgram(&prog);
    return false;
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  code.assign(code_size, '\0');
  hiprtcGetCode(prog, &code[0]);
  hiprtcDestroyProgram(&prog);
  return true;
}

namespace {

bool load(int device, const std::string& src, Entry& out, std::string& err) {
  std::string code;
  if (!compile_code(src, arch_of(device), code, err)) return false;
  const hipError_t e = hipModuleLoadData(&out.module, code.data());
  if (e != hipSuccess) {
    err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
    out.module = nullptr;
    return false;
  }
  return true;
}

}  // namespace

bool enabled() {
  const char* v = std::getenv("PRISM_JIT");
  return !(v && v[0] == '0' && v[1] == '\0');
}

hipFunction_t get(int device, const std::string& src, const char* name) {
  std::string key = std::to_string(device) + '\n' + src;
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    ++g_stats.hits;
  } else {
    const auto t0 = std::chrono::steady_clock::now();
    Entry e;
    std::string err;
    if (!load(device, src, e, err)) {
      ++g_stats.failures;
      g_last_error = err;
    } else {
      g_stats.compile_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      ++g_stats.kernels;
    }
    it = g_cache.emplace(std::move(key), std::move(e)).first;
  }
  Entry& e = it->second;
  if (!e.module) return nullptr;
  auto f = e.fns.find(name);
  if (f != e.fns.end()) return f->second;
  hipFunction_t fn = nullptr;
  if (hipModuleGetFunction(&fn, e.module, name) != hipSuccess) {
    ++g_stats.failures;
    g_last_error = std::string("hipModuleGetFunction: no kernel ") + name;
    fn = nullptr;
  }
  e.fns.emplace(name, fn);
  return fn;
}

Stats stats() {
  std::lock_guard<std::mutex> lock(g_mu);
  return g_stats;
}

std::string last_error() {
  std::lock_guard<std::mutex> lock(g_mu);
  return g_last_error;
}

}  // namespace jit
}  // namespace prism

extern "C" {

prism_status prism_jit_get_stats(prism_jit_stats* out) {
  return prism::guarded([&] {
    if (!out) prism::fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL output");
    const prism::jit::Stats s = prism::jit::stats();
    out->kernels = s.kernels;
    out->failures = s.failures;
    out->hits = s.hits;
    out->compile_ms = s.compile_ms;
  });
}

const char* prism_jit_last_error(void) {
  static thread_local std::string msg;
  msg = prism::jit::last_error();
  return msg.c_str();
}

}  // extern "C"
