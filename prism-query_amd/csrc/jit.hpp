// Kernels generated at run time: HIP source specialised to one plan (column
// types, constants, aggregate lanes, null-free columns), compiled with hipRTC
// for the device's own architecture and loaded as a module.  Compiled kernels
// are cached per device and source text for the life of the process.
#pragma once

#include <string>

#include "common.hpp"

namespace prism {
namespace jit {

// PRISM_JIT=0 in the environment turns generated kernels off (the prebuilt
// interpreting kernels run instead).  Read at every call.
bool enabled();

// Kernel `name` of `src`, compiled and loaded once per device and source (a
// failed compile is remembered too).  On a compile or load failure returns
// nullptr and keeps the compiler log as the last JIT error.
hipFunction_t get(int device, const std::string& src, const char* name);

// Compile `src` for `arch` (e.g. "gfx950") into a code object; no device is
// needed.  On failure returns false with the compiler log in `err`.
bool compile_code(const std::string& src, const std::string& arch, std::string& code, std::string& err);

struct Stats {
  uint64_t kernels = 0;   // distinct sources compiled
  uint64_t failures = 0;  // failed compiles / loads
  uint64_t hits = 0;      // lookups served from the cache
  double compile_ms = 0;  // total compile + load time
};
Stats stats();
std::string last_error();

}  // namespace jit
}  // namespace prism
