// Group-by update kernels generated per plan (gb_jit.hip) and compiled with
// hipRTC (jit.hpp).
#pragma once

#include <string>

#include "groupby.hpp"

namespace prism {
namespace gb {

// Accumulation forms of a generated kernel.
constexpr int kJitWave = 0;    // wave-level pre-aggregation (few groups per wave)
constexpr int kJitShared = 1;  // per-row LDS / HBM slots

constexpr const char* kJitGroupbyKernel = "prism_gb_update";  // k_groupby's arguments
constexpr const char* kJitGroupbyTiles = "prism_gb_tiles";    // (TileQueue*, RowTile*, TileReq*)
// Line of the generated source naming the occupancy hint of the kernel;
// replaced by one asking for at least N waves per SIMD when the register
// count sits just above the budget of N waves.
constexpr const char* kJitWavesHint = "#define PRISM_WAVES_HINT";

// HIP source of the update kernels of `g` for one launch shape: the form,
// rows per lane and the columns that may hold nulls (bit c = column c).  It
// holds kJitGroupbyKernel (one launch over a row range, k_groupby's
// arguments) and kJitGroupbyTiles (persistent tile-queue form).
std::string jit_groupby_source(const prism_groupby* g, int form, int rows, uint32_t nullable);

struct JitKernels {
  hipFunction_t update = nullptr;
  hipFunction_t tiles = nullptr;
};
// Compiled kernels of jit_groupby_source, cached for the process by plan and
// launch shape; empty when generated kernels are off or the compile failed.
// A kernel whose registers sit just above the budget of one more wave per
// SIMD is compiled again asking for that occupancy, and kept if it spills
// nothing.
JitKernels jit_kernels(const prism_groupby* g, int form, int rows, uint32_t nullable);

}  // namespace gb
}  // namespace prism
