// Shared helpers of libf3d: error reporting, launch checks, wave reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/f3d.h"

namespace f3d {

// Thread-local last error message, read back through f3d_last_error().
void set_error(const char* fmt, ...);

#define F3D_REQUIRE(cond, code, ...)     \
  do {                                   \
    if (!(cond)) {                       \
      ::f3d::set_error(__VA_ARGS__);     \
      return (code);                     \
    }                                    \
  } while (0)

#define F3D_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::f3d::set_error("%s failed: %s", #expr, hipGetErrorString(e_));        \
      return F3D_EHIP;                                                        \
    }                                                                         \
  } while (0)

// Launch a kernel and report a launch failure as F3D_EHIP.
#define F3D_LAUNCH(kernel, grid, block, lds, stream, ...)                     \
  do {                                                                        \
    hipLaunchKernelGGL(kernel, grid, block, lds, stream, __VA_ARGS__);        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ::f3d::set_error("launch of %s failed: %s", #kernel,                    \
                       hipGetErrorString(e_));                                \
      return F3D_EHIP;                                                        \
    }                                                                         \
  } while (0)

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

constexpr double kPi = 3.14159265358979323846;

// ---- wave64 reductions (device) -----------------------------------------
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace f3d
