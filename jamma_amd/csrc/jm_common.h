// Shared helpers for the jamma_amd HIP library (gfx950 / CDNA4 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/jamma_hip.h"

namespace jm {

// Per-thread last error message (C-ABI: jm_last_error).
void set_error(const char* fmt, ...);

#define JM_CHECK_ARG(cond, ...)          \
  do {                                   \
    if (!(cond)) {                       \
      ::jm::set_error(__VA_ARGS__);      \
      return JM_ERR_ARG;                 \
    }                                    \
  } while (0)

#define JM_CHECK_HIP(expr)                                                          \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::jm::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
      return JM_ERR_HIP;                                                            \
    }                                                                               \
  } while (0)

#define JM_CHECK_LAUNCH() JM_CHECK_HIP(hipGetLastError())

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

constexpr int kWave = 64;

// Wave-level helpers -----------------------------------------------------------
__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_d(v, m);
  return v;
}

}  // namespace jm
