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

// Global -> LDS copies (global_load_lds_dwordx4) -----------------------------------
// Each lane copies 16 bytes from its own global address to lds_base + 16 * lane
// (lds_base is wave-uniform and goes through M0). The copy is issued from inline
// asm on purpose: the compiler treats the builtin form as a store to LDS that may
// alias every later LDS read of the kernel and waits (vmcnt(0)) for the copy
// before the next fragment read, which serialises copy and compute in a
// multi-buffered loop. Kernels using this helper therefore order the copies
// themselves: wait_vmcnt<N>() before the barrier that publishes a stage, with N
// = the copies this wave issued for stages that may still be in flight.
// The instruction has no VGPR result, so nothing in the asm block is written
// late. M0 is written only here in these kernels.
typedef __attribute__((address_space(3))) uint8_t lds_u8_t;

__device__ __forceinline__ void lds_copy16(const void* gsrc, const uint8_t* lds_base) {
  const uint32_t m0v = (uint32_t)(uintptr_t)(const lds_u8_t*)lds_base;
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(gsrc), "s"(m0v)
               : "memory", "m0");
}

// s_waitcnt vmcnt(N) leaving expcnt / lgkmcnt alone (gfx9 encoding: vmcnt[3:0] in
// bits 3:0, vmcnt[5:4] in bits 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

}  // namespace jm
