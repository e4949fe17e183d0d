// Shared helpers for the d2ic_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/d2ic_amd.h"

namespace d2ic {

// Thread-local last error message; every extern "C" entry point returns a
// status code and records the message here.
void set_error(const char* fmt, ...);
const char* last_error();

#define D2IC_HIP(call)                                                                 \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            ::d2ic::set_error("%s failed at %s:%d: %s", #call, __FILE__, __LINE__,     \
                              hipGetErrorString(_e));                                  \
            return D2IC_ERR_DEVICE;                                                    \
        }                                                                              \
    } while (0)

#define D2IC_CHECK_ARG(cond, ...)                                                      \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::d2ic::set_error(__VA_ARGS__);                                            \
            return D2IC_ERR_INVALID;                                                   \
        }                                                                              \
    } while (0)

#define D2IC_LAUNCH_CHECK()                                                            \
    do {                                                                               \
        hipError_t _e = hipGetLastError();                                             \
        if (_e != hipSuccess) {                                                        \
            ::d2ic::set_error("kernel launch failed at %s:%d: %s", __FILE__, __LINE__, \
                              hipGetErrorString(_e));                                  \
            return D2IC_ERR_DEVICE;                                                    \
        }                                                                              \
    } while (0)

constexpr int kWave = 64;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Wave64 reductions through DPP/permute (no LDS, no barriers).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

}  // namespace d2ic
