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

// true when no pointer is NULL (argument checks of the entry points)
inline bool all_set() { return true; }
template <class P, class... R>
inline bool all_set(const P* p, const R*... r) {
    return p != nullptr && all_set(r...);
}

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

// Pointers read from descriptor structs are generic; accessing them through
// the global address space gives global_load/store (vmcnt only) instead of
// flat accesses that also count against lgkmcnt and so serialise with LDS.
#define D2IC_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const D2IC_GLOBAL T* gld(const T* p) {
    return (const D2IC_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ D2IC_GLOBAL T* gst(T* p) {
    return (D2IC_GLOBAL T*)p;
}

// Wave64 sums through DPP (VALU, no LDS): quad_perm [1,0,3,2] and [2,3,0,1],
// row_half_mirror, row_mirror reduce every 16-lane row; row_bcast15 and
// row_bcast31 (rows masked) fold the rows.  The total is exact in lane 63
// only; the association order is fixed.
template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, Ctrl, RowMask, 0xF, false);
}

template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(dpp_i<Ctrl, RowMask>(__float_as_int(v)));
}

template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<Ctrl, RowMask>((int)(b & 0xffffffffll));
    const int hi = dpp_i<Ctrl, RowMask>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ float wave_sum63(float v) {
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    v += dpp_f<0x142, 0xA>(v);
    v += dpp_f<0x143, 0xC>(v);
    return v;
}

template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
    const unsigned lo = (unsigned)dpp_i<Ctrl, RowMask>((int)(unsigned)v);
    const unsigned hi = (unsigned)dpp_i<Ctrl, RowMask>((int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long wave_sum63(unsigned long long v) {
    v += dpp_u64<0xB1>(v);
    v += dpp_u64<0x4E>(v);
    v += dpp_u64<0x141>(v);
    v += dpp_u64<0x140>(v);
    v += dpp_u64<0x142, 0xA>(v);
    v += dpp_u64<0x143, 0xC>(v);
    return v;
}

__device__ __forceinline__ double wave_sum63(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    v += dpp_d<0x142, 0xA>(v);
    v += dpp_d<0x143, 0xC>(v);
    return v;
}

}  // namespace d2ic
