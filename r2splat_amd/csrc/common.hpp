// Shared device/host helpers for the gfx950 CT splatting kernels.
//
// Everything in this library is plain HIP for CDNA4 (wave64). Host entry points
// are `extern "C"` and take raw device pointers, sizes and a hipStream_t; the
// caller owns every allocation (outputs and workspaces), so no entry point
// allocates device memory and all of them can be captured into a hipGraph.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <algorithm>

#include "../../include/r2splat_hip.h"

namespace r2s {

constexpr int kWave = 64;

// A dispatch's grid size counts work-items in 32 bits, so blocks * threads
// must stay below 2^32. Kernels with one wave per element (intersections can
// number in the hundreds of millions) loop over a grid capped at this many
// blocks instead of launching one block per few elements.
constexpr uint32_t kIsectBlocks = 1u << 16;

// ---------------------------------------------------------------------------
// Error reporting: thread-local message + status code.
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int check_launch(const char* where);

#define R2S_REQUIRE(cond, ...)                 \
    do {                                       \
        if (!(cond)) {                         \
            ::r2s::set_error(__VA_ARGS__);     \
            return R2S_ERR_INVALID;            \
        }                                      \
    } while (0)

#define R2S_HIP(call)                                                         \
    do {                                                                      \
        hipError_t _e = (call);                                               \
        if (_e != hipSuccess) {                                               \
            ::r2s::set_error("%s failed: %s", #call, hipGetErrorString(_e));  \
            return R2S_ERR_HIP;                                               \
        }                                                                     \
    } while (0)

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Precision-dependent math. Float uses the hardware transcendental units
// (v_exp_f32 / v_log_f32 are base-2); double uses the IEEE library routines.
// ---------------------------------------------------------------------------
template <typename S> struct Math;

template <> struct Math<float> {
    static __device__ __forceinline__ float exp(float x) { return __expf(x); }
    static __device__ __forceinline__ float log(float x) { return __logf(x); }
    static __device__ __forceinline__ float pow(float b, float e) { return __expf(e * __logf(b)); }
    static __device__ __forceinline__ float sqrt(float x) { return sqrtf(x); }
};

template <> struct Math<double> {
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double pow(double b, double e) { return ::pow(b, e); }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
};

// Type-dispatch helper for the C-ABI dtype code.
template <typename F>
inline int dispatch_dtype(int dtype, F&& f) {
    if (dtype == R2S_F32) return f(float{});
    if (dtype == R2S_F64) return f(double{});
    set_error("unsupported dtype code %d (expected R2S_F32 or R2S_F64)", dtype);
    return R2S_ERR_INVALID;
}

// ---------------------------------------------------------------------------
// Wave64 reductions (butterfly over 64 lanes; the compiler lowers the xor
// shuffles to DPP / ds_swizzle on gfx950).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Device atomics on float / double with hardware fp atomics (-munsafe-fp-atomics).
template <typename S>
__device__ __forceinline__ void atomic_add(S* p, S v) { atomicAdd(p, v); }

}  // namespace r2s
