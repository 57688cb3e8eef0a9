// Shared device/host helpers of libmphip (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mphip.h"

#define MP_API extern "C" __attribute__((visibility("default")))

namespace mp {

// ---- error plumbing (thread-local message, integer status across the ABI) -------------
void set_error(const char *fmt, ...);
int fail_arg(const char *fmt, ...);
int check_hip(hipError_t e, const char *what);
hipStream_t stream();
// Temporarily redirect launches to `s`; returns the previous stream.
hipStream_t swap_stream(hipStream_t s);

// per-kernel timing hooks (no-ops unless mp_profile_enable(1))
struct KernelTimer {
    explicit KernelTimer(const char *name);
    ~KernelTimer();
    const char *name_;
    void *start_ = nullptr;
};

#define MP_HIP(call)                                          \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return ::mp::check_hip(e_, #call); \
    } while (0)

#define MP_LAUNCH_CHECK(name)                                 \
    do {                                                      \
        hipError_t e_ = hipGetLastError();                    \
        if (e_ != hipSuccess) return ::mp::check_hip(e_, name); \
    } while (0)

// ---- storage types ----------------------------------------------------------------------
struct bf16 { uint16_t v; };
struct f16s { __half v; };

// Load a stored value as the compute type; store with round-to-nearest-even.
template <typename C, typename S> __device__ __forceinline__ C ld(const S &x) { return (C)x; }
template <typename S, typename C> __device__ __forceinline__ S st(C x) { return (S)x; }

template <> __device__ __forceinline__ float ld<float, bf16>(const bf16 &x) {
    return __uint_as_float(((uint32_t)x.v) << 16);
}
template <> __device__ __forceinline__ bf16 st<bf16, float>(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return bf16{(uint16_t)((u >> 16) | 0x40u)};  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16{(uint16_t)(u >> 16)};
}
template <> __device__ __forceinline__ float ld<float, f16s>(const f16s &x) { return __half2float(x.v); }
template <> __device__ __forceinline__ f16s st<f16s, float>(float x) { return f16s{__float2half(x)}; }

// ---- Julia float semantics ----------------------------------------------------------------
// max/min propagate NaN and order -0.0 < 0.0 (Base.max / Base.min for IEEE floats);
// clamp(x, lo, hi) = x > hi ? hi : (x < lo ? lo : x).
__device__ __forceinline__ double jl_max(double x, double y) {
    if ((y > x) || (__builtin_signbit(y) < __builtin_signbit(x))) return __builtin_isnan(x) ? x : y;
    return __builtin_isnan(y) ? y : x;
}
__device__ __forceinline__ double jl_clamp(double x, double lo, double hi) {
    return x > hi ? hi : (x < lo ? lo : x);
}

template <typename C> __device__ __forceinline__ C fmaT(C a, C b, C c);
template <> __device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename C> __device__ __forceinline__ bool finiteT(C x) { return __builtin_isfinite(x); }

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace mp
