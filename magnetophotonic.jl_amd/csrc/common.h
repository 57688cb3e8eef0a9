// Shared device/host helpers of libmphip (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
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
struct f16s { uint16_t v; };  // IEEE binary16 bit pattern

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
// binary16 in integer arithmetic: round to nearest even, subnormals kept, |x| >= 65520 to
// infinity, NaN quieted, and the sign of zero kept (the hardware conversion maps -0 to +0).
template <> __device__ __forceinline__ float ld<float, f16s>(const f16s &x) {
    const uint32_t sign = (uint32_t)(x.v & 0x8000u) << 16;
    const uint32_t e = (x.v >> 10) & 0x1fu, m = x.v & 0x3ffu;
    if (e == 0) return __uint_as_float(__float_as_uint((float)m * 5.9604644775390625e-8f) | sign);
    if (e == 31) return __uint_as_float(sign | 0x7f800000u | (m << 13));
    return __uint_as_float(sign | ((e + 112u) << 23) | (m << 13));
}
template <> __device__ __forceinline__ f16s st<f16s, float>(float x) {
    uint32_t u = __float_as_uint(x);
    const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return f16s{(uint16_t)(sign | 0x7e00u | ((u >> 13) & 0x3ffu))};
    if (u >= 0x477ff000u) return f16s{(uint16_t)(sign | 0x7c00u)};
    if (u >= 0x38800000u) {
        u += 0xfffu + ((u >> 13) & 1u);
        return f16s{(uint16_t)(sign | ((u - 0x38000000u) >> 13))};
    }
    const float q = __uint_as_float(u) * 16777216.0f;  // |x| / 2^-24, exact
    return f16s{(uint16_t)(sign | (uint16_t)__builtin_rintf(q))};
}

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
