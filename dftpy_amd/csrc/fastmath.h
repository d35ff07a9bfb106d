// Transcendental functions of the point-wise and G-space kernels, by
// precision.  FP64 keeps the correctly rounded library functions.  FP32 (the
// FP32-bulk pipeline, whose accuracy is set by single-precision storage and
// FFTs anyway) uses the hardware approximations v_sqrt_f32, v_rsq_f32,
// v_rcp_f32, v_log_f32 (log2) and v_exp_f32 (exp2), each within about 1 ulp,
// instead of the library's range-reduced forms of 10 to 40 instructions.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace ofdft {
namespace fm {

template <typename T> __device__ __forceinline__ T sqrt(T x) { return ::sqrt(x); }
template <> __device__ __forceinline__ float sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T> __device__ __forceinline__ T rcp(T x) { return (T)1 / x; }
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T> __device__ __forceinline__ T rsq(T x) { return (T)1 / ::sqrt(x); }
template <> __device__ __forceinline__ float rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// a / b
template <typename T> __device__ __forceinline__ T div(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// natural logarithm for arguments away from 1 (absolute error of log2 about
// 1 ulp of its result's scale)
template <typename T> __device__ __forceinline__ T log(T x) { return ::log(x); }
template <> __device__ __forceinline__ float log<float>(float x) {
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}

// x^y for x > 0
template <typename T> __device__ __forceinline__ T pow(T x, T y) { return ::pow(x, y); }
template <> __device__ __forceinline__ float pow<float>(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// cube root for x >= 0: exp2(log2(x) / 3) refined by one Newton step
// y <- y - (y^3 - x) / (3 y^2)
template <typename T> __device__ __forceinline__ T cbrt(T x) { return ::cbrt(x); }
template <> __device__ __forceinline__ float cbrt<float>(float x) {
    const float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f));
    const float r = __builtin_amdgcn_rcpf(y * y);
    const float z = y - (y - x * r) * (1.0f / 3.0f);
    return x > 0.0f ? z : 0.0f;
}

}  // namespace fm
}  // namespace ofdft
