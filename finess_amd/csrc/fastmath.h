// fastmath.h -- reciprocal, division and square root for the FP32 stencil
// form.  The FP32 form keeps its values to FP32 resolution, so the IEEE
// division and square-root sequences (about ten instructions each) are
// replaced by the hardware estimates (within one ulp); the FP64 overloads
// keep the IEEE operations, so the plain FP64 form is unchanged bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace finess {

__host__ __device__ __forceinline__ float frcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// 1/x to full FP64 accuracy (within an ulp) without the IEEE division
// sequence: hardware reciprocal estimate refined by two Newton steps.
__host__ __device__ __forceinline__ double rcp_nr(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
#else
    return 1.0 / x;
#endif
}
__host__ __device__ __forceinline__ float rcp_nr(float x) { return frcp(x); }

// a / b: IEEE in FP64, a times the reciprocal estimate in FP32.
__host__ __device__ __forceinline__ double div_lp(double a, double b) { return a / b; }
__host__ __device__ __forceinline__ float div_lp(float a, float b) { return a * frcp(b); }

// sqrt: IEEE in FP64, the hardware square root (within one ulp) in FP32.
__host__ __device__ __forceinline__ double sqrt_lp(double x) { return sqrt(x); }
__host__ __device__ __forceinline__ float sqrt_lp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

} // namespace finess
