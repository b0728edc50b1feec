// weno.h -- fifth-order WENO reconstructions on the device.
//
// JS5 : lib/WenoReconstruct.cpp:90-141 (Jiang & Shu 1996)
// Z5  : lib/WenoReconstruct.cpp:385-438 (Borges et al. 2008)
// FD5 : lib/WenoReconstruct.cpp:580-616 (linear weights 0.1, 0.6, 0.3)
//
// g[0..4] = u_{i-2..i+2}; the result is the value at x_{i+1/2}.
//
// For the common power_param == 2 the JS weights avoid pow():
//  FP64: omega_k = d_k/a_k / sum_j d_j/a_j with a_k = (eps+beta_k)^2 is formed
//        with one division as d_k prod_{j!=k} a_j / sum_j d_j prod_{i!=j} a_i
//        (products of two a's stay inside the FP64 range for eps >= 1e-100);
//  FP32: the same weights as d_k r_k / sum_j d_j r_j with r_k = min(a)/a_k in
//        [0,1], so nothing overflows and a weight that underflows is one whose
//        exact value is below FP32 resolution anyway (eps = 1e-29 gives
//        a = 1e-58, outside the FP32 range).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "fastmath.h"

namespace finess {

enum WenoKind : int {
    WENO_JS5_P2 = 0, WENO_JS5_POW = 1, WENO_Z5 = 2, WENO_FD5 = 3,
    // the face-list path of stage2d_wide.hip
    WENO_JS7 = 4, WENO_JS9 = 5, WENO_JS11 = 6, WENO_Z7 = 7, WENO_Z9 = 8, WENO_FD7 = 9, WENO_FD9 = 10,
    WENO_O1 = 11
};

template <typename R>
struct WenoParams {
    R eps;
    R power;
};

template <typename R>
__host__ __device__ __forceinline__ void weno5_candidates(const R g[5], R& u1, R& u2, R& u3)
{
    u1 = (R(1) / R(3)) * g[0] - (R(7) / R(6)) * g[1] + (R(11) / R(6)) * g[2];
    u2 = (R(-1) / R(6)) * g[1] + (R(5) / R(6)) * g[2] + (R(1) / R(3)) * g[3];
    u3 = (R(1) / R(3)) * g[2] + (R(5) / R(6)) * g[3] - (R(1) / R(6)) * g[4];
}

template <typename R>
__host__ __device__ __forceinline__ void weno5_betas(const R g[5], R& b0, R& b1, R& b2)
{
    const R d0 = g[0] - R(2) * g[1] + g[2], e0 = g[0] - R(4) * g[1] + R(3) * g[2];
    const R d1 = g[1] - R(2) * g[2] + g[3], e1 = g[1] - g[3];
    const R d2 = g[2] - R(2) * g[3] + g[4], e2 = R(3) * g[2] - R(4) * g[3] + g[4];
    b0 = (R(13) / R(12)) * (d0 * d0) + R(0.25) * (e0 * e0);
    b1 = (R(13) / R(12)) * (d1 * d1) + R(0.25) * (e1 * e1);
    b2 = (R(13) / R(12)) * (d2 * d2) + R(0.25) * (e2 * e2);
}

template <typename R>
__host__ __device__ __forceinline__ R fmin3(R a, R b, R c)
{
    return fmin(fmin(a, b), c);
}


// Unnormalised JS5 weights for power_param == 2 (see the header comment).
__host__ __device__ __forceinline__ void js5_p2_terms(double b0, double b1, double b2, double eps,
                                                      double& t0, double& t1, double& t2)
{
    double a0 = eps + b0, a1 = eps + b1, a2 = eps + b2;
    a0 *= a0; a1 *= a1; a2 *= a2;
    t0 = 0.1 * (a1 * a2); t1 = 0.6 * (a0 * a2); t2 = 0.3 * (a0 * a1);
}
__host__ __device__ __forceinline__ void js5_p2_terms(float b0, float b1, float b2, float eps,
                                                      float& t0, float& t1, float& t2)
{
    float a0 = eps + b0, a1 = eps + b1, a2 = eps + b2;
    a0 *= a0; a1 *= a1; a2 *= a2;
    const float amin = fmin3(a0, a1, a2);
    t0 = 0.1f * (amin * frcp(a0)); t1 = 0.6f * (amin * frcp(a1)); t2 = 0.3f * (amin * frcp(a2));
}

// Nonlinear JS weights (normalised, summing to one).
template <int KIND, typename R>
__host__ __device__ __forceinline__ void weno5_weights(const R g[5], const WenoParams<R>& wp, R w[3])
{
    R b0, b1, b2;
    weno5_betas(g, b0, b1, b2);
    if (KIND == WENO_FD5) {
        w[0] = R(0.1); w[1] = R(0.6); w[2] = R(0.3);
    } else if (KIND == WENO_JS5_P2) {
        R t0, t1, t2;
        js5_p2_terms(b0, b1, b2, wp.eps, t0, t1, t2);
        const R inv = R(1) / (t0 + t1 + t2);
        w[0] = t0 * inv; w[1] = t1 * inv; w[2] = t2 * inv;
    } else if (KIND == WENO_JS5_POW) {
        const R t0 = R(0.1) * pow(wp.eps + b0, -wp.power);
        const R t1 = R(0.6) * pow(wp.eps + b1, -wp.power);
        const R t2 = R(0.3) * pow(wp.eps + b2, -wp.power);
        const R inv = R(1) / (t0 + t1 + t2);
        w[0] = t0 * inv; w[1] = t1 * inv; w[2] = t2 * inv;
    } else { // Z5
        const R tau5 = fabs(b0 - b2);
        const R t0 = R(0.1) * (R(1) + pow(tau5 / (wp.eps + b0), wp.power));
        const R t1 = R(0.6) * (R(1) + pow(tau5 / (wp.eps + b1), wp.power));
        const R t2 = R(0.3) * (R(1) + pow(tau5 / (wp.eps + b2), wp.power));
        const R inv = R(1) / (t0 + t1 + t2);
        w[0] = t0 * inv; w[1] = t1 * inv; w[2] = t2 * inv;
    }
}

template <int KIND, typename R>
__host__ __device__ __forceinline__ R weno5(const R g[5], const WenoParams<R>& wp)
{
    R u1, u2, u3;
    weno5_candidates(g, u1, u2, u3);
    if (KIND == WENO_FD5) return R(0.1) * u1 + R(0.6) * u2 + R(0.3) * u3;
    R b0, b1, b2;
    weno5_betas(g, b0, b1, b2);
    if (KIND == WENO_JS5_P2) {
        R t0, t1, t2;
        js5_p2_terms(b0, b1, b2, wp.eps, t0, t1, t2);
        return (t0 * u1 + t1 * u2 + t2 * u3) * rcp_nr(t0 + t1 + t2);
    } else if (KIND == WENO_JS5_POW) {
        const R t0 = R(0.1) * pow(wp.eps + b0, -wp.power);
        const R t1 = R(0.6) * pow(wp.eps + b1, -wp.power);
        const R t2 = R(0.3) * pow(wp.eps + b2, -wp.power);
        return (t0 * u1 + t1 * u2 + t2 * u3) / (t0 + t1 + t2);
    } else {
        const R tau5 = fabs(b0 - b2);
        const R t0 = R(0.1) * (R(1) + pow(tau5 / (wp.eps + b0), wp.power));
        const R t1 = R(0.6) * (R(1) + pow(tau5 / (wp.eps + b1), wp.power));
        const R t2 = R(0.3) * (R(1) + pow(tau5 / (wp.eps + b2), wp.power));
        return (t0 * u1 + t1 * u2 + t2 * u3) / (t0 + t1 + t2);
    }
}

} // namespace finess
