// rk_low_storage.hip -- the elementwise parts of the 10-stage fourth-order
// and the 8-stage fifth-order Runge-Kutta schemes (lib/2d/FinSolveRK.cpp:
// 213-305, lib/2d/UpdateSoln.cpp:40-80); their stages are the stage kernel.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace finess {

// SSP(10,4) after stage 5 (lib/2d/FinSolveRK.cpp:240-249): q2 starts as q^n,
// q2 = (q2 + 9 q1)/25, q1 = 15 q2 - 5 q1.
__global__ __launch_bounds__(256) void k_rk4_combine(const double* qn, double* q2, double* q1, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double tmp = (qn[k] + 9.0 * q1[k]) / 25.0;
    q2[k] = tmp;
    q1[k] = 15.0 * tmp - 5.0 * q1[k];
}

// One stage of the two-register fifth-order scheme: q2 = q2 + delta q1,
// out = g1 q1 + g2 q2 + g3 q^n + beta dt L (q2 starts at zero: first stage).
__global__ __launch_bounds__(256) void k_rk5_update(const double* q1, double* q2, const double* qold,
                                                    const double* L, double* out, long long n, double g1, double g2,
                                                    double g3, double delta, double bdt, int first)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double s1 = q1[k];
    const double s2 = (first ? 0.0 : q2[k]) + delta * s1;
    q2[k] = s2;
    out[k] = g1 * s1 + g2 * s2 + g3 * qold[k] + bdt * L[k];
}

int launch_rk4_combine(const double* qn, double* q2, double* q1, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_rk4_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, qn, q2, q1, n);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_rk5_update(const double* q1, double* q2, const double* qold, const double* L, double* out, long long n,
                      double g1, double g2, double g3, double delta, double bdt, int first, hipStream_t st)
{
    hipLaunchKernelGGL(k_rk5_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, q1, q2, qold, L, out, n,
                       g1, g2, g3, delta, bdt, first);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace finess
