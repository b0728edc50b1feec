// step_control.hip -- the device-side time-step controller of FinSolveRK as
// its own 64-lane launch after each step attempt (the controller itself is
// step_control_wave), and ConSoln's per-step records.
#include <hip/hip_runtime.h>

#include "../../include/finess_hip.h"
#include "kernels.h"
#include "stage2d_body.h"
#include "step_control.h"
#include "step_control_body.h"

namespace finess {

__global__ __launch_bounds__(64) void k_step_control(StepCtl* c, unsigned long long* smax, StepCtlParams p)
{
    step_control_wave(c, smax, p, threadIdx.x);
}

int launch_step_control(StepCtl* c, unsigned long long* smax, const StepCtlParams& p, hipStream_t st)
{
    hipLaunchKernelGGL(k_step_control, dim3(1), dim3(64), 0, st, c, smax, p);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// ConSoln's sums (lib/2d/ConSoln.cpp:44-57, lib/1d/ConSoln.cpp): for every
// field the sum over the interior of dx*dy*q, reduced in a fixed order (no
// atomics, so the result does not depend on scheduling).  Pass 1: NB blocks
// each sum a contiguous slice of the cells; pass 2: one block adds the
// partial sums and writes the record (t, sums).  Under the device controller
// the record is written only for an accepted attempt not yet recorded
// (ctl->acc > ctl->nrec), at slot nrec - base of the log; without it (ctl
// null) it is written to slot 0.
constexpr int SUM_BLOCKS = 240;

__global__ __launch_bounds__(256) void k_sum_partial(const double* q, long long plane, int M, double w,
                                                     double* partial)
{
    __shared__ double red[4];
    const long long per = (plane + SUM_BLOCKS - 1) / SUM_BLOCKS;
    const long long lo = (long long)blockIdx.x * per, hi = lo + per < plane ? lo + per : plane;
    for (int m = 0; m < M; m++) {
        double s = 0.0;
        for (long long k = lo + threadIdx.x; k < hi; k += blockDim.x) s += w * q[m * plane + k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) partial[m * SUM_BLOCKS + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
        __syncthreads();
    }
}

__global__ __launch_bounds__(64) void k_sum_final(const double* partial, int M, StepCtl* ctl, long long base,
                                                  double t_host, double* log)
{
    long long slot = 0;
    double t = t_host;
    if (ctl) {
        if (!(ctl->acc > ctl->nrec)) return;
        slot = ctl->nrec - base;
        t = ctl->t;
    }
    double* rec = log + slot * (M + 1);
    const int l = threadIdx.x;
    for (int m = 0; m < M; m++) {
        double s = 0.0;
        for (int b = l; b < SUM_BLOCKS; b += 64) s += partial[m * SUM_BLOCKS + b];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (l == 0) rec[1 + m] = s;
    }
    if (l == 0) {
        rec[0] = t;
        if (ctl) ctl->nrec = ctl->nrec + 1;
    }
}

int launch_con_soln(const double* q, long long plane, int M, double w, double* partial, StepCtl* ctl,
                    long long base, double t, double* log, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_partial, dim3(SUM_BLOCKS), dim3(256), 0, st, q, plane, M, w, partial);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, st, partial, M, ctl, base, t, log);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace finess
