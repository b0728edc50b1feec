// step_control.h -- the time-step controller run on the device, one small
// launch after each step attempt, so that a run of steps is enqueued without
// a host round trip per step.  The host reads the block back once per chunk
// of steps.
#pragma once
#include <hip/hip_runtime.h>

namespace finess {

enum : unsigned long long { HALT_NONE = 0, HALT_DONE = 1, HALT_REJECT = 2, HALT_ERROR = 3 };

struct StepCtl {
    double dt;          // dt of the next attempt (clipped to tend)
    double dt_next;     // the controller's dt before clipping (dtv[1] on return)
    double t;           // time after the last accepted step
    double cfl;         // CFL number of the last attempt
    double dtmin, dtmax;
    long long n_step;   // outer iterations so far (the reference's n_step)
    long long acc, rej; // accepted / rejected attempts
    long long done;     // attempts the controller has processed
    unsigned long long halt;
    long long err;      // FINESS_ERR_* when halt == HALT_ERROR
    long long nrec;     // accepted steps whose conservation record is written
    long long buf;      // 1D run kernel: the plane set holding the last accepted state
};

// The SSP Runge-Kutta schemes of orders 1-3 as stages over three plane sets
// (0 = A = q^n, never written during a step; 1 = B; 2 = C; -1 = none):
// stage k of order n forms out = a1 qa + a2 qb + beta dt L(src), with the
// coefficients of lib/2d/SetRKinfo.cpp:50-64 and the operand order of
// lib/2d/FinSolveRK.cpp:158-211 (lib/1d/FinSolveRK.cpp:129-200).  One table
// for every path that runs these stages.
struct SspStage {
    int src, qa, qb, out;
    double a1, a2, beta;
};

__host__ __device__ __forceinline__ SspStage ssp_stage(int n, int k)
{
    if (k == 0) return {0, 0, -1, 1, 1.0, 0.0, 1.0};
    if (n == 2) return {1, 1, 0, 2, 0.5, 0.5, 0.5};
    if (k == 1) return {1, 0, 1, 2, 0.75, 0.25, 0.25};
    return {2, 2, 0, 1, 2.0e0 / 3.0e0, 1.0e0 / 3.0e0, 2.0e0 / 3.0e0};
}

// the plane set that holds q^{n+1} after the last stage
__host__ __device__ __forceinline__ int ssp_result(int n) { return n == 2 ? 2 : 1; }

struct StepCtlParams {
    double tend, max_dt, cfl_max, cfl_target, dx, dy;
    int one_d;          // 1D GetCFL (lib/1d/GetCFL.cpp): x speeds only, bound 1e10
    long long nv;
    long long target;   // accepted steps to make; 0: run until t reaches tend
};

} // namespace finess
