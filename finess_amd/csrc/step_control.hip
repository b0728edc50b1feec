// step_control.hip -- device-side time-step controller of FinSolveRK
// (lib/2d/FinSolveRK.cpp:77-106 and 318-350, with GetCFL from
// lib/2d/GetCFL.cpp:31-74).  One 64-lane workgroup per step attempt: it
// reduces the attempt's wave-speed maxima (and clears them for the next
// attempt), forms the CFL number, accepts or rejects, sets the next dt and
// stops the run (halt) at the end, at a rejection (the host re-plans from
// the rejected attempt) or at an error.  The arithmetic is the host
// controller's, operation for operation.
#include <hip/hip_runtime.h>

#include "../../include/finess_hip.h"
#include "kernels.h"
#include "stage2d_body.h"
#include "step_control.h"

namespace finess {

__global__ __launch_bounds__(64) void k_step_control(StepCtl* c, unsigned long long* smax, StepCtlParams p)
{
    static_assert(CFL_SPREAD == 64, "one maxima word per lane and direction");
    const int l = threadIdx.x;
    // every load is issued before anything waits on one: the maxima, and
    // the whole control block in lane 0 (decided on a register copy and
    // stored back in one piece)
    unsigned long long bx = smax[l], by = smax[CFL_SPREAD + l];
    StepCtl h;
    if (l == 0) h = *c;
    const unsigned long long halt = __shfl(l == 0 ? h.halt : 0ull, 0, 64);
    if (halt != HALT_NONE) return;
    smax[l] = 0ull;
    smax[CFL_SPREAD + l] = 0ull;
    // bit patterns of non-negative doubles (or +inf) order as the values
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long ox = (unsigned long long)__double_as_longlong(
            __shfl_xor(__longlong_as_double((long long)bx), o, 64));
        const unsigned long long oy = (unsigned long long)__double_as_longlong(
            __shfl_xor(__longlong_as_double((long long)by), o, 64));
        bx = ox > bx ? ox : bx;
        by = oy > by ? oy : by;
    }
    if (l != 0) return;
    const double sx = __longlong_as_double((long long)bx), sy = __longlong_as_double((long long)by);
    const double dt = h.dt, told = h.t;
    h.done = h.done + 1;
    auto finish = [&]() { *c = h; };
    auto error = [&](long long code, double t) {
        h.err = code;
        h.t = t;
        h.halt = HALT_ERROR;
        finish();
    };
    // GetCFL's checks, in its order
    if (dt > p.max_dt) { error(FINESS_ERR_DT, told + dt); return; }
    if (!isfinite(sx) || !isfinite(sy)) { error(FINESS_ERR_NONFINITE, told + dt); return; }
    double cfl = -100.0;
    cfl = rmax(dt * sx / p.dx, cfl);
    if (!p.one_d) cfl = rmax(dt * sy / p.dy, cfl);
    h.cfl = cfl;
    if (cfl > (p.one_d ? 1.0e10 : 1.0e8)) { error(FINESS_ERR_CFL, told + dt); return; }
    double dtn;
    if (cfl > 0.0) {
        dtn = rmin(p.max_dt, dt * p.cfl_target / cfl);
        h.dtmin = rmin(dtn, h.dtmin);
        h.dtmax = rmax(dtn, h.dtmax);
    } else {
        dtn = p.max_dt;
    }
    double t = told;
    if (cfl <= p.cfl_max) {
        t = told + dt;
        h.acc = h.acc + 1;
        if (p.target ? h.acc >= p.target : !(t < p.tend)) {
            h.halt = HALT_DONE;
        } else if (!p.target) {
            h.n_step = h.n_step + 1;
            if (h.n_step > p.nv) { h.dt_next = dtn; error(FINESS_ERR_NV, t); return; }
        }
    } else {
        h.rej = h.rej + 1;
        h.halt = HALT_REJECT;
    }
    h.t = t;
    h.dt_next = dtn;
    h.dt = (!p.target && t + dtn > p.tend) ? p.tend - t : dtn;
    finish();
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
