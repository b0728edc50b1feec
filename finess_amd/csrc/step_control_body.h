// step_control_body.h -- the device-side time-step controller of FinSolveRK
// (lib/2d/FinSolveRK.cpp:77-106 and 318-350, with GetCFL from
// lib/2d/GetCFL.cpp:31-74 and lib/1d/GetCFL.cpp:9-45), run by one full wave
// (lanes l = 0..63) after a step attempt: it reduces the attempt's wave-speed
// maxima (and clears them for the next attempt), forms the CFL number,
// accepts or rejects, sets the next dt and stops the run (halt) at the end,
// at a rejection (the host re-plans from the rejected attempt) or at an
// error.  The arithmetic is the host controller's, operation for operation.
// Callers: k_step_control (its own launch) and the 1D step kernel (wave 0,
// after the step's last stage).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/finess_hip.h"
#include "physics.h"
#include "stage2d_body.h"
#include "step_control.h"

namespace finess {

__device__ __forceinline__ void step_control_wave(StepCtl* c, unsigned long long* smax, const StepCtlParams& p, int l)
{
    // every load is issued before anything waits on one: the maxima (words
    // l, l + 64, ... of each direction), and the whole control block in lane 0
    // (decided on a register copy and stored back in one piece)
    unsigned long long bx = 0ull, by = 0ull;
    for (int k = l; k < CFL_SPREAD; k += 64) {
        const unsigned long long x = __hip_atomic_load(smax + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long y = __hip_atomic_load(smax + CFL_SPREAD + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bx = x > bx ? x : bx;
        by = y > by ? y : by;
    }
    StepCtl h;
    if (l == 0) h = *c;
    const unsigned long long h0 = l == 0 ? h.halt : 0ull;
    const unsigned long long halt = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(h0 >> 32), 0) << 32) |
                                    (unsigned)__builtin_amdgcn_readlane((int)h0, 0);
    if (halt != HALT_NONE) return;
    for (int k = l; k < CFL_SPREAD; k += 64) {
        smax[k] = 0ull;
        smax[CFL_SPREAD + k] = 0ull;
    }
    // the words hold non-negative doubles or +inf (never NaN), whose maximum
    // is the maximum of their bit patterns; wave_max64 returns it in every lane
    bx = (unsigned long long)__double_as_longlong(wave_max64(__longlong_as_double((long long)bx)));
    by = (unsigned long long)__double_as_longlong(wave_max64(__longlong_as_double((long long)by)));
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

} // namespace finess
