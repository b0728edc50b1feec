// step_control.h -- the time-step controller run on the device, one small
// launch after each step attempt, so that a run of steps is enqueued without
// a host round trip per step.  The host reads the block back once per chunk
// of steps.
#pragma once

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
};

struct StepCtlParams {
    double tend, max_dt, cfl_max, cfl_target, dx, dy;
    int one_d;          // 1D GetCFL (lib/1d/GetCFL.cpp): x speeds only, bound 1e10
    long long nv;
    long long target;   // accepted steps to make; 0: run until t reaches tend
};

} // namespace finess
