// batch.h -- device-side description of a batch of independent 2D problems
// (no counterpart in the reference): one table entry per problem, with its
// grid, its three state plane sets and its own time-step controller state.
#pragma once
#include <hip/hip_runtime.h>

#include "physics.h"

namespace finess {

constexpr int BATCH_SPREAD = 8;  // words per wave-speed maximum of one problem

struct BatchProb {
    int mx, my;
    int ntj;                      // tiles along j for the current tile shape
    int tile0;                    // first entry of the problem in the tile map
    long long plane;              // mx * my
    double* buf[3];               // state plane sets; buf[cur] holds q^n
    unsigned long long* smax;     // [2][BATCH_SPREAD] wave-speed maxima of the step
    double dx, dy;
    double t, dt;                 // time of q^n, dt of the next step
    long long accepted, rejected, target;
    int cur, done, err;           // err: a FINESS_ERR_* status, 0 if none
    double last_cfl;
};

struct BatchArgs {
    BatchProb* probs;
    const int* tile_prob;         // tile -> problem
    int ntiles, nprob;
    int bc[4];
    PhysParams pp;
    double eps, power, alpha_scaling;
    double desired_cfl, max_cfl, max_dt;
};

// shapes the autotuner chooses from: (TI + 1) * TJ <= 256 and TI * (TJ + 1) <= 256
constexpr int BATCH_NSHAPES = 3;
constexpr int BATCH_SHAPES[BATCH_NSHAPES][2] = {{7, 32}, {15, 16}, {3, 64}};

// stage: 1..3 of SSP-RK3.  grid: persistent workgroups.
int launch_batch_stage(int physics, int fp32, int shape, int stage, const BatchArgs& b, int grid, hipStream_t st);
int launch_batch_control(const BatchArgs& b, hipStream_t st);
// persistent grid size for (physics, fp32, shape): resident workgroups on the device
int batch_grid(int physics, int fp32, int shape, int device);

} // namespace finess
