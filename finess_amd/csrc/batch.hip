// batch.hip -- kernels of the batched mode (no counterpart in the reference):
// one persistent launch per SSP-RK3 stage covers the tiles of every problem
// of the batch, and one small launch runs every problem's time-step
// controller on the device, so a whole step sequence needs no host
// round trip and can be captured in a HIP graph.
//
// The per-tile work is the fused stage of the single-problem kernel (same
// phase functions: BC-on-load halo, characteristic WENO5-LLF faces, FP64
// stage update), so each problem's result does not depend on how it was
// batched or on the tile shape.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "kernels.h"
#include "stage2d_body.h"

namespace finess {

#ifndef FINESS_BATCH_FY
#define FINESS_BATCH_FY 1
#endif

template <class P, typename R, int TI, int TJ, int STAGE>
__global__ __launch_bounds__(256) void k_batch_stage(BatchArgs b)
{
    constexpr int NT = 256, M = P::MEQN;
    static_assert((TI + 1) * TJ <= NT && TI * (TJ + 1) <= NT, "one face per thread and direction");
    __shared__ StageSmem<P, R, TI, TJ> sm;
    __shared__ double red[2][NT / 64];
#if FINESS_BATCH_FY
    constexpr int NFY = TI * (TJ + 1);
    __shared__ R fy[M][NFY];
#endif
    const int t = threadIdx.x;
    for (int tile = blockIdx.x; tile < b.ntiles; tile += gridDim.x) {
        const BatchProb* pr = b.probs + b.tile_prob[tile];
        if (pr->done) continue;  // uniform over the workgroup
        const int lt = tile - pr->tile0;
        const int ti = lt / pr->ntj, tj = lt - ti * pr->ntj;
        const int i0 = ti * TI, j0 = tj * TJ;
        StageArgs a = {};
        a.mx = pr->mx;
        a.my = pr->my;
        a.plane = pr->plane;
        for (int k = 0; k < 4; k++) a.bc[k] = b.bc[k];
        a.dx = pr->dx;
        a.dy = pr->dy;
        a.idx = 1.0 / pr->dx;
        a.idy = 1.0 / pr->dy;
        a.alpha_scaling = b.alpha_scaling;
        a.eps = b.eps;
        a.power = b.power;
        a.pp = b.pp;
        // SSP-RK3 with A = q^n never written (coefficients of lib/2d/SetRKinfo.cpp:50-64)
        const int A = pr->cur, B = (A + 1) % 3, C = (A + 2) % 3;
        double beta;
        if (STAGE == 1) {
            a.src = pr->buf[A]; a.qa = pr->buf[A]; a.a_is_src = 1; a.out = pr->buf[B];
            a.a1 = 1.0; a.a2 = 0.0; beta = 1.0;
        } else if (STAGE == 2) {
            a.src = pr->buf[B]; a.qa = pr->buf[A]; a.qb = pr->buf[B]; a.b_is_src = 1; a.use_b = 1;
            a.out = pr->buf[C]; a.a1 = 0.75; a.a2 = 0.25; beta = 0.25;
        } else {
            a.src = pr->buf[C]; a.qa = pr->buf[C]; a.a_is_src = 1; a.qb = pr->buf[A]; a.use_b = 1;
            a.out = pr->buf[B]; a.a1 = 2.0e0 / 3.0e0; a.a2 = 1.0e0 / 3.0e0; beta = 2.0e0 / 3.0e0;
        }
        a.bdt = beta * pr->dt;
        double pa[M], pb[M], Lx[M];
        phase_prefetch<P, R, TI, TJ>(t, i0, j0, a, pa, pb);
        phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, sm);
        __syncthreads();
        R ax = R(0), ay = R(0);  // wave-speed maxima in the stencil precision
#if FINESS_BATCH_FY
        // the y fluxes in their own array: no barrier between the face passes
        phase_faces<P, R, WENO_JS5_P2, false, TI, TJ>(t, 0, i0, j0, a, sm, ax);
        phase_faces<P, R, WENO_JS5_P2, false, TI, TJ>(t, 1, i0, j0, a, sm, ay, false, &fy[0][0], NFY);
        __syncthreads();
        phase_lx<P, R, TI, TJ>(t, a, sm, Lx);
        phase_update<P, R, TI, TJ>(t, i0, j0, a, sm, Lx, pa, pb, &fy[0][0], NFY);
#else
        phase_faces<P, R, WENO_JS5_P2, false, TI, TJ>(t, 0, i0, j0, a, sm, ax);
        __syncthreads();
        phase_lx<P, R, TI, TJ>(t, a, sm, Lx);
        __syncthreads();
        phase_faces<P, R, WENO_JS5_P2, false, TI, TJ>(t, 1, i0, j0, a, sm, ay);
        __syncthreads();
        phase_update<P, R, TI, TJ>(t, i0, j0, a, sm, Lx, pa, pb);
#endif
        if (ax != ax) ax = R(__builtin_inf());
        if (ay != ay) ay = R(__builtin_inf());
        if constexpr (sizeof(R) == sizeof(float)) {
            ax = wave_max64_xor(ax);
            ay = wave_max64_xor(ay);
        } else {
            ax = wave_max64(ax);
            ay = wave_max64(ay);
        }
        if ((t & 63) == 0) { red[0][t >> 6] = double(ax); red[1][t >> 6] = double(ay); }
        __syncthreads();
        if (t == 0) {
            double mx_ = red[0][0], my_ = red[1][0];
            for (int k = 1; k < NT / 64; k++) {
                mx_ = red[0][k] > mx_ ? red[0][k] : mx_;
                my_ = red[1][k] > my_ ? red[1][k] : my_;
            }
            const int s = tile % BATCH_SPREAD;
            atomicMax(&pr->smax[s], (unsigned long long)__double_as_longlong(mx_));
            atomicMax(&pr->smax[BATCH_SPREAD + s], (unsigned long long)__double_as_longlong(my_));
        }
        __syncthreads();  // the tile's LDS is reused by the next tile
    }
}

// The a-posteriori controller of lib/2d/FinSolveRK.cpp:77-367 for every
// problem (one thread each): CFL from the step's wave speeds over the
// interior faces (lib/2d/GetCFL.cpp:31-74), accept when cfl <= max_cfl,
// next dt = min(max_dt, dt * desired_cfl / cfl).
__global__ void k_batch_control(BatchArgs b)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.nprob) return;
    BatchProb* pr = b.probs + k;
    if (pr->done) return;
    unsigned long long bx = 0, by = 0;
    for (int s = 0; s < BATCH_SPREAD; s++) {
        bx = pr->smax[s] > bx ? pr->smax[s] : bx;
        by = pr->smax[BATCH_SPREAD + s] > by ? pr->smax[BATCH_SPREAD + s] : by;
        pr->smax[s] = 0ull;
        pr->smax[BATCH_SPREAD + s] = 0ull;
    }
    const double sx = __longlong_as_double((long long)bx), sy = __longlong_as_double((long long)by);
    const double dt = pr->dt;
    int err = 0;
    if (dt > b.max_dt) err = -5;                                    // FINESS_ERR_DT
    else if (!isfinite(sx) || !isfinite(sy)) err = -2;              // FINESS_ERR_NONFINITE
    double cfl = -100.0;
    cfl = fmax(dt * sx / pr->dx, cfl);
    cfl = fmax(dt * sy / pr->dy, cfl);
    if (!err && cfl > 1.0e8) err = -6;                              // FINESS_ERR_CFL
    if (err) {
        pr->err = err;
        pr->done = 1;
        return;
    }
    pr->last_cfl = cfl;
    pr->dt = cfl > 0.0 ? fmin(b.max_dt, dt * b.desired_cfl / cfl) : b.max_dt;
    if (cfl <= b.max_cfl) {
        pr->cur = (pr->cur + 1) % 3;   // the step's result is the first stage's output set
        pr->t = pr->t + dt;
        pr->accepted = pr->accepted + 1;
        if (pr->accepted >= pr->target) pr->done = 1;
    } else {
        pr->rejected = pr->rejected + 1;
    }
}

template <class P, typename R, int SH>
static int launch_shape(int stage, const BatchArgs& b, int grid, hipStream_t st)
{
    constexpr int TI = BATCH_SHAPES[SH][0], TJ = BATCH_SHAPES[SH][1];
    switch (stage) {
    case 1: hipLaunchKernelGGL((k_batch_stage<P, R, TI, TJ, 1>), dim3(grid), dim3(256), 0, st, b); break;
    case 2: hipLaunchKernelGGL((k_batch_stage<P, R, TI, TJ, 2>), dim3(grid), dim3(256), 0, st, b); break;
    case 3: hipLaunchKernelGGL((k_batch_stage<P, R, TI, TJ, 3>), dim3(grid), dim3(256), 0, st, b); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <class P, typename R>
static int launch_pr(int shape, int stage, const BatchArgs& b, int grid, hipStream_t st)
{
    switch (shape) {
    case 0: return launch_shape<P, R, 0>(stage, b, grid, st);
    case 1: return launch_shape<P, R, 1>(stage, b, grid, st);
    case 2: return launch_shape<P, R, 2>(stage, b, grid, st);
    default: return -1;
    }
}

int launch_batch_stage(int physics, int fp32, int shape, int stage, const BatchArgs& b, int grid, hipStream_t st)
{
    if (physics == PHYS_SHALLOW_WATER)
        return fp32 ? launch_pr<ShallowWater2D, float>(shape, stage, b, grid, st)
                    : launch_pr<ShallowWater2D, double>(shape, stage, b, grid, st);
    if (physics == PHYS_EULER)
        return fp32 ? launch_pr<Euler2D, float>(shape, stage, b, grid, st)
                    : launch_pr<Euler2D, double>(shape, stage, b, grid, st);
    return -1;
}

int launch_batch_control(const BatchArgs& b, hipStream_t st)
{
    hipLaunchKernelGGL(k_batch_control, dim3((b.nprob + 255) / 256), dim3(256), 0, st, b);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <class P, typename R, int SH>
static int occupancy(int* blocks)
{
    constexpr int TI = BATCH_SHAPES[SH][0], TJ = BATCH_SHAPES[SH][1];
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_batch_stage<P, R, TI, TJ, 1>, 256, 0) ==
                   hipSuccess
               ? 0
               : -4;
}

template <class P, typename R>
static int occupancy_pr(int shape, int* blocks)
{
    switch (shape) {
    case 0: return occupancy<P, R, 0>(blocks);
    case 1: return occupancy<P, R, 1>(blocks);
    case 2: return occupancy<P, R, 2>(blocks);
    default: return -1;
    }
}

int batch_grid(int physics, int fp32, int shape, int device)
{
    int blocks = 0, rc;
    if (physics == PHYS_SHALLOW_WATER)
        rc = fp32 ? occupancy_pr<ShallowWater2D, float>(shape, &blocks) : occupancy_pr<ShallowWater2D, double>(shape, &blocks);
    else
        rc = fp32 ? occupancy_pr<Euler2D, float>(shape, &blocks) : occupancy_pr<Euler2D, double>(shape, &blocks);
    int cus = 0;
    if (rc || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return -1;
    return (blocks > 0 ? blocks : 1) * cus;
}

} // namespace finess
