// stage1d.hip -- the 1D path: lib/1d/ConstructL.cpp:16-224 with the
// advection callbacks of apps/1d/advection/lib (f = aux(1) q, identity
// eigenvectors, s1 = min(a_l, (a_l+a_r)/2), s2 = max(a_r, (a_l+a_r)/2)),
// UpdateSoln (lib/1d/UpdateSoln.cpp:17-54) and the periodic halo
// (apps/1d/advection/smooth_example/SetBndValues.cpp:6-63).
//
// 1D grids are small (the advection case has 200 cells): up to
// small1d_cells<WK>() cells a stage is one workgroup (faces, barrier, cells),
// and a time step's stages run in one launch of that workgroup; beyond it
// each stage is two plain launches, one thread per face, then one thread
// per cell.
#include <hip/hip_runtime.h>
#include <type_traits>
#include "kernels.h"
#include "stage1d_body.h"
#include "step_control_body.h"

namespace finess {

template <int WK>
__global__ __launch_bounds__(256) void k_faces1d(Stage1DArgs a)
{
    if (halted1d(a)) return;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    double alpha = 0.0;
    if (f <= a.mx) alpha = face1d_any<WK>(f, a);
    if (a.smax_slot) {
        if (alpha != alpha) alpha = __builtin_inf();
        // GetCFL (lib/1d/GetCFL.cpp:30-33) reads faces 1..mx (1-based), i.e. f < mx
        if (f < a.mx)
            atomicMax(a.smax_slot + blockIdx.x % CFL_SPREAD, (unsigned long long)__double_as_longlong(alpha));
    }
}

__global__ __launch_bounds__(256) void k_update1d(Stage1DArgs a)
{
    if (halted1d(a)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.mx) update1d(i, a);
}

// GlobalWaveSpd (lib/1d/GlobalWaveSpd.cpp): max |aux(1)| over the interior
// (the advection SetWaveSpd with both sides equal), one workgroup.
__global__ __launch_bounds__(256) void k_global_speed1d(Stage1DArgs a)
{
    if (halted1d(a)) return;
    double v = 0.0;
    for (int i = threadIdx.x; i < a.mx; i += 256) {
        const double s = fabs(a.aux[idx1d(a, i)]);
        v = s != s ? __builtin_inf() : (v > s ? v : s);
    }
    v = wave_max64(v);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int k = 1; k < 4; k++) m = red[k] > m ? red[k] : m;
        *a.galpha = m;
    }
}

// Grids up to small1d_cells<WK>() cells run the whole stage in one
// workgroup: the faces, a workgroup barrier (which orders the face fluxes
// written to global memory before they are read), then the cells.  One
// launch instead of two, but one CU: the limits are the measured crossovers
// with the two-launch path, which takes 31-34 us per SSP-RK3 step (with the
// controller) at any size up to 8192 cells.  WENO5: 21.5 against 30.6 us
// at 1024 cells, 34.6 against 31.0 at 2048; WENO11: 31.6 against 33.2 at
// 512, 44.2 against 33.6 at 1024.  Orders 7 and 9 lie between.
template <int WK>
constexpr int small1d_cells()
{
#ifdef FINESS_SMALL1D
    return FINESS_SMALL1D;
#else
    return WK == WENO_JS11                                       ? 512
           : (WK == WENO_JS9 || WK == WENO_Z9 || WK == WENO_FD9) ? 768
           : (WK == WENO_JS7 || WK == WENO_Z7 || WK == WENO_FD7) ? 1024
                                                                 : 1536;
#endif
}

constexpr int SMALL1D_NT = 512;  // 2 waves per SIMD: the wide orders fit without spilling

#ifndef FINESS_PAIR1D
#define FINESS_PAIR1D 1
#endif

// One stage on a single workgroup: faces, barrier, cells, the CFL maximum.
template <int WK>
__device__ __forceinline__ void stage1d_small_body(const Stage1DArgs& a, double* red)
{
    double amax = 0.0;
    bool paired = false;
    if constexpr (WK <= WENO_FD5) paired = FINESS_PAIR1D && 2 * (a.mx + 1) <= SMALL1D_NT;
    if (paired) {
        // a grid of at most SMALL1D_NT / 2 faces: two lanes per face, one
        // split part each, so that a face's two reconstructions run side by
        // side; the even lane adds them in face1d's order
        const int f = threadIdx.x >> 1, part = threadIdx.x & 1;
        if (f <= a.mx) {
            double alpha;
            const double w = face1d_part<WK <= WENO_FD5 ? WK : WENO_JS5_P2>(f, a, part, alpha);
            // the partner lane's part (quad_perm [1,0,3,2]: both lanes of a
            // pair take this branch together)
            const long long wb = __double_as_longlong(w);
            const int lo = dpp_move<0xB1>((int)wb), hi = dpp_move<0xB1>((int)(wb >> 32));
            const double wo = __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
            if (!part) {
                a.fhat[f] = w + wo;
                if (a.smax_faces) a.smax_faces[f] = alpha;
                if (alpha != alpha) alpha = __builtin_inf();
                // GetCFL (lib/1d/GetCFL.cpp:30-33) reads faces 1..mx (1-based), i.e. f < mx
                if (f < a.mx) amax = alpha > amax ? alpha : amax;
            }
        }
    } else {
        for (int f = threadIdx.x; f <= a.mx; f += SMALL1D_NT) {
            double alpha = face1d_any<WK>(f, a);
            if (alpha != alpha) alpha = __builtin_inf();
            // GetCFL (lib/1d/GetCFL.cpp:30-33) reads faces 1..mx (1-based), i.e. f < mx
            if (f < a.mx) amax = alpha > amax ? alpha : amax;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.mx; i += SMALL1D_NT) update1d(i, a);
    if (!a.smax_slot) return;
    amax = wave_max64(amax);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = amax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int k = 1; k < SMALL1D_NT / 64; k++) m = red[k] > m ? red[k] : m;
        atomicMax(a.smax_slot, (unsigned long long)__double_as_longlong(m));
    }
}

template <int WK>
__global__ __launch_bounds__(SMALL1D_NT) void k_stage1d_small(Stage1DArgs a)
{
    if (halted1d(a)) return;
    __shared__ double red[SMALL1D_NT / 64];
    stage1d_small_body<WK>(a, red);
}

// Every stage of one time step in one workgroup: the barrier after each stage
// orders its cell writes (and its reads of the shared face fluxes) before the
// next stage's faces.  One launch per step instead of one per stage.
template <int WK>
__global__ __launch_bounds__(SMALL1D_NT) void k_step1d_small(Stage1DStep s)
{
    if (halted1d(s.stage[0])) return;
    __shared__ double red[SMALL1D_NT / 64];
    for (int k = 0; k < s.n; k++) {
        stage1d_small_body<WK>(s.stage[k], red);
        __syncthreads();
    }
    // wave 0 runs the controller on the last stage's maxima; the barrier
    // above has ordered the stages' reads of dt before the block changes
    if (s.ctl && threadIdx.x < 64) step_control_wave(s.ctl, s.stage[s.n - 1].smax_slot, s.cp, threadIdx.x);
}

// A run of steps (Stage1DRun).  Every thread takes the controller's decision
// from the block after the barrier that follows it: a step that raised the
// accepted count rotates the plane sets (an nv stop happens after its step
// is accepted, as on the host); a rejection is cleared here and retried.
//
// The run works in LDS (run1d_lds_bytes): the three plane sets, the face
// fluxes, the advection speeds, the CFL words and the controller block.  The stage code reaches
// them through the same pointers it takes for global memory (generic
// addresses), the state at entry and the block are read once, and the state
// at exit, the CFL words and the block are written back once: each stage
// then waits on LDS instead of on a round trip through L2 after every
// barrier.
#ifndef FINESS_RUN1D_LDS
#define FINESS_RUN1D_LDS 1
#endif
template <int WK>
__global__ __launch_bounds__(SMALL1D_NT) void k_run1d_small(Stage1DRun r)
{
    __shared__ double red[SMALL1D_NT / 64];
    int cur = r.cur;
#if FINESS_RUN1D_LDS
    extern __shared__ double lds1d[];
    __shared__ StepCtl lctl;
    __shared__ unsigned long long lsmax[2 * CFL_SPREAD];
    const int mx = r.base.mx;
    double* const buf[3] = {lds1d, lds1d + mx, lds1d + 2 * mx};
    double* const laux = lds1d + 4 * mx + 1;
    for (int i = threadIdx.x; i < mx; i += SMALL1D_NT) {
        buf[cur][i] = r.buf[cur][i];
        laux[i] = r.base.aux[i];
    }
    for (int k = threadIdx.x; k < 2 * CFL_SPREAD; k += SMALL1D_NT) lsmax[k] = r.smax_slot[k];
    if (threadIdx.x == 0) lctl = *r.ctl;
    __syncthreads();
    StepCtl* const ctl = &lctl;
    unsigned long long* const smax = lsmax;
    Stage1DArgs base = r.base;
    base.fhat = lds1d + 3 * mx;
    base.aux = laux;
    base.dtp = &lctl.dt;
    base.halt = &lctl.halt;
#else
    double* const* const buf = r.buf;
    StepCtl* const ctl = r.ctl;
    unsigned long long* const smax = r.smax_slot;
    const Stage1DArgs base = r.base;
#endif
    for (int att = 0; att < r.max_attempts; att++) {
        if (__hip_atomic_load(&ctl->halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != HALT_NONE) break;
        const long long acc0 = __hip_atomic_load(&ctl->acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int k = 0; k < r.n; k++) {
            const SspStage g = ssp_stage(r.n, k);
            Stage1DArgs a = base;
            a.src = buf[(cur + g.src) % 3];
            a.qa = buf[(cur + g.qa) % 3];
            a.qb = g.qb >= 0 ? buf[(cur + g.qb) % 3] : nullptr;
            a.use_b = g.qb >= 0 && g.a2 != 0.0;
            a.out = buf[(cur + g.out) % 3];
            a.a1 = g.a1;
            a.a2 = g.a2;
            a.beta = g.beta;
            a.smax_slot = k == r.n - 1 ? smax : nullptr;
            stage1d_small_body<WK>(a, red);
            __syncthreads();
        }
        if (threadIdx.x < 64) step_control_wave(ctl, smax, r.cp, threadIdx.x);
        __syncthreads();
        const unsigned long long h = __hip_atomic_load(&ctl->halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_load(&ctl->acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != acc0)
            cur = (cur + ssp_result(r.n)) % 3;
        if (h == HALT_REJECT) {
            __syncthreads();  // every thread has read the halt word before it is cleared
            if (threadIdx.x == 0) __hip_atomic_store(&ctl->halt, (unsigned long long)HALT_NONE, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            continue;
        }
        if (h != HALT_NONE) break;
    }
#if FINESS_RUN1D_LDS
    for (int i = threadIdx.x; i < mx; i += SMALL1D_NT) r.buf[cur][i] = buf[cur][i];
    for (int k = threadIdx.x; k < 2 * CFL_SPREAD; k += SMALL1D_NT) r.smax_slot[k] = lsmax[k];
    if (threadIdx.x == 0) {
        lctl.buf = cur;
        *r.ctl = lctl;
    }
#else
    if (threadIdx.x == 0) r.ctl->buf = cur;
#endif
}

// Dynamic LDS of the run kernel: three plane sets of mx cells, mx + 1 face
// fluxes and mx advection speeds (61 KB at the largest single-workgroup
// grid, 1536 cells).
static size_t run1d_lds_bytes(int mx)
{
    return FINESS_RUN1D_LDS ? (size_t)(5 * mx + 1) * sizeof(double) : 0;
}

template <int WK>
static int launch_wk(const Stage1DArgs& a, hipStream_t st)
{
    if (a.galpha) hipLaunchKernelGGL(k_global_speed1d, dim3(1), dim3(256), 0, st, a);
    if (a.mx <= small1d_cells<WK>()) {
        hipLaunchKernelGGL(k_stage1d_small<WK>, dim3(1), dim3(SMALL1D_NT), 0, st, a);
        return hipGetLastError() == hipSuccess ? 0 : -4;
    }
    hipLaunchKernelGGL(k_faces1d<WK>, dim3((a.mx + 1 + 255) / 256), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_update1d, dim3((a.mx + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// GlobalWaveSpd in 1D reads aux only, which no stage changes: one launch
// serves all stages of a step.
template <int WK>
static int launch_step_wk(const Stage1DStep& s, hipStream_t st)
{
    if (s.stage[0].mx > small1d_cells<WK>()) {
        for (int k = 0; k < s.n; k++)
            if (const int rc = launch_wk<WK>(s.stage[k], st)) return rc;
        return s.ctl ? launch_step_control(s.ctl, s.stage[s.n - 1].smax_slot, s.cp, st) : 0;
    }
    if (s.stage[0].galpha) hipLaunchKernelGGL(k_global_speed1d, dim3(1), dim3(256), 0, st, s.stage[0]);
    hipLaunchKernelGGL(k_step1d_small<WK>, dim3(1), dim3(SMALL1D_NT), 0, st, s);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// Time-integrated fluxes of the 1D Lax-Wendroff step on the extended cells.
__global__ __launch_bounds__(256) void k_lxw_flux1d(Stage1DArgs a, double dt, int time_order, double* F)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < a.mx + 6) F[e] = integrated_f1d(e - 3, a, dt, time_order);
}

int launch_lxw_flux1d(const Stage1DArgs& a, double dt, int time_order, double* F, hipStream_t st)
{
    hipLaunchKernelGGL(k_lxw_flux1d, dim3((a.mx + 6 + 255) / 256), dim3(256), 0, st, a, dt, time_order, F);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <class F>
static int with_kind(int wk, F f)
{
    switch (wk) {
    case WENO_JS5_P2: return f(std::integral_constant<int, WENO_JS5_P2>());
    case WENO_JS5_POW: return f(std::integral_constant<int, WENO_JS5_POW>());
    case WENO_Z5: return f(std::integral_constant<int, WENO_Z5>());
    case WENO_FD5: return f(std::integral_constant<int, WENO_FD5>());
    case WENO_JS7: return f(std::integral_constant<int, WENO_JS7>());
    case WENO_JS9: return f(std::integral_constant<int, WENO_JS9>());
    case WENO_JS11: return f(std::integral_constant<int, WENO_JS11>());
    case WENO_Z7: return f(std::integral_constant<int, WENO_Z7>());
    case WENO_Z9: return f(std::integral_constant<int, WENO_Z9>());
    case WENO_FD7: return f(std::integral_constant<int, WENO_FD7>());
    case WENO_FD9: return f(std::integral_constant<int, WENO_FD9>());
    case WENO_O1: return f(std::integral_constant<int, WENO_O1>());
    default: return -1;
    }
}

int launch_stage1d(int wk, const Stage1DArgs& a, hipStream_t st)
{
    return with_kind(wk, [&](auto k) { return launch_wk<decltype(k)::value>(a, st); });
}

int launch_run1d(int wk, const Stage1DRun& r, hipStream_t st)
{
    if (r.n < 1 || r.n > 3 || !run1d_fits(wk, r.base.mx)) return -1;
    if (r.base.galpha) hipLaunchKernelGGL(k_global_speed1d, dim3(1), dim3(256), 0, st, r.base);
    return with_kind(wk, [&](auto k) {
        hipLaunchKernelGGL(k_run1d_small<decltype(k)::value>, dim3(1), dim3(SMALL1D_NT), run1d_lds_bytes(r.base.mx),
                           st, r);
        return hipGetLastError() == hipSuccess ? 0 : -4;
    });
}

bool run1d_fits(int wk, int mx)
{
    // (and the run's LDS within the 64 KB a launch gets without an attribute)
    return run1d_lds_bytes(mx) + 2048 <= 65536 &&
           with_kind(wk, [&](auto k) { return mx <= small1d_cells<decltype(k)::value>() ? 1 : 0; }) == 1;
}

int launch_step1d(int wk, const Stage1DStep& s, hipStream_t st)
{
    if (s.n < 1 || s.n > 3) return -1;
    return with_kind(wk, [&](auto k) { return launch_step_wk<decltype(k)::value>(s, st); });
}

} // namespace finess
