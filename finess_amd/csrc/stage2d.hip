// stage2d.hip -- gfx950 kernels of the 2D path: the fused stage kernel and
// the small helpers around it (layout conversion, CFL-slot reset).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "stage2d_body.h"

namespace finess {

// Reduces two per-thread maxima over the workgroup and folds them into
// slot[b] and slot[CFL_SPREAD + b], b = block index mod CFL_SPREAD
// (non-negative doubles compare like their bit patterns).  The per-thread
// maxima are FP64 or FP32 values; the wave reduction runs in their type.
template <int NT, typename A>
__device__ __forceinline__ void block_max2(A ax, A ay, unsigned long long* slot)
{
    __shared__ double red[2][NT / 64];
    ax = wave_max64(ax);
    ay = wave_max64(ay);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { red[0][wid] = double(ax); red[1][wid] = double(ay); }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx_ = red[0][0], my_ = red[1][0];
        for (int w = 1; w < NT / 64; w++) {
            mx_ = red[0][w] > mx_ ? red[0][w] : mx_;
            my_ = red[1][w] > my_ ? red[1][w] : my_;
        }
        const int b = blockIdx.x % CFL_SPREAD;
        atomicMax(&slot[b], (unsigned long long)__double_as_longlong(mx_));
        atomicMax(&slot[CFL_SPREAD + b], (unsigned long long)__double_as_longlong(my_));
    }
}

// The minimum waves per SIMD asked of the register allocator.  The FP64
// kernel compiles to 156-166 VGPRs without spills, so 3 waves per SIMD run,
// as many as the tile's LDS (about 40 KB, 3 workgroups per CU) allows.
#ifndef FINESS_WPE
#define FINESS_WPE 2
#endif
constexpr int WPE = FINESS_WPE;
// The y-face fluxes in an LDS array of their own rather than in the x
// fluxes' array after them (two barriers fewer, 10 KB more LDS in FP64, the
// same 3 workgroups per CU): 1.6% less time for the FP64 Runge-Kutta stage
// and 1% for the FP32 Lax-Wendroff stage on MI355X, but 3% more for the FP64
// Lax-Wendroff stage, which keeps the shared array.
#ifndef FINESS_STAGE_FY
#define FINESS_STAGE_FY 1
#endif
template <typename R, bool LXW>
constexpr bool SEPARATE_FY = FINESS_STAGE_FY && !(LXW && sizeof(R) == sizeof(double));

template <class P, typename R, int WK, bool LXW, int TI, int TJ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_stage2d(StageArgs a)
{
    constexpr int NT = 256;
    static_assert((TI + 1) * TJ <= NT && TI * (TJ + 1) <= NT, "one face per thread and direction");
    __shared__ StageSmem<P, R, TI, TJ> sm;
    if (stage_halted(a)) return;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int ti = tile / a.ntj, tj = tile - ti * a.ntj;
    const int i0 = ti * TI, j0 = tj * TJ;
    const int t = threadIdx.x;
    double pa[P::MEQN], pb[P::MEQN];
    // a tile whose cells all have rho w = 0 (every 2D Euler problem that
    // starts without it), and on the Lax-Wendroff path zero time-integrated
    // fluxes of it, skips that field's reconstructions; not in the FP32
    // Lax-Wendroff form, whose stage measured 3% slower with the skip
    constexpr bool SKIPPABLE = !(LXW && sizeof(R) == 4);
    R f3[3] = {R(0), R(0), R(0)};
    if constexpr (LXW && SKIPPABLE) lxw_fetch3<P, R, TI, TJ>(t, i0, j0, a, f3);
    phase_prefetch<P, R, TI, TJ>(t, i0, j0, a, pa, pb);
    bool nz = phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, sm);
    if constexpr (LXW) nz = nz || f3[0] != R(0) || f3[1] != R(0) || f3[2] != R(0);
    const bool skip3 = !__syncthreads_or(nz) && SKIPPABLE;
    R ax = R(0), ay = R(0);  // wave-speed maxima in the stencil precision
    double Lx[P::MEQN];
    if constexpr (SEPARATE_FY<R, LXW>) {
        // the y fluxes in an array of their own: both face passes run
        // without a barrier between them, and Lx is formed after them
        constexpr int NFY = TI * (TJ + 1);
        __shared__ R fy[P::MEQN][NFY];
        phase_faces<P, R, WK, LXW, TI, TJ>(t, 0, i0, j0, a, sm, ax, skip3);
        phase_faces<P, R, WK, LXW, TI, TJ>(t, 1, i0, j0, a, sm, ay, skip3, &fy[0][0], NFY);
        __syncthreads();
        phase_lx<P, R, TI, TJ>(t, a, sm, Lx);
        phase_update<P, R, TI, TJ>(t, i0, j0, a, sm, Lx, pa, pb, &fy[0][0], NFY);
    } else {
        phase_faces<P, R, WK, LXW, TI, TJ>(t, 0, i0, j0, a, sm, ax, skip3);
        __syncthreads();
        phase_lx<P, R, TI, TJ>(t, a, sm, Lx);
        __syncthreads();
        phase_faces<P, R, WK, LXW, TI, TJ>(t, 1, i0, j0, a, sm, ay, skip3);
        __syncthreads();
        phase_update<P, R, TI, TJ>(t, i0, j0, a, sm, Lx, pa, pb);
    }
    if (a.smax) {
        // a NaN speed must not vanish in the max: report it as +inf
        if (ax != ax) ax = R(__builtin_inf());
        if (ay != ay) ay = R(__builtin_inf());
        block_max2<NT>(ax, ay, a.smax);
    }
}

template <class P, typename R, int WK, bool LXW, int TI, int TJ>
static int launch_stage(const StageArgs& a0, hipStream_t st)
{
    StageArgs a = a0;
    const int nti = (a.mx + TI - 1) / TI, ntj = (a.my + TJ - 1) / TJ;
    a.ntj = ntj;
    hipLaunchKernelGGL((k_stage2d<P, R, WK, LXW, TI, TJ>), dim3(nti * ntj), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// The tile of the fused stage kernel (cells in i, cells in j; at most 256
// faces of each direction).  15 x 16 computes 511 faces and 426 cell
// records for 240 cells, 7 x 32 487 faces and 458 records for 224: 4%
// less time per FP64 stage on MI355X (1024^2 Euler).
#ifndef FINESS_STAGE_TI
#define FINESS_STAGE_TI 15
#define FINESS_STAGE_TJ 16
#endif
constexpr int STI = FINESS_STAGE_TI, STJ = FINESS_STAGE_TJ;
// the tile of the FP32 stencil form (15 x 16: 2% less time per stage than
// 7 x 32 for the Runge-Kutta form; the same for Lax-Wendroff)
#ifndef FINESS_STAGE32_TI
#define FINESS_STAGE32_TI 15
#define FINESS_STAGE32_TJ 16
#endif
constexpr int STI32 = FINESS_STAGE32_TI, STJ32 = FINESS_STAGE32_TJ;

template <class P, typename R, bool LXW>
static int dispatch_weno(int wk, const StageArgs& a, hipStream_t st)
{
    switch (wk) {
    case WENO_JS5_P2: return launch_stage<P, R, WENO_JS5_P2, LXW, STI, STJ>(a, st);
    case WENO_FD5: return launch_stage<P, R, WENO_FD5, LXW, STI, STJ>(a, st);
    case WENO_JS5_POW: return launch_stage<P, R, WENO_JS5_POW, LXW, STI, STJ>(a, st);
    case WENO_Z5: return launch_stage<P, R, WENO_Z5, LXW, STI, STJ>(a, st);
    default: return -1;
    }
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v);

// GlobalWaveSpd (lib/2d/GlobalWaveSpd.cpp:20-80): max over the interior
// cells of the SetWaveSpd speed with both sides equal to the cell, per
// direction, folded into gspread[dir][block % CFL_SPREAD].
template <class P>
__global__ __launch_bounds__(256) void k_global_speed(StageArgs a)
{
    if (stage_halted(a)) return;
    double ax = 0.0, ay = 0.0;
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < a.plane; c += (long long)gridDim.x * 256) {
        const int i = (int)(c / a.my), j = (int)(c - (long long)i * a.my);
        double q[P::MEQN], rec[P::NV], qx[P::MEQN], qy[P::MEQN];
        load_state<P>(a, i, j, q);
        P::cell_record(q, a.pp, rec);
        P::state(rec, 0, qx);
        P::state(rec, 1, qy);
        const double sx = P::wave_speed(qx, qx, rec, rec, a.pp), sy = P::wave_speed(qy, qy, rec, rec, a.pp);
        // Max(alpha, s) as the reference folds it; a NaN speed must not vanish
        ax = sx != sx ? __builtin_inf() : (ax > sx ? ax : sx);
        ay = sy != sy ? __builtin_inf() : (ay > sy ? ay : sy);
    }
    block_max2<256>(ax, ay, a.gspread);
}

// Folds the spread words into the two speeds (one workgroup).
__global__ __launch_bounds__(64) void k_global_speed_final(StageArgs a)
{
    if (stage_halted(a)) return;
    const int t = threadIdx.x;
    const double vx = wave_max64(spread_lane_max(a.gspread, t));
    const double vy = wave_max64(spread_lane_max(a.gspread + CFL_SPREAD, t));
    if (t == 0) {
        a.galpha[0] = vx;
        a.galpha[1] = vy;
    }
}

static int launch_global_speed(int physics, const StageArgs& a, hipStream_t st)
{
    if (launch_fill_u64(a.gspread, 2 * CFL_SPREAD, 0ull, st)) return -4;
    const unsigned nb = (unsigned)std::min<long long>(1024, (a.plane + 255) / 256);
    if (physics == PHYS_EULER) hipLaunchKernelGGL(k_global_speed<Euler2D>, dim3(nb), dim3(256), 0, st, a);
    else if (physics == PHYS_SHALLOW_WATER) hipLaunchKernelGGL(k_global_speed<ShallowWater2D>, dim3(nb), dim3(256), 0, st, a);
    else return -1;
    hipLaunchKernelGGL(k_global_speed_final, dim3(1), dim3(64), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_stage2d(int physics, int fp32, int wk, int lxw, const StageArgs& a, hipStream_t st)
{
    if (a.galpha) {
        if (!a.gspread) return -1;
        const int rc = launch_global_speed(physics, a, st);
        if (rc) return rc;
    }
    if (wk >= WENO_JS7) return (fp32 || lxw) ? -1 : launch_stage2d_wide(physics, wk, a, st);
    if (physics == PHYS_EULER) {
        if (!fp32) return lxw ? dispatch_weno<Euler2D, double, true>(wk, a, st)
                              : dispatch_weno<Euler2D, double, false>(wk, a, st);
        if (wk != WENO_JS5_P2) return -1;
        return lxw ? launch_stage<Euler2D, float, WENO_JS5_P2, true, STI32, STJ32>(a, st)
                   : launch_stage<Euler2D, float, WENO_JS5_P2, false, STI32, STJ32>(a, st);
    }
    if (physics == PHYS_SHALLOW_WATER) {
        // the Lax-Wendroff form of the stage serves RK4-PIF's final stage
        if (!fp32) return lxw ? dispatch_weno<ShallowWater2D, double, true>(wk, a, st)
                              : dispatch_weno<ShallowWater2D, double, false>(wk, a, st);
        if (lxw) return -1;
        if (wk != WENO_JS5_P2) return -1;
        return launch_stage<ShallowWater2D, float, WENO_JS5_P2, false, STI32, STJ32>(a, st);
    }
    return -1;
}

// ---- helpers ------------------------------------------------------------

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// Reference layout (i slowest, m fastest, mbc ghosts) -> field planes.
// ndims 2: src index ((i+g)*(my+2g) + (j+g))*meqn + m for 0-based interior
// (i, j); dst planes m*(mx+2*dg)*(my+2*dg) + (i+dg)*(my+2dg) + (j+dg) where dg
// = 0 copies the interior only and dg = g copies ghosts too.
__global__ void k_ref_to_planes(const double* __restrict__ src, double* __restrict__ dst, int mx, int my,
                                int meqn, int g, int dg)
{
    const long long n = (long long)(mx + 2 * dg) * (my + 2 * dg);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int ld = my + 2 * dg;
    const int ii = (int)(k / ld), jj = (int)(k - (long long)ii * ld);
    const int i = ii - dg, j = jj - dg;
    const long long s = ((long long)(i + g) * (my + 2 * g) + (j + g)) * meqn;
    for (int m = 0; m < meqn; m++) dst[m * n + k] = src[s + m];
}

int launch_ref_to_planes(const double* src, double* dst, int mx, int my, int meqn, int g, int dg, hipStream_t st)
{
    const long long n = (long long)(mx + 2 * dg) * (my + 2 * dg);
    hipLaunchKernelGGL(k_ref_to_planes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, mx, my, meqn,
                       g, dg);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// Field planes (interior) -> reference layout with every ghost (corners too)
// filled through the boundary map.
__global__ void k_planes_to_ref(const double* __restrict__ src, double* __restrict__ dst, int mx, int my, int meqn,
                                int g, int bc0, int bc1, int bc2, int bc3, int nx_mom, int ny_mom)
{
    const long long n = (long long)(mx + 2 * g) * (my + 2 * g);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int ld = my + 2 * g;
    const int i = (int)(k / ld) - g, j = (int)(k - (long long)(k / ld) * ld) - g;
    int fi, fj;
    const int si = bc_map(bc0, bc1, i, mx, fi);
    const int sj = bc_map(bc2, bc3, j, my, fj);
    const long long plane = (long long)mx * my;
    for (int m = 0; m < meqn; m++) {
        double v = src[m * plane + (long long)si * my + sj];
        if ((fi && m == nx_mom) || (fj && m == ny_mom)) v = -v;
        dst[k * meqn + m] = v;
    }
}

int launch_planes_to_ref(const double* src, double* dst, int mx, int my, int meqn, int g, const int bc[4],
                         int nx_mom, int ny_mom, hipStream_t st)
{
    const long long n = (long long)(mx + 2 * g) * (my + 2 * g);
    hipLaunchKernelGGL(k_planes_to_ref, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, mx, my, meqn,
                       g, bc[0], bc[1], bc[2], bc[3], nx_mom, ny_mom);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// Edge strips exchanged with a host SetBndValues (BC_HOST sides), packed side
// after side: left and right hold meqn x g x my values (field, layer, j),
// bottom and top meqn x g x mx (field, layer, i).  Layer l is the l-th cell
// from the boundary: inward for the interior ring the host reads, outward for
// the ghosts it returns.
__host__ __device__ __forceinline__ void edge_slot(long long k, int mx, int my, int meqn, int g, int& side, int& m,
                                                   int& l, int& along)
{
    const long long nx = (long long)meqn * g * my, ny = (long long)meqn * g * mx;
    long long r;
    if (k < 2 * nx) {
        side = (int)(k / nx);
        r = k - side * nx;
    } else {
        r = k - 2 * nx;
        side = 2 + (int)(r / ny);
        r -= (side - 2) * ny;
    }
    const int n = side < 2 ? my : mx;
    m = (int)(r / ((long long)g * n));
    r -= (long long)m * g * n;
    l = (int)(r / n);
    along = (int)(r - (long long)l * n);
}

__global__ void k_ring_pack(const double* __restrict__ src, double* __restrict__ ring, int mx, int my, int meqn, int g)
{
    const long long n = 2ll * meqn * g * (mx + my);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int side, m, l, a;
    edge_slot(k, mx, my, meqn, g, side, m, l, a);
    const int i = side == 0 ? l : side == 1 ? mx - 1 - l : a;
    const int j = side == 2 ? l : side == 3 ? my - 1 - l : a;
    ring[k] = src[(long long)m * mx * my + (long long)i * my + j];
}

int launch_ring_pack(const double* src, double* ring, int mx, int my, int meqn, int g, hipStream_t st)
{
    const long long n = 2ll * meqn * g * (mx + my);
    hipLaunchKernelGGL(k_ring_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, ring, mx, my, meqn, g);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// Field planes with g ghosts (plane (mx+2g) x (my+2g), the stage kernels'
// ghosted source): the interior from src, the edge ghosts of BC_HOST sides
// from the packed strips, every other ghost through the boundary map.
__global__ void k_ghost_fill(const double* __restrict__ src, const double* __restrict__ ghost, double* __restrict__ dst,
                             int mx, int my, int meqn, int g, int bc0, int bc1, int bc2, int bc3, int nx_mom, int ny_mom)
{
    const int ld = my + 2 * g;
    const long long pl = (long long)(mx + 2 * g) * ld;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= pl) return;
    const int i = (int)(k / ld) - g, j = (int)(k - (long long)(k / ld) * ld) - g;
    const bool oi = i < 0 || i >= mx, oj = j < 0 || j >= my;
    int side = -1, l = 0, a = 0;
    if (oi && !oj) {
        side = i < 0 ? 0 : 1;
        l = i < 0 ? -1 - i : i - mx;
        a = j;
    } else if (oj && !oi) {
        side = j < 0 ? 2 : 3;
        l = j < 0 ? -1 - j : j - my;
        a = i;
    }
    const int bcs = side == 0 ? bc0 : side == 1 ? bc1 : side == 2 ? bc2 : bc3;
    if (side >= 0 && bcs == BC_HOST) {
        const long long nx = (long long)meqn * g * my, ny = (long long)meqn * g * mx;
        const long long off = side < 2 ? side * nx : 2 * nx + (side - 2) * ny;
        const int n = side < 2 ? my : mx;
        for (int m = 0; m < meqn; m++) dst[m * pl + k] = ghost[off + ((long long)m * g + l) * n + a];
        return;
    }
    int fi, fj;
    const int si = bc_map(bc0, bc1, i, mx, fi);
    const int sj = bc_map(bc2, bc3, j, my, fj);
    const long long plane = (long long)mx * my;
    for (int m = 0; m < meqn; m++) {
        double v = src[m * plane + (long long)si * my + sj];
        if ((fi && m == nx_mom) || (fj && m == ny_mom)) v = -v;
        dst[m * pl + k] = v;
    }
}

int launch_ghost_fill(const double* src, const double* ghost, double* dst, int mx, int my, int meqn, int g,
                      const int bc[4], int nx_mom, int ny_mom, hipStream_t st)
{
    const long long n = (long long)(mx + 2 * g) * (my + 2 * g);
    hipLaunchKernelGGL(k_ghost_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, ghost, dst, mx, my,
                       meqn, g, bc[0], bc[1], bc[2], bc[3], nx_mom, ny_mom);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace finess
