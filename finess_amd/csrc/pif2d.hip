// pif2d.hip -- the RK4-PIF step of FinSolveRK_PIF for 2D Euler and shallow
// water (lib/2d/FinSolveRK_PIF.cpp:151-206).  Per stage:
//
//   k_pif_flux_sum   per cell of the extended grid [-3, mx+2]^2: adds w
//                    times the flux of the boundary-mapped state to the
//                    running sums F, G (weights 1, 2, 2, 1; the last call
//                    divides by 6), the time-averaged fluxes fstar, gstar
//                    (:155-186), laid out as the Lax-Wendroff stage kernel
//                    reads its time-integrated fluxes;
//   k_pif_noc_faces  per face: ConstructL_NOC's component-wise WENO of f/2
//                    without characteristic projection or wave-speed
//                    splitting (:723-977);
//   k_pif_noc_update per cell: EulerStep, q* = q^n + c dt L_NOC (:980-1004).
//
// The final stage (characteristic WENO on fstar, gstar with the
// eigen-systems and wave speeds of q^n, then q^n + dt L) is the Lax-Wendroff
// stage kernel.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "physics.h"
#include "stage2d_body.h"

namespace finess {

struct PifArgs {
    const double* q;     // the stage's state (planes)
    const double* qold;  // q^n
    double* out;
    int mx, my;
    long long plane;
    int bc[4];
    double dx, dy, cdt, w;
    int first, last;
    double eps, power;
    PhysParams pp;
    double *F, *G;       // extended planes (mx+6)(my+6)
    double *fx, *fy;     // interface fluxes: x faces [m][(mx+1) my], y faces [m][mx (my+1)]
};

template <class P>
__device__ __forceinline__ void pif_load(const PifArgs& a, int i, int j, double q[P::MEQN])
{
    int fi, fj;
    const int si = bc_map(a.bc[0], a.bc[1], i, a.mx, fi);
    const int sj = bc_map(a.bc[2], a.bc[3], j, a.my, fj);
    const long long idx = (long long)si * a.my + sj;
    for (int m = 0; m < P::MEQN; m++) q[m] = a.q[m * a.plane + idx];
    if (fi) q[P::NORMAL_X] = -q[P::NORMAL_X];
    if (fj) q[P::NORMAL_Y] = -q[P::NORMAL_Y];
}

// the flux of a state in direction dir, components in the state's order
template <class P>
__device__ __forceinline__ void pif_flux(const double q[P::MEQN], int dir, const PhysParams& pp,
                                         double f[P::MEQN])
{
    double rec[P::NV], s[P::MEQN], fr[P::MEQN];
    P::cell_record(q, pp, rec);
    P::state(rec, dir, s);
    P::flux(s, rec, fr);
    for (int r = 0; r < P::MEQN; r++) f[P::rot_q(r, dir)] = fr[r];
}

template <class P>
__global__ __launch_bounds__(256) void k_pif_flux_sum(PifArgs a)
{
    constexpr int M = P::MEQN;
    const int ld = a.my + 6;
    const long long lpl = (long long)(a.mx + 6) * ld;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= lpl) return;
    const int i = (int)(e / ld) - 3, j = (int)(e - (long long)(i + 3) * ld) - 3;
    double q[M], f[M], g[M];
    pif_load<P>(a, i, j, q);
    pif_flux<P>(q, 0, a.pp, f);
    pif_flux<P>(q, 1, a.pp, g);
    for (int m = 0; m < M; m++) {
        double sf = a.first ? 0.0 : a.F[m * lpl + e];
        double sg = a.first ? 0.0 : a.G[m * lpl + e];
        sf += a.w * f[m];  // w = 1 or 2: the product is exact, so a fused add rounds the same
        sg += a.w * g[m];
        if (a.last) { sf = sf / 6.0; sg = sg / 6.0; }
        a.F[m * lpl + e] = sf;
        a.G[m * lpl + e] = sg;
    }
}

template <class P, int WK>
__global__ __launch_bounds__(256) void k_pif_noc_faces(PifArgs a)
{
    constexpr int M = P::MEQN;
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfx + nfy) return;
    const int dir = k < nfx ? 0 : 1;
    int i, j;
    if (dir == 0) { i = (int)(k / a.my); j = (int)(k - (long long)i * a.my); }
    else { const long long h = k - nfx; i = (int)(h / (a.my + 1)); j = (int)(h - (long long)i * (a.my + 1)); }
    double fs[6][M];
    for (int s = 0; s < 6; s++) {
        double q[M];
        if (dir == 0) pif_load<P>(a, i - 3 + s, j, q);
        else pif_load<P>(a, i, j - 3 + s, q);
        pif_flux<P>(q, dir, a.pp, fs[s]);
    }
    const WenoParams<double> wp{a.eps, a.power};
    double* dst = dir ? a.fy : a.fx;
    const long long nf = dir ? nfy : nfx, f = dir ? k - nfx : k;
    for (int m = 0; m < M; m++) {
        double gp[5], gm[5];
        for (int s = 0; s < 5; s++) {
            gp[s] = 0.5 * fs[s][m];
            gm[s] = 0.5 * fs[5 - s][m];
        }
        dst[m * nf + f] = weno5<WK>(gp, wp) + weno5<WK>(gm, wp);
    }
}

template <int M>
__global__ __launch_bounds__(256) void k_pif_noc_update(PifArgs a)
{
    const long long n = a.plane;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c / a.my), j = (int)(c - (long long)i * a.my);
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long fx0 = (long long)i * a.my + j, fx1 = fx0 + a.my;
    const long long fy0 = (long long)i * (a.my + 1) + j, fy1 = fy0 + 1;
    for (int m = 0; m < M; m++) {
        double L = -(a.fx[m * nfx + fx1] - a.fx[m * nfx + fx0]) / a.dx;
        L = L - (a.fy[m * nfy + fy1] - a.fy[m * nfy + fy0]) / a.dy;
        a.out[m * a.plane + c] = a.qold[m * a.plane + c] + a.cdt * L;
    }
}

static PifArgs pif_args(const PifLaunch& L)
{
    PifArgs a;
    a.q = L.q; a.qold = L.qold; a.out = L.out; a.mx = L.mx; a.my = L.my; a.plane = (long long)L.mx * L.my;
    for (int k = 0; k < 4; k++) a.bc[k] = L.bc[k];
    a.dx = L.dx; a.dy = L.dy; a.cdt = L.cdt; a.w = L.w; a.first = L.first; a.last = L.last;
    a.eps = L.eps; a.power = L.power;
    a.pp.gamma = L.gamma; a.pp.sw_compat_hl = L.sw_compat_hl;
    a.F = L.F; a.G = L.G; a.fx = L.fx; a.fy = L.fy;
    return a;
}

template <class P>
static int pif_flux_sum_t(const PifArgs& a, hipStream_t st)
{
    const long long lpl = (long long)(a.mx + 6) * (a.my + 6);
    hipLaunchKernelGGL(k_pif_flux_sum<P>, dim3((unsigned)((lpl + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_pif_flux_sum(const PifLaunch& L, hipStream_t st)
{
    const PifArgs a = pif_args(L);
    if (L.physics == PHYS_EULER) return pif_flux_sum_t<Euler2D>(a, st);
    if (L.physics == PHYS_SHALLOW_WATER) return pif_flux_sum_t<ShallowWater2D>(a, st);
    return -1;
}

template <class P>
static int pif_noc_stage_t(int wk, const PifArgs& a, hipStream_t st)
{
    const long long faces = (long long)(a.mx + 1) * a.my + (long long)a.mx * (a.my + 1);
    const dim3 bf((unsigned)((faces + 255) / 256)), bc((unsigned)((a.plane + 255) / 256)), blk(256);
    switch (wk) {
    case WENO_JS5_P2: hipLaunchKernelGGL((k_pif_noc_faces<P, WENO_JS5_P2>), bf, blk, 0, st, a); break;
    case WENO_JS5_POW: hipLaunchKernelGGL((k_pif_noc_faces<P, WENO_JS5_POW>), bf, blk, 0, st, a); break;
    case WENO_Z5: hipLaunchKernelGGL((k_pif_noc_faces<P, WENO_Z5>), bf, blk, 0, st, a); break;
    case WENO_FD5: hipLaunchKernelGGL((k_pif_noc_faces<P, WENO_FD5>), bf, blk, 0, st, a); break;
    default: return -1;
    }
    hipLaunchKernelGGL(k_pif_noc_update<P::MEQN>, bc, blk, 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_pif_noc_stage(int wk, const PifLaunch& L, hipStream_t st)
{
    const PifArgs a = pif_args(L);
    if (L.physics == PHYS_EULER) return pif_noc_stage_t<Euler2D>(wk, a, st);
    if (L.physics == PHYS_SHALLOW_WATER) return pif_noc_stage_t<ShallowWater2D>(wk, a, st);
    return -1;
}

} // namespace finess
