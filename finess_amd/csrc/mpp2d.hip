// mpp2d.hip -- the positivity-preserving flux limiter of the 2D Lax-Wendroff
// step for Euler (lib/2d/FinSolveLxW.cpp:126-157).  After the Lax-Wendroff
// stage kernel has written the high-order interface fluxes (instead of
// updating the state), five launches:
//
//   k_mpp_speed   per cell: the global wave speeds max |s| of SetWaveSpd
//                 with the cell against itself (lib/2d/GlobalWaveSpd.cpp:20-80),
//                 folded into spread maxima words;
//   k_mpp_lf      per face: the Lax-Friedrichs flux with those speeds
//                 (lib/2d/ConstructLFL.cpp:112-128); the speeds also enter the
//                 step's CFL maxima (ConstructLFL sets smax(1,1,.), :62-64);
//   k_mpp_theta   per cell: the largest fractions of the four high-order flux
//                 corrections that keep density and pressure positive
//                 (apps/2d/euler/lib/ApplyMPPLimiter2D.cpp:56-232);
//   k_mpp_limit   per face: theta from its two cells and the limited flux
//                 theta (fHat - fLF) + fLF, in place (:236-284);
//   k_mpp_update  per cell: q - dt/dx (F(i+1) - F(i)), then - dt/dy (G(j+1) - G(j)).
//
// The limiter's arithmetic is FP64 without contraction, in the reference's
// operation order, so that its branch decisions (signs, the 9-step bisection)
// follow the CPU code's.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "physics.h"
#include "stage2d_body.h"

namespace finess {

struct MppArgs {
    const double* q;           // stage source (field planes, interior)
    double* out;               // updated state
    int mx, my;
    long long plane;
    int bc[4];
    double dx, dy, dt;
    PhysParams pp;
    double* fh;                // x faces [m][(mx+1) my]: high-order in, limited out
    double* gh;                // y faces [m][mx (my+1)]
    double* fl;                // Lax-Friedrichs fluxes, same layout
    double* gl;
    double* amin;              // [4][mx my]
    unsigned long long* gmax;  // [2][CFL_SPREAD] global wave-speed words
    unsigned long long* smax;  // the step's CFL words (slot 0)
};

// the state of cell (i, j) through the boundary map
__device__ __forceinline__ void mpp_load(const MppArgs& a, int i, int j, double q[5])
{
    int fi, fj;
    const int si = bc_map(a.bc[0], a.bc[1], i, a.mx, fi);
    const int sj = bc_map(a.bc[2], a.bc[3], j, a.my, fj);
    const long long idx = (long long)si * a.my + sj;
    for (int m = 0; m < 5; m++) q[m] = a.q[m * a.plane + idx];
    if (fi) q[Euler2D::NORMAL_X] = -q[Euler2D::NORMAL_X];
    if (fj) q[Euler2D::NORMAL_Y] = -q[Euler2D::NORMAL_Y];
}

__global__ __launch_bounds__(256) void k_mpp_speed(MppArgs a)
{
    __shared__ double red[2][4];
    const long long n = (long long)a.mx * a.my;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double sx = 0.0, sy = 0.0;
    if (k < n) {
        double q[5], rec[8];
        for (int m = 0; m < 5; m++) q[m] = a.q[m * a.plane + k];
        Euler2D::cell_record(q, a.pp, rec);
        double qs[5];
        Euler2D::state(rec, 0, qs);
        sx = Euler2D::wave_speed(qs, qs, rec, rec, a.pp);
        Euler2D::state(rec, 1, qs);
        sy = Euler2D::wave_speed(qs, qs, rec, rec, a.pp);
        if (sx != sx) sx = __builtin_inf();
        if (sy != sy) sy = __builtin_inf();
    }
    sx = wave_max64(sx);
    sy = wave_max64(sy);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[0][w] = sx; red[1][w] = sy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < 4; v++) {
            sx = red[0][v] > sx ? red[0][v] : sx;
            sy = red[1][v] > sy ? red[1][v] : sy;
        }
        const int b = blockIdx.x % CFL_SPREAD;
        atomicMax(&a.gmax[b], (unsigned long long)__double_as_longlong(sx));
        atomicMax(&a.gmax[CFL_SPREAD + b], (unsigned long long)__double_as_longlong(sy));
    }
}

// the global speeds, reduced by a whole wave from the spread words
__device__ __forceinline__ void mpp_global_speeds(const MppArgs& a, double& ax, double& ay)
{
    const int lane = threadIdx.x & 63;
    ax = wave_max64(spread_lane_max(a.gmax, lane));
    ay = wave_max64(spread_lane_max(a.gmax + CFL_SPREAD, lane));
}

__global__ __launch_bounds__(256) void k_mpp_lf(MppArgs a)
{
#pragma clang fp contract(off)
    double ax, ay;
    mpp_global_speeds(a, ax, ay);
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        atomicMax(&a.smax[0], (unsigned long long)__double_as_longlong(ax));
        atomicMax(&a.smax[CFL_SPREAD], (unsigned long long)__double_as_longlong(ay));
    }
    if (k >= nfx + nfy) return;
    const int dir = k < nfx ? 0 : 1;
    int i, j;
    if (dir == 0) { i = (int)(k / a.my); j = (int)(k - (long long)i * a.my); }
    else { const long long h = k - nfx; i = (int)(h / (a.my + 1)); j = (int)(h - (long long)i * (a.my + 1)); }
    double qa[5], qb[5], ra[8], rb[8];
    if (dir == 0) { mpp_load(a, i - 1, j, qa); mpp_load(a, i, j, qb); }
    else { mpp_load(a, i, j - 1, qa); mpp_load(a, i, j, qb); }
    Euler2D::cell_record(qa, a.pp, ra);
    Euler2D::cell_record(qb, a.pp, rb);
    double sa[5], sb[5], fa[5], fb[5];
    Euler2D::state(ra, dir, sa);
    Euler2D::state(rb, dir, sb);
    Euler2D::flux(sa, ra, fa);
    Euler2D::flux(sb, rb, fb);
    const double al = dir ? ay : ax;
    double* dst = dir ? a.gl : a.fl;
    const long long nf = dir ? nfy : nfx, fidx = dir ? k - nfx : k;
    for (int r = 0; r < 5; r++) {
        const int m = Euler2D::rot_q(r, dir);
        dst[m * nf + fidx] = 0.5 * (fa[r] + fb[r]) + 0.5 * al * (qa[m] - qb[m]);
    }
}

__device__ __forceinline__ double mpp_pressure(double gm1, const double Q[5])
{
#pragma clang fp contract(off)
    const double rho = Q[0], u1 = Q[1] / rho, u2 = Q[2] / rho, u3 = Q[3] / rho, energy = Q[4];
    return gm1 * (energy - 0.5 * rho * (u1 * u1 + u2 * u2 + u3 * u3));
}

__global__ __launch_bounds__(256) void k_mpp_theta(MppArgs a)
{
#pragma clang fp contract(off)
    const long long n = (long long)a.mx * a.my;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c / a.my), j = (int)(c - (long long)i * a.my);
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    // faces 0: x at i, 1: x at i+1, 2: y at j, 3: y at j+1
    const long long fx0 = (long long)i * a.my + j, fx1 = fx0 + a.my;
    const long long fy0 = (long long)i * (a.my + 1) + j, fy1 = fy0 + 1;
    double q0[5], fhl[4][5], fll[4][5];
    for (int m = 0; m < 5; m++) {
        q0[m] = a.q[m * a.plane + c];
        fhl[0][m] = a.fh[m * nfx + fx0];
        fhl[1][m] = a.fh[m * nfx + fx1];
        fhl[2][m] = a.gh[m * nfy + fy0];
        fhl[3][m] = a.gh[m * nfy + fy1];
        fll[0][m] = a.fl[m * nfx + fx0];
        fll[1][m] = a.fl[m * nfx + fx1];
        fll[2][m] = a.gl[m * nfy + fy0];
        fll[3][m] = a.gl[m * nfy + fy1];
    }
    const double eps = 1.0e-12, rho_min = eps;
    const double dtx = a.dt / a.dx, dty = a.dt / a.dy;
    const double gm1 = a.pp.gamma - 1.0;
    // density (:56-95)
    const double gmin = rho_min - (q0[0] + dtx * (fll[0][0] - fll[1][0]) + dty * (fll[2][0] - fll[3][0]));
    double ff[4];
    ff[0] = dtx * (fhl[0][0] - fll[0][0]);
    ff[1] = -dtx * (fhl[1][0] - fll[1][0]);
    ff[2] = dty * (fhl[2][0] - fll[2][0]);
    ff[3] = -dty * (fhl[3][0] - fll[3][0]);
    bool neg[4];
    for (int k = 0; k < 4; k++) neg[k] = ff[k] < 0.0;
    double ffsum = double(neg[0]) * ff[0];
    for (int k = 1; k < 4; k++) ffsum = ffsum + double(neg[k]) * ff[k];
    double am[4];
    for (int k = 0; k < 4; k++) {
        const double r = gmin / (ffsum - eps);
        am[k] = neg[k] ? (r < 1.0 ? r : 1.0) : 1.0;
    }
    // pressure, every combination of the four faces (:96-232)
    double qlf[5];
    for (int m = 0; m < 5; m++) qlf[m] = q0[m] + dtx * (fll[0][m] - fll[1][m]) + dty * (fll[2][m] - fll[3][m]);
    double r2[4] = {1.0, 1.0, 1.0, 1.0};
    for (int cmb = 0; cmb < 16; cmb++) {
        double aa[4];
        for (int k = 0; k < 4; k++) aa[k] = double((cmb >> (3 - k)) & 1) * am[k];
        double qh[5];
        for (int m = 0; m < 5; m++) {
            double f4[4];
            for (int k = 0; k < 4; k++) f4[k] = aa[k] * fhl[k][m] + (1.0 - aa[k]) * fll[k][m];
            qh[m] = q0[m] + dtx * (f4[0] - f4[1]) + dty * (f4[2] - f4[3]);
        }
        double phigh = mpp_pressure(gm1, qh);
        double rescale = 1.0;
        if (phigh < 0.0) {
            double root1 = 0.0, root2 = 1.0, rate = 0.0;
            for (int nn = 1; nn < 10; nn++) {
                rate = (root1 + root2) / 2.0;
                double qt[5];
                for (int m = 0; m < 5; m++) qt[m] = rate * qh[m] + (1.0 - rate) * qlf[m];
                phigh = mpp_pressure(gm1, qt);
                if (phigh < 0.0) root2 = rate;
                else root1 = rate;
            }
            if (!(phigh > 0.0)) rate = root1;
            rescale = rate;
        }
        for (int k = 0; k < 4; k++)
            if ((cmb >> (3 - k)) & 1) r2[k] = r2[k] < rescale ? r2[k] : rescale;
    }
    for (int k = 0; k < 4; k++) a.amin[k * n + c] = r2[k] * am[k];
}

__global__ __launch_bounds__(256) void k_mpp_limit(MppArgs a)
{
#pragma clang fp contract(off)
    const long long n = (long long)a.mx * a.my;
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfx + nfy) return;
    double th;
    double *h, *l;
    long long nf, f;
    if (k < nfx) {
        const int i = (int)(k / a.my), j = (int)(k - (long long)i * a.my);
        const double* A0 = a.amin;          // left faces of the cells
        const double* A1 = a.amin + n;      // right faces
        const long long cr = (long long)i * a.my + j, cl = cr - a.my;
        if (i == 0) th = A0[cr];
        else if (i == a.mx) th = A1[cl];
        else th = A0[cr] < A1[cl] ? A0[cr] : A1[cl];
        h = a.fh; l = a.fl; nf = nfx; f = k;
    } else {
        const long long hh = k - nfx;
        const int i = (int)(hh / (a.my + 1)), j = (int)(hh - (long long)i * (a.my + 1));
        const double* A2 = a.amin + 2 * n;  // bottom faces
        const double* A3 = a.amin + 3 * n;  // top faces
        const long long ct = (long long)i * a.my + j, cb = ct - 1;
        if (j == 0) th = A2[ct];
        else if (j == a.my) th = A3[cb];
        else th = A2[ct] < A3[cb] ? A2[ct] : A3[cb];
        h = a.gh; l = a.gl; nf = nfy; f = hh;
    }
    for (int m = 0; m < 5; m++) h[m * nf + f] = th * (h[m * nf + f] - l[m * nf + f]) + l[m * nf + f];
}

__global__ __launch_bounds__(256) void k_mpp_update(MppArgs a)
{
#pragma clang fp contract(off)
    const long long n = (long long)a.mx * a.my;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c / a.my), j = (int)(c - (long long)i * a.my);
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long fx0 = (long long)i * a.my + j, fx1 = fx0 + a.my;
    const long long fy0 = (long long)i * (a.my + 1) + j, fy1 = fy0 + 1;
    const double dtdx = a.dt / a.dx, dtdy = a.dt / a.dy;
    for (int m = 0; m < 5; m++) {
        double v = a.q[m * a.plane + c];
        double tmp = a.fh[m * nfx + fx1] - a.fh[m * nfx + fx0];
        v = v - dtdx * tmp;
        tmp = a.gh[m * nfy + fy1] - a.gh[m * nfy + fy0];
        v = v - dtdy * tmp;
        a.out[m * a.plane + c] = v;
    }
}

static MppArgs mpp_args(const MppLaunch& L)
{
    MppArgs a;
    a.q = L.q; a.out = L.out; a.mx = L.mx; a.my = L.my; a.plane = (long long)L.mx * L.my;
    for (int k = 0; k < 4; k++) a.bc[k] = L.bc[k];
    a.dx = L.dx; a.dy = L.dy; a.dt = L.dt;
    a.pp.gamma = L.gamma; a.pp.sw_compat_hl = 0;
    a.fh = L.fh; a.gh = L.gh; a.fl = L.fl; a.gl = L.gl; a.amin = L.amin; a.gmax = L.gmax; a.smax = L.smax;
    return a;
}

int launch_mpp(const MppLaunch& L, hipStream_t st)
{
    const MppArgs a = mpp_args(L);
    const long long cells = a.plane, faces = (long long)(a.mx + 1) * a.my + (long long)a.mx * (a.my + 1);
    const dim3 bc((unsigned)((cells + 255) / 256)), bf((unsigned)((faces + 255) / 256)), blk(256);
    if (launch_fill_u64(a.gmax, 2 * CFL_SPREAD, 0ull, st)) return -4;
    hipLaunchKernelGGL(k_mpp_speed, bc, blk, 0, st, a);
    hipLaunchKernelGGL(k_mpp_lf, bf, blk, 0, st, a);
    hipLaunchKernelGGL(k_mpp_theta, bc, blk, 0, st, a);
    hipLaunchKernelGGL(k_mpp_limit, bf, blk, 0, st, a);
    hipLaunchKernelGGL(k_mpp_update, bc, blk, 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_mpp_limit(const MppLaunch& L, hipStream_t st)
{
    const MppArgs a = mpp_args(L);
    const long long cells = a.plane, faces = (long long)(a.mx + 1) * a.my + (long long)a.mx * (a.my + 1);
    const dim3 bc((unsigned)((cells + 255) / 256)), bf((unsigned)((faces + 255) / 256)), blk(256);
    hipLaunchKernelGGL(k_mpp_theta, bc, blk, 0, st, a);
    hipLaunchKernelGGL(k_mpp_limit, bf, blk, 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace finess
