// stage1d_body.h -- per-face and per-cell work of the 1D stage (see
// stage1d.hip); plain functions so that a CPU harness can run them too.
#pragma once
#include "lxw_body.h"
#include "physics.h"
#include "stage2d_body.h"
#include "step_control.h"
#include "weno.h"
#include "weno_wide.h"

namespace finess {

struct Stage1DArgs {
    const double* src;   // q planes (meqn=1): q[i], i in [0, mx)
    const double* qa;
    const double* qb;
    double* out;
    const double* aux;   // aux(1)[i] (advection speed), i in [0, mx)
    double* fhat;        // mx+1 face fluxes
    int mx;
    int bc_lo, bc_hi;
    int src_ghosts;      // >0: src/aux include that many ghosts per side
    int a_is_src, b_is_src, use_b;
    double a1, a2, bdt, dx, alpha_scaling, eps, power;
    unsigned long long* smax_slot;
    double* smax_faces;  // optional: alpha of each face f in [0, mx]
    double* Lout;        // optional: L of each cell
    double* galpha;      // global_alpha (null: off): the global wave speed, formed before the faces
    // device-controlled stepping (null: off), as in the 2D stage: bdt = beta *
    // (*dtp), and the kernels return at once while *halt is non-zero
    const double* dtp;
    double beta;
    const unsigned long long* halt;
    // source term (SRC_OFF: none): psi of the source's cell added to L
    int src_kind;
    double src_coef;
    // Lax-Wendroff (null: off): the time-integrated fluxes of the extended
    // cells [-3, mx+2], cell i at lxwF[i + 3], read as the stencil fluxes
    // (lib/1d/ConstructLxWL.cpp:14-237)
    const double* lxwF;
};

__host__ __device__ __forceinline__ bool halted1d(const Stage1DArgs& a) { return a.halt && *a.halt != 0ull; }

// The stages of one time step, launched together (stage k reads what stage
// k-1 wrote).
struct Stage1DStep {
    Stage1DArgs stage[3];
    int n;
    // the device controller after the last stage, in the same launch (null:
    // the host launches it, or there is none)
    StepCtl* ctl;
    StepCtlParams cp;
};

// A run of 1D time steps in one launch (grids up to the single-workgroup
// size): the stages of each attempt over the three plane sets, the device
// controller after it, and the next attempt -- from the accepted state, or
// from the same state with the controller's smaller dt after a rejection --
// until the controller halts the run (the end, an error) or `max_attempts`.
struct Stage1DRun {
    Stage1DArgs base;   // the common fields; dt, halt from the controller block
    double* buf[3];
    int cur;            // plane set holding the state at entry
    int n;              // SSP order 1..3
    unsigned long long* smax_slot;
    StepCtl* ctl;
    StepCtlParams cp;
    int max_attempts;
};

// SetWaveSpd of the advection callbacks with both sides equal (|a|) and the
// Max(alpha1, local) of lib/1d/ConstructL.cpp:148 when global_alpha is set.
__host__ __device__ __forceinline__ double alpha1d(const Stage1DArgs& a, double local)
{
    const double g = a.galpha ? *a.galpha : 0.0;
    return g > local ? g : local;  // a NaN speed stays NaN
}

__host__ __device__ __forceinline__ long long idx1d(const Stage1DArgs& a, int g)
{
    if (a.src_ghosts) return g + a.src_ghosts;
    int fl;
    return bc_map(a.bc_lo, a.bc_hi, g, a.mx, fl);
}

// Face f sits between cells f-1 and f (0-based); lib/1d/ConstructL.cpp:49-165.
// The reconstruction of split part `part` alone (0: f + alpha q over cells
// 0..4, 1: f - alpha q over cells 5..1), for two lanes per face; alpha
// returned through `alpha`.  The same arithmetic as face1d below.
template <int WK>
__host__ __device__ __forceinline__ double face1d_part(int f, const Stage1DArgs& a, int part, double& alpha)
{
    const WenoParams<double> wp{a.eps, a.power};
    double q[6], fl[6];
    for (int s = 0; s < 6; s++) {
        const long long k = idx1d(a, f - 3 + s);
        q[s] = a.src[k];
        fl[s] = a.lxwF ? a.lxwF[f + s] : q[s] * a.aux[k];  // FluxFunc: Q(i,1)*Aux(i,1)
    }
    const double al = a.aux[idx1d(a, f - 1)], ar = a.aux[idx1d(a, f)];
    const double s1 = rmin(al, 0.5 * (al + ar));
    const double s2 = rmax(ar, 0.5 * (al + ar));
    alpha = alpha1d(a, rmax(fabs(s1), fabs(s2)));
    const double la = a.alpha_scaling * alpha;
    double g[5];
    for (int s = 0; s < 5; s++) g[s] = part ? 0.5 * (fl[5 - s] - la * q[5 - s]) : 0.5 * (fl[s] + la * q[s]);
    return weno5<WK>(g, wp);
}

// Both parts in one pass (the stencil read and alpha formed once).
template <int WK>
__host__ __device__ __forceinline__ double face1d(int f, const Stage1DArgs& a)
{
    const WenoParams<double> wp{a.eps, a.power};
    double q[6], fl[6];
    for (int s = 0; s < 6; s++) {
        const long long k = idx1d(a, f - 3 + s);
        q[s] = a.src[k];
        fl[s] = a.lxwF ? a.lxwF[f + s] : q[s] * a.aux[k];  // FluxFunc: Q(i,1)*Aux(i,1)
    }
    const double al = a.aux[idx1d(a, f - 1)], ar = a.aux[idx1d(a, f)];
    const double s1 = rmin(al, 0.5 * (al + ar));
    const double s2 = rmax(ar, 0.5 * (al + ar));
    const double alpha = alpha1d(a, rmax(fabs(s1), fabs(s2)));
    const double la = a.alpha_scaling * alpha;
    double gp[5], gm[5];
    for (int s = 0; s < 5; s++) {
        gp[s] = 0.5 * (fl[s] + la * q[s]);
        gm[s] = 0.5 * (fl[5 - s] - la * q[5 - s]);
    }
    a.fhat[f] = weno5<WK>(gp, wp) + weno5<WK>(gm, wp);
    if (a.smax_faces) a.smax_faces[f] = alpha;
    return alpha;
}

// The other orders (lib/WenoReconstruct.cpp:39-73): a stencil of 2R cells,
// f-R .. f+R-1, split as lib/1d/ConstructL.cpp:153-158 does for any ws;
// R = 1 is the first-order value.
template <int R, int VER>
__host__ __device__ __forceinline__ double face1d_wide(int f, const Stage1DArgs& a)
{
    constexpr int WS = 2 * R - 1, NS = 2 * R;
    double q[NS], fl[NS];
    for (int s = 0; s < NS; s++) {
        const long long k = idx1d(a, f - R + s);
        q[s] = a.src[k];
        fl[s] = q[s] * a.aux[k];  // FluxFunc: Q(i,1)*Aux(i,1)
    }
    const double al = a.aux[idx1d(a, f - 1)], ar = a.aux[idx1d(a, f)];
    const double s1 = rmin(al, 0.5 * (al + ar));
    const double s2 = rmax(ar, 0.5 * (al + ar));
    const double alpha = alpha1d(a, rmax(fabs(s1), fabs(s2)));
    const double la = a.alpha_scaling * alpha;
    double gp[WS], gm[WS];
    for (int s = 0; s < WS; s++) {
        gp[s] = 0.5 * (fl[s] + la * q[s]);
        gm[s] = 0.5 * (fl[WS - s] - la * q[WS - s]);
    }
    if constexpr (R == 1) a.fhat[f] = gp[0] + gm[0];
    else a.fhat[f] = weno_wide<R, VER>(gp, a.eps, a.power) + weno_wide<R, VER>(gm, a.eps, a.power);
    if (a.smax_faces) a.smax_faces[f] = alpha;
    return alpha;
}

template <int WK>
__host__ __device__ __forceinline__ double face1d_any(int f, const Stage1DArgs& a)
{
    if constexpr (WK <= WENO_FD5) return face1d<WK>(f, a);
    else return face1d_wide<wide_radius(WK), wide_version(WK)>(f, a);
}

// ConstructIntegratedF (lib/1d/ConstructIntegratedF.cpp:54-212) with the
// advection callbacks (FluxFunc aux(1) q, DFluxFunc aux(1), D2FluxFunc 0):
// F = f + dt/2 (f_t + dt/3 f_tt), f_t = -a f_x, f_tt = a (a f_xx) for
// time_order > 2, with the 5-point differences Diff1 / Diff2; cell i of the
// extended range [-3, mx+2].
__host__ __device__ __forceinline__ double integrated_f1d(int i, const Stage1DArgs& a, double dt, int time_order)
{
    double fv[5];
    for (int s = 0; s < 5; s++) {
        const long long k = idx1d(a, i - 2 + s);
        fv[s] = a.src[k] * a.aux[k];
    }
    const double fx = d1(a.dx, fv[0], fv[1], fv[3], fv[4]);
    const double fxx = d2(a.dx, fv[0], fv[1], fv[2], fv[3], fv[4]);
    const double A = a.aux[idx1d(a, i)];
    const double ft = -A * fx;
    const double ftt = time_order > 2 ? A * (A * fxx) : 0.0;
    return fv[2] + 0.5 * dt * (ft + dt / 3.0 * ftt);
}

__host__ __device__ __forceinline__ void update1d(int i, const Stage1DArgs& a)
{
    double L = -(a.fhat[i + 1] - a.fhat[i]) / a.dx;
    if (a.src_kind != SRC_OFF) {  // Lstar = psi + tmp (lib/1d/ConstructL.cpp:194-205)
        const double q = a.src[idx1d(a, i)];
        double psi;
        source_psi<1>(a.src_kind, a.src_coef, &q, nullptr, 0, &psi);
        L = psi + L;
    }
    if (a.Lout) a.Lout[i] = L;
    if (!a.out) return;
    double v = a.a1 * a.qa[i];
    if (a.use_b) v = v + a.a2 * a.qb[i];
    a.out[i] = v + (a.dtp ? a.beta * *a.dtp : a.bdt) * L;
}

} // namespace finess
