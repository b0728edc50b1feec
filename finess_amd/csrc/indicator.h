// indicator.h -- the troubled-interface indicator: its input features, the
// label the solver generates for training, and the cheap face flux used at
// interfaces the indicator calls smooth.  No counterpart in the reference.
//
// Label.  An interface is troubled when the nonlinear WENO5-JS weights of any
// characteristic field, for either Lax-Friedrichs split part, depart from the
// linear weights d = (0.1, 0.6, 0.3) by more than tau:
//     max_{fields, +/-} max_k |omega_k - d_k| > tau.
// The weights are those the reference computes at the interface
// (lib/WenoReconstruct.cpp:131-134 applied to the split characteristic
// stencils of lib/2d/ConstructL.cpp:177-192).
//
// Features (24).  The 6 stencil cells (i-3..i+2 across the face) times the
// rotated primitives (rho, u_n, u_t, p).  Each variable is centred on its
// stencil mean and divided by (its stencil range + 1% of a reference
// magnitude): mean rho for rho, mean p for p, the mean sound speed for both
// velocities.  Features are dimensionless, O(1) across a jump and near zero
// in flat regions, whatever the units.
//
// Smooth-interface flux.  Component-wise Lax-Friedrichs splitting with the
// linear fifth-order upwind reconstruction (WenoReconstruct_FD5,
// lib/WenoReconstruct.cpp:580-616) on the conserved variables: no
// characteristic projection, no nonlinear weights.
#pragma once
#include "physics.h"
#include "weno.h"

namespace finess {

constexpr int MLP_IN = 24, MLP_H = 32;
constexpr int MLP_NPARAM = MLP_IN * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;  // 1889

// Parameter vector layout (FP32 master copy):
//   W1[k][n] (k < 24, n < 32) | b1[32] | W2[k][n] (32x32) | b2[32] | w3[32] | b3
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;

// Harvest of training interfaces from a 2D Euler state (mlp.hip).  Faces k <
// (mx+1)*my are x-faces (i = k / my, j = k % my), the rest y-faces; every
// `stride`-th face (offset `phase`) is written at row out0 + k / stride.
struct HarvestArgs {
    const double* q;
    int mx, my;
    long long plane;
    int bc[4];
    PhysParams pp;
    double eps, alpha_scaling, tau;
    int stride, phase;
    float* X;          // [n][24] features
    float* y;          // [n] labels
    float* dep;        // [n] weight departure (may be null)
    long long out0, cap;
};

// Features of the stencil records rec[s] (s = 0..5, cells i-3..i+2) for a
// face of direction dir.  Euler records: q0..q4, 1/rho, p, c.
template <class P, typename R>
__host__ __device__ __forceinline__ void face_features(const R (*rec)[P::NV], int dir, const PhysParams& pp,
                                                       float feat[MLP_IN]);

template <>
__host__ __device__ __forceinline__ void face_features<Euler2D, double>(const double (*rec)[8], int dir,
                                                                       const PhysParams& pp, float feat[MLP_IN])
{
    double v[6][4];
    for (int s = 0; s < 6; s++) {
        const double ir = rec[s][5];
        v[s][0] = rec[s][0];
        v[s][1] = (dir ? rec[s][2] : rec[s][1]) * ir;
        v[s][2] = (dir ? rec[s][1] : rec[s][2]) * ir;
        v[s][3] = rec[s][6];
    }
    double mean[4], lo[4], hi[4];
    for (int k = 0; k < 4; k++) {
        mean[k] = 0.0; lo[k] = v[0][k]; hi[k] = v[0][k];
        for (int s = 0; s < 6; s++) {
            mean[k] += v[s][k];
            lo[k] = fmin(lo[k], v[s][k]);
            hi[k] = fmax(hi[k], v[s][k]);
        }
        mean[k] *= (1.0 / 6.0);
    }
    const double cref = sqrt(fabs(pp.gamma * mean[3] / mean[0]));
    const double ref[4] = {fabs(mean[0]), cref, cref, fabs(mean[3])};
    for (int k = 0; k < 4; k++) {
        const double inv = 1.0 / (hi[k] - lo[k] + 0.01 * ref[k] + 1e-300);
        for (int s = 0; s < 6; s++) feat[s * 4 + k] = float((v[s][k] - mean[k]) * inv);
    }
}

template <>
__host__ __device__ __forceinline__ void face_features<Euler2D, float>(const float (*rec)[8], int dir,
                                                                      const PhysParams& pp, float feat[MLP_IN])
{
    float v[6][4];
    for (int s = 0; s < 6; s++) {
        const float ir = rec[s][5];
        v[s][0] = rec[s][0];
        v[s][1] = (dir ? rec[s][2] : rec[s][1]) * ir;
        v[s][2] = (dir ? rec[s][1] : rec[s][2]) * ir;
        v[s][3] = rec[s][6];
    }
    float mean[4], lo[4], hi[4];
    for (int k = 0; k < 4; k++) {
        mean[k] = 0.0f; lo[k] = v[0][k]; hi[k] = v[0][k];
        for (int s = 0; s < 6; s++) {
            mean[k] += v[s][k];
            lo[k] = fminf(lo[k], v[s][k]);
            hi[k] = fmaxf(hi[k], v[s][k]);
        }
        mean[k] *= (1.0f / 6.0f);
    }
    const float cref = sqrt_lp(fabsf(div_lp(float(pp.gamma) * mean[3], mean[0])));
    const float ref[4] = {fabsf(mean[0]), cref, cref, fabsf(mean[3])};
    for (int k = 0; k < 4; k++) {
        const float inv = frcp(hi[k] - lo[k] + 0.01f * ref[k] + 1e-30f);
        for (int s = 0; s < 6; s++) feat[s * 4 + k] = (v[s][k] - mean[k]) * inv;
    }
}

// Largest departure of the JS weights from the linear weights over all
// characteristic fields and both split parts, in the arithmetic of R.
template <class P, typename R>
__host__ __device__ __forceinline__ R face_departure(const R (*rec)[P::NV], int dir, const PhysParams& pp, R eps,
                                                     R ascale)
{
    constexpr int M = P::MEQN;
    R ql[M], qr[M], qa[M];
    P::state(rec[2], dir, ql);
    P::state(rec[3], dir, qr);
    for (int k = 0; k < M; k++) qa[k] = R(0.5) * (qr[k] + ql[k]);
    typename P::template Eig<R> e;
    P::eig(qa, pp, e);
    const R la = ascale * P::wave_speed(ql, qr, rec[2], rec[3], pp);
    R gp[M][5], gm[M][5];
    for (int s = 0; s < 6; s++) {
        R q[M], f[M], vp[M], vm[M], t[M];
        P::state(rec[s], dir, q);
        P::flux(q, rec[s], f);
        for (int k = 0; k < M; k++) { vp[k] = f[k] + la * q[k]; vm[k] = f[k] - la * q[k]; }
        if (s < 5) { P::left(e, vp, t); for (int k = 0; k < M; k++) gp[k][s] = t[k]; }
        if (s > 0) { P::left(e, vm, t); for (int k = 0; k < M; k++) gm[k][5 - s] = t[k]; }
    }
    const WenoParams<R> wp4{R(4) * eps, R(2)};
    const R d[3] = {R(0.1), R(0.6), R(0.3)};
    R dep = R(0);
    for (int k = 0; k < M; k++) {
        R w[3];
        weno5_weights<WENO_JS5_POW>(gp[k], wp4, w);
        for (int j = 0; j < 3; j++) dep = rmax(dep, R(fabs(w[j] - d[j])));
        weno5_weights<WENO_JS5_POW>(gm[k], wp4, w);
        for (int j = 0; j < 3; j++) dep = rmax(dep, R(fabs(w[j] - d[j])));
    }
    return dep;
}

// Smooth-interface flux: component-wise LLF + linear fifth-order upwind
// reconstruction of the split conserved fluxes.  Returns alpha (the same
// wave speed the characteristic path uses, so the CFL is unchanged).
// The FD5 weights (2, -13, 47, 27, -3)/60 on f + la q over cells 0..4 and on
// f - la q over cells 5..1 fold into one symmetric stencil on f and one
// antisymmetric stencil on q:
//   fhat = [f0 - 8f1 + 37f2 + 37f3 - 8f4 + f5]/60 + la [q0 - 5q1 + 10q2 - 10q3 + 5q4 - q5]/60.
template <class P, typename R, int NC>
__host__ __device__ __forceinline__ R face_flux_linear(const R (*cv)[NC], int base, int stride, int dir,
                                                       const PhysParams& pp, R ascale, R fh[P::MEQN])
{
    constexpr int M = P::MEQN, NV = P::NV;
    // accumulated cell by cell, so only the sums and one cell stay live
    constexpr float A[6] = {1.f, -8.f, 37.f, 37.f, -8.f, 1.f};
    constexpr float B[6] = {1.f, -5.f, 10.f, -10.f, 5.f, -1.f};
    R fa[M], qb[M], cl[NV], cr[NV], ql[M], qr[M];
#pragma unroll
    for (int s = 0; s < 6; s++) {
        R rec[NV], q[M], f[M];
        for (int v = 0; v < NV; v++) rec[v] = cv[v][base + s * stride];
        P::state(rec, dir, q);
        P::flux(q, rec, f);
        for (int k = 0; k < M; k++) {
            fa[k] = s == 0 ? f[k] : fa[k] + R(A[s]) * f[k];
            qb[k] = s == 0 ? q[k] : qb[k] + R(B[s]) * q[k];
        }
        if (s == 2) {
            for (int v = 0; v < NV; v++) cl[v] = rec[v];
            for (int k = 0; k < M; k++) ql[k] = q[k];
        }
        if (s == 3) {
            for (int v = 0; v < NV; v++) cr[v] = rec[v];
            for (int k = 0; k < M; k++) qr[k] = q[k];
        }
    }
    const R alpha = P::wave_speed(ql, qr, cl, cr, pp);
    const R la = ascale * alpha;
    const R k60 = R(1) / R(60);
    for (int k = 0; k < M; k++) fh[P::rot_q(k, dir)] = (fa[k] + la * qb[k]) * k60;
    return alpha;
}

} // namespace finess
