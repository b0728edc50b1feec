/*
 * finess_oracle.h -- CPU restatement of FINESS's WENO operator and time
 * steppers.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke) and the CPU baseline timed
 * beside the GPU (bench.py).  No product code links or calls it.
 *
 * Arrays use the reference layout (lib/tensors.h:267-273): 2D fields are
 * dTensorBC3(mx,my,meqn,mbc), 1D fields dTensorBC2(mx,meqn,mbc).
 * Return codes are those of include/finess_hip.h.
 */
#ifndef FINESS_ORACLE_H
#define FINESS_ORACLE_H

#include "../include/finess_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* WENO reconstructions (lib/WenoReconstruct.cpp:90-141 JS5, :385-438 Z5,
 * :580-616 FD5): g[0..4] = stencil, returns the value at the right edge. */
double orc_weno5(int version, double eps, double power, const double* g);
/* Nonlinear weights of JS5 (omega_k = omt_k/omts, lib/WenoReconstruct.cpp:131-134). */
void   orc_weno5_js_weights(double eps, double power, const double* g, double* w);
/* Any order GetWenoReconstruct has (lib/WenoReconstruct.cpp:39-73): 1, 5,
 * 7, 9, 11 (JS), 7, 9 (Z, FD); g[0..order-1]. */
double orc_weno(int version, int order, double eps, double power, const double* g);
int    orc_weno_supported(int version, int order);
/* Candidate, smoothness-indicator and linear-weight tables of order 7, 9, 11
 * (cand[r][r], beta[r][r][r] upper triangles, d[r]); returns r or 0. */
int    orc_weno_tables(int order, double* cand, double* beta, double* d);

/* Physics callbacks of the configured app. Q, F, G are meqn-vectors. */
void orc_flux(const finess_params* p, const double* Q, const double* aux, double* F, double* G);
void orc_project_left (const finess_params* p, int ixy, const double* Qavg, int npts,
                       const double* Qvals /* [meqn][npts] */, double* Wvals);
void orc_project_right(const finess_params* p, int ixy, const double* Qavg, int npts,
                       const double* Wvals /* [meqn][npts] */, double* Qvals);
void orc_wave_speed(const finess_params* p, int ixy, const double* Ql, const double* Qr,
                    const double* auxl, const double* auxr, double* s1, double* s2);
/* Jacobians A[d][m1][m2] = d f_d,m1 / d q_m2 and Hessians H[d][m][m1][m2]. */
void orc_dflux (const finess_params* p, const double* Q, double* A /* [2][meqn][meqn] */);
void orc_d2flux(const finess_params* p, const double* Q, double* H /* [2][meqn][meqn][meqn] */);

/* SetBndValues for the configured per-side boundary conditions. */
void orc_set_bnd(const finess_params* p, double* q, double* aux);

/* ConstructL (2D or 1D by p->ndims). */
int orc_construct_l(const finess_params* p, const double* q, const double* aux,
                    double* Lstar, double* smax);

/* GetCFL (lib/2d/GetCFL.cpp:31-74, lib/1d/GetCFL.cpp:9-45). */
int orc_get_cfl(const finess_params* p, double dt, double dtmax, const double* smax, double* cfl);

typedef struct orc_stats {
    long long steps_accepted, steps_rejected;
    double dt_min, dt_max, last_cfl;
} orc_stats;

/* FinSolveRK / FinSolveLxW: q in/out (reference layout), *t in/out. */
int orc_fin_solve_rk (const finess_params* p, double* q, double* aux, double* t,
                      double tend, double* dtv, orc_stats* st);
int orc_fin_solve_lxw(const finess_params* p, double* q, double* aux, double* t,
                      double tend, double* dtv, orc_stats* st);
/* FinSolveRK_PIF (lib/2d/FinSolveRK_PIF.cpp:66-295; 2D, time_order 4). */
int orc_fin_solve_rk_pif(const finess_params* p, double* q, double* aux, double* t,
                         double tend, double* dtv, orc_stats* st);
/* Fixed number of steps of size dt, no controller (for timing and parity at
 * a stated step count).  Returns the CFL of the last step in *cfl. */
int orc_run_fixed_steps(const finess_params* p, double* q, double* aux, double dt,
                        int nsteps, double* cfl);

/* ConstructIntegratedR + ConstructLxWL (lib/2d/ConstructIntegratedR.cpp:59-296,
 * lib/2d/ConstructLxWL.cpp:30-414). */
int orc_construct_integrated_r(const finess_params* p, double dt, const double* q,
                               const double* aux, double* F, double* G);
int orc_construct_lxw_l(const finess_params* p, const double* q, const double* aux,
                        const double* F, const double* G, double* Lstar, double* smax);

/* Positivity-preserving limiter of the LxW path (2D Euler): GlobalWaveSpd
 * (lib/2d/GlobalWaveSpd.cpp:20-80), and the interface fluxes of one limited
 * step of q (boundary values set): high-order Fh, Gh, Lax-Friedrichs Fl, Gl
 * (lib/2d/ConstructLFL.cpp:25-159), limited Fm, Gm
 * (apps/2d/euler/lib/ApplyMPPLimiter2D.cpp:27-286).  x-face arrays
 * (mx+1+2mbc)(my+2mbc)meqn, y-face arrays (mx+2mbc)(my+1+2mbc)meqn, index
 * (i, j, m) as the state's with my+1 in place of my for y faces.  Any output
 * may be NULL except Fm, Gm. */
void orc_global_wave_spd(const finess_params* p, const double* q, const double* aux, double* alpha1,
                         double* alpha2);
int orc_lxw_mpp_faces(const finess_params* p, double dt, const double* q, const double* aux, double* Fh,
                      double* Gh, double* Fl, double* Gl, double* Fm, double* Gm);

/* ---- troubled-interface indicator (no counterpart in the reference) ----
 * qs: the 6 stencil states q(i-3..i+2) across a face, [6][meqn]. */
void   orc_face_features(const finess_params* p, int ixy, const double* qs, float* feat /* [24] */);
double orc_face_departure(const finess_params* p, int ixy, const double* qs);
/* The smooth-interface flux (component-wise LLF, FD5 of lib/WenoReconstruct.cpp:580-616
 * on each split part); returns alpha. */
double orc_face_flux_linear(const finess_params* p, int ixy, const double* qs, double* fhat);
/* ConstructL with routing: mode 0 characteristic WENO everywhere, 1 the
 * component-wise linear fifth-order flux everywhere, 2 characteristic WENO
 * where orc_face_departure > tau else linear.  *n_troubled counts faces. */
int orc_construct_l_routed(const finess_params* p, const double* q, const double* aux, int mode, double tau,
                           double* Lstar, double* smax, long long* n_troubled);
/* Fixed SSP-RK3 steps with mode-2 routing.  per_step = k >= 1: the first
 * stage of every k-th step decides the routing, which the later stages of
 * that step and the next k-1 steps reuse; 0: decided at every stage. */
int orc_run_fixed_steps_routed(const finess_params* p, double* q, double* aux, double dt, int nsteps, double tau,
                               int per_step, double* cfl);
/* MLP 24-32-32-1 (parameter layout of include/finess_hip.h's finess_mlp),
 * every matrix product with both operands rounded to BF16 as the matrix cores
 * do, accumulated in FP64. */
float orc_bf16(float x);
/* 0: exact products (for finite-difference checks of the gradient). */
void  orc_mlp_set_rounding(int on);
void  orc_mlp_forward(const float* params, const float* X, int n, float* logits);
/* Gradient of the mean binary cross-entropy; returns the mean loss. */
double orc_mlp_gradient(const float* params, const float* X, const float* y, int n, double* grad);

/* Number of OpenMP threads the oracle runs with. */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
