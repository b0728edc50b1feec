/*
 * finess_hip.h -- C-ABI of the MI355X (gfx950) implementation of FINESS's
 * finite-difference WENO spatial operator and its time steppers.
 *
 * Every entry point takes plain pointers and sizes.  State arrays crossing
 * this boundary use the reference's own tensor layout (lib/tensors.h:267-273,
 * lib/tensors.cpp:650-667): a 2D field q(i,j,m), 1-based, with mbc ghost
 * layers on the two spatial indices, lives at
 *     ((i+mbc-1)*(my+2*mbc) + (j+mbc-1))*meqn + (m-1)
 * and a 1D field q(i,m) at (i+mbc-1)*meqn + (m-1).  The library keeps its own
 * device copy in a field-major layout; the conversion happens on the device.
 *
 * Status codes: every function returns FINESS_OK (0) or a negative code.  The
 * reference reports the same conditions with exit(1) or an exception; the C++
 * drop-in layer (finess_amd/compat) maps the codes back to that behaviour.
 */
#ifndef FINESS_HIP_H
#define FINESS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define FINESS_OK               0
#define FINESS_ERR_CONFIG      -1  /* unsupported parameter combination (lib/WenoReconstruct.cpp:70) */
#define FINESS_ERR_NONFINITE   -2  /* non-finite state detected on the device                        */
#define FINESS_ERR_NV          -3  /* n_step > nv            (lib/2d/FinSolveRK.cpp:84-92)           */
#define FINESS_ERR_HIP         -4  /* HIP runtime error                                               */
#define FINESS_ERR_DT          -5  /* dt > max_dt            (lib/2d/GetCFL.cpp:43-50)               */
#define FINESS_ERR_CFL         -6  /* cfl > 1e8 (2D) / 1e10 (1D) (lib/2d/GetCFL.cpp:60-67)           */
#define FINESS_ERR_ARG         -7  /* null pointer / bad size                                        */
#define FINESS_ERR_NODEVICE    -8  /* no GPU visible                                                  */

/* ---- enumerations (IniParams::WenoVersion, TimeSteppingMethod) ------- */
#define FINESS_WENO_JS 0
#define FINESS_WENO_FD 1
#define FINESS_WENO_Z  2

#define FINESS_TS_RK   0
#define FINESS_TS_LXW  1
#define FINESS_TS_PIF  2  /* FinSolveRK_PIF (lib/2d/FinSolveRK_PIF.cpp:66-295): 2D, time_order 4 */

/* physics compiled into the kernels (the app callback libraries):
 *   ADVECTION      apps/1d/advection/lib  f = aux(1)*q           (1D, meqn=1)
 *   EULER          apps/2d/euler/lib      (rho,rho u,rho v,rho w,E) (2D, meqn=5)
 *   SHALLOW_WATER  apps/2d/shallow_water/lib (h,hu,hv), g=1        (2D, meqn=3) */
#define FINESS_PHYS_ADVECTION     0
#define FINESS_PHYS_EULER         1
#define FINESS_PHYS_SHALLOW_WATER 2

/* boundary condition per side (the SetBndValues templates):
 *   EXTRAP   zeroth-order extrapolation, apps/2d/euler/riemann_test/SetBndValues.cpp:9-114
 *   PERIODIC apps/2d/advection/smooth_example/SetBndValues.cpp:28-194 (with corners),
 *            apps/1d/advection/smooth_example/SetBndValues.cpp:6-63
 *   REFLECT  solid wall: mirror image, normal momentum negated                    */
#define FINESS_BC_EXTRAP   0
#define FINESS_BC_PERIODIC 1
#define FINESS_BC_REFLECT  2
/*   HOST     any other SetBndValues rule (time-dependent inflow such as
 *            apps/2d/euler/mach_reflection/SetBndValues.cpp:14-146): the
 *            ghost cells of the side come from the callback installed with
 *            finess_set_bnd_callback, called on the host before every stage
 *            (2D Runge-Kutta of orders 1-3, plain or routed, host controller) */
#define FINESS_BC_HOST     3

/* device forms of the app's SourceTermFunc (used when source_term = 1):
 *   ZERO           psi = 0: the SourceTermFunc of every reference app except
 *                  shallow water's (e.g. apps/2d/euler/mach_reflection/SourceTermFunc.cpp)
 *   SW_BATHYMETRY  psi = (0, -h aux(2), -h aux(3)), aux = (b, b_x, b_y), maux >= 3:
 *                  apps/2d/shallow_water/riemann_test/SourceTermFunc.cpp
 *   DAMPING        psi = -source_coef * q, every field (no reference app uses
 *                  it: a linear relaxation term with a closed-form solution)     */
#define FINESS_SRC_ZERO          0
#define FINESS_SRC_SW_BATHYMETRY 1
#define FINESS_SRC_DAMPING       2

/* stencil arithmetic of the device path */
#define FINESS_PREC_FP64 0   /* plain form, FP64 end to end                       */
#define FINESS_PREC_FP32 1   /* FP32 stencil, FP64 stage accumulation              */

/* interface routing */
#define FINESS_ROUTE_NONE    0 /* full characteristic WENO everywhere              */
#define FINESS_ROUTE_LEARNED 1 /* MLP indicator: smooth -> component-wise FD5     */
#define FINESS_ROUTE_ORACLE  2 /* route by the weight-departure label itself       */

/*
 * Parameters.  Field names follow the accessors of the reference's generated
 * IniParams (schema: python/finess/params/{finess,weno,dim1,dim2,euler}.py).
 */
typedef struct finess_params {
    /* [grid] (python/finess/params/dim2.py:8-65) */
    int    ndims;
    int    mx, my, mbc;
    double xlow, xhigh, ylow, yhigh;
    double dx, dy;                      /* derived: (xhigh-xlow)/mx, ... */
    /* [finess] (python/finess/params/finess.py:13-170) */
    int    meqn, maux;
    int    space_order;                 /* 1, 5, 7, 9, 11 (JS), 1, 5, 7, 9 (Z, FD), as
                                           GetWenoReconstruct (lib/WenoReconstruct.cpp:39-73);
                                           orders other than 5: Runge-Kutta and ConstructL in
                                           FP64, mbc >= (space_order + 1) / 2 */
    int    time_order;                  /* RK: 1..5 (1D and 2D); LxW: 2..3; PIF: 4 */
    int    time_stepping_method;        /* FINESS_TS_* */
    double initial_dt, max_dt, desired_cfl, max_cfl;
    int    nv;
    int    global_alpha;                /* 1: GlobalWaveSpd before every operator call and
                                           alpha = Max(global, local) at every interface
                                           (lib/2d/ConstructL.cpp:48-54,172,323); not with routing
                                           or the batched mode */
    int    source_term;                 /* 1: the SourceTermFunc values psi(q, aux) of the
                                           stage's state are added to L
                                           (lib/2d/ConstructL.cpp:371-387,
                                           lib/1d/ConstructL.cpp:194-205), psi being the
                                           device form named by source_kind; Runge-Kutta
                                           (1D, 2D, every order) and ConstructL.  Lax-Wendroff
                                           refuses it as the reference does
                                           (lib/2d/ConstructLxWL.cpp:376-380); RK4-PIF, routing
                                           and the batched mode refuse it too */
    int    verbosity;
    /* [weno] (python/finess/params/weno.py:15-51) */
    int    weno_version;                /* FINESS_WENO_* */
    double epsilon, power_param, alpha_scaling;
    /* physics */
    int    physics;                     /* FINESS_PHYS_* */
    double gamma;                       /* [euler] gamma (euler.py:13-19) */
    int    sw_compat_hl;                /* 1: keep the reference's u_r = (hu)_r / h_l */
    int    bc[4];                       /* left, right, bottom, top: FINESS_BC_* */
    /* device-path options (no counterpart in the reference) */
    int    precision;                   /* FINESS_PREC_* */
    int    routing;                     /* FINESS_ROUTE_* */
    double route_tau;                   /* label threshold on max_k |w_k - d_k| */
    int    tile_i, tile_j;              /* 0 = library default */
    int    route_per_step;              /* k >= 1: the indicator runs at the first stage of
                                           every k-th time step and its routing serves the
                                           later stages of that step and the next k-1 steps;
                                           0: it runs at every stage */
    int    host_controller;             /* 1: the host runs the time-step controller after
                                           every step; 0 (default): Runge-Kutta runs of
                                           time_order 1-3 (2D and 1D) without verbosity
                                           are controlled on the device
                                           (same decisions, no host round trip per step) */
    int    mpp_limiter;                 /* Lax-Wendroff, 2D Euler: limit the fluxes toward the
                                           Lax-Friedrichs fluxes so that density and pressure stay
                                           positive (lib/2d/FinSolveLxW.cpp:126-157); Runge-Kutta
                                           ignores it, as the reference does */
    int    source_kind;                 /* FINESS_SRC_*: the device form of SourceTermFunc */
    double source_coef;                 /* k of FINESS_SRC_DAMPING */
} finess_params;

/* Fills defaults exactly as the reference's IniParams defaults do. */
void finess_params_default(finess_params* p);
/* Recomputes dx, dy and validates the combination; returns a status code. */
int  finess_params_finalize(finess_params* p);

/* ---- context ------------------------------------------------------------ */
typedef struct finess_ctx finess_ctx;

/* Creates a context on the current HIP device.  *out receives the handle. */
int  finess_create(const finess_params* p, finess_ctx** out);
void finess_destroy(finess_ctx* ctx);
/* Last error message of this context (static storage inside the context). */
const char* finess_last_error(const finess_ctx* ctx);

/* ---- state transfer (reference layout, host memory) -------------------- */
/* Uploads q (and aux, may be NULL when maux==0 or unused) with ghosts; t. */
int finess_set_state(finess_ctx* ctx, const double* q, const double* aux, double t);
/* Downloads the state into q (reference layout; ghost cells are filled by the
 * configured boundary condition, corners included; FINESS_BC_HOST sides by
 * extrapolation, the callback running only before stages). */
int finess_get_state(finess_ctx* ctx, double* q, double* t);

/* SetBndValues (lib/2d/FinSolveRK.cpp:171-206 calls it before every stage)
 * for the sides whose bc is FINESS_BC_HOST.  Before each stage the library
 * calls fn(user, t, q) on the host: q is a reference-layout array with ghosts
 * ((mx+2mbc) x (my+2mbc) x meqn) whose interior cells within mbc of the
 * boundary hold the stage's source state and whose every other value is NaN;
 * t is the time UpdateSoln gives that state (lib/2d/UpdateSoln.cpp:20-27).  fn
 * fills the ghost cells as the app's SetBndValues does and returns 0.  The
 * library then takes the edge ghosts of the FINESS_BC_HOST sides (not the
 * corners, which the Runge-Kutta stencils never read, and nothing the callback
 * writes into interior cells); a non-finite one, such as
 * a value computed from a cell deeper than mbc, fails the step with
 * FINESS_ERR_CONFIG, as does a nonzero return.  fn = NULL removes it. */
typedef int (*finess_bnd_fn)(void* user, double t, double* q);
int finess_set_bnd_callback(finess_ctx* ctx, finess_bnd_fn fn, void* user);

/* ---- drop-in operators -------------------------------------------------- */

/* ConstructL (2D: lib/2d/ConstructL.cpp:13-408, 1D: lib/1d/ConstructL.cpp:16-224).
 * q, aux: host arrays in the reference layout with ghosts already filled by
 * the caller (as FinSolveRK does before each call).  Lstar receives the
 * interior; smax is max-accumulated (2D, dTensorBC3(mx+1,my+1,2,mbc)) or
 * overwritten (1D, dTensorBC1(mx,mbc)) exactly as the reference does. */
int finess_construct_l(finess_ctx* ctx, const double* q, const double* aux,
                       double* Lstar, double* smax);

/* FinSolveRK (lib/2d/FinSolveRK.cpp:13-371, lib/1d/FinSolveRK.cpp:27-363).
 * Advances q (host, reference layout) from *t to tend.  dtv[1] is the in/out
 * time step, dtv[2] the max_dt (reference convention, dtv has 3 entries). */
int finess_fin_solve_rk(finess_ctx* ctx, double* q, const double* aux,
                        double* t, double tend, double* dtv);

/* FinSolveLxW (lib/2d/FinSolveLxW.cpp:16-245, lib/1d/FinSolveLxW.cpp:20-218),
 * same conventions. */
int finess_fin_solve_lxw(finess_ctx* ctx, double* q, const double* aux,
                         double* t, double tend, double* dtv);

/* FinSolveRK_PIF (lib/2d/FinSolveRK_PIF.cpp:66-295; the reference links it as
 * FinSolveUser, :31-35), same conventions; a FINESS_TS_PIF context. */
int finess_fin_solve_rk_pif(finess_ctx* ctx, double* q, const double* aux,
                            double* t, double tend, double* dtv);

/* ConstructIntegratedR (lib/2d/ConstructIntegratedR.cpp:59-296): time-
 * integrated fluxes F, G (dTensorBC3(mx,my,meqn,mbc)) on cells [1-3, m+3].
 * 1D: ConstructIntegratedF (lib/1d/ConstructIntegratedF.cpp:54-212), F only
 * (dTensorBC2(mx,meqn,mbc), cells [1-3, mx+3]); G may be NULL. */
int finess_construct_integrated_r(finess_ctx* ctx, double dt, const double* q,
                                  const double* aux, double* F, double* G);

/* ConSoln after every accepted step (lib/2d/ConSoln.cpp:13-95, called from
 * lib/2d/FinSolveRK.cpp:361-362 and lib/2d/FinSolveLxW.cpp:228-229): with the
 * log on, each accepted step of the context's steppers records (t, then for
 * every field the sum over the interior of dx*dy*q (2D) or dx*q (1D)),
 * reduced on the device.  finess_read_step_log copies up to max_records
 * records (meqn+1 doubles each, oldest first) into rec, removes them from the
 * log and returns how many it copied. */
int finess_set_step_log(finess_ctx* ctx, int on);
int finess_read_step_log(finess_ctx* ctx, double* rec, int max_records);

/* ApplyMPPLimiter2D for 2D Euler (apps/2d/euler/lib/ApplyMPPLimiter2D.cpp:27-286;
 * an overridable object of the reference, lib/2d/Makefile.defs:109): limits the
 * high-order interface fluxes fHat (dTensorBC3(mx+1,my,meqn,mbc)) and gHat
 * (dTensorBC3(mx,my+1,meqn,mbc)) in place toward the Lax-Friedrichs fluxes fLF,
 * gLF (same shapes) so that the update of q (dTensorBC3(mx,my,meqn,mbc)) by dt
 * keeps density and pressure positive.  Interior faces only.  The context must
 * be a 2D Euler Lax-Wendroff context with mpp_limiter set; its state is kept. */
int finess_apply_mpp_limiter_2d(finess_ctx* ctx, double dt, const double* q,
                                const double* fLF, const double* gLF,
                                double* fHat, double* gHat);

/* ConstructL on device-resident field planes, allocating nothing.
 * No counterpart in the reference: it exists for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  It computes what finess_construct_l computes for
 * a 2D fifth-order context without routing, on device memory:
 *   q      device, field planes of the interior, q[m*mx*my + i*my + j]
 *          (0-based i, j); the halo is formed through p->bc
 *   L      device, same layout, receives L = Lstar of the interior
 *   smax2  device, 2 doubles: the maxima of the x- and y-face wave speeds that
 *          GetCFL takes (faces mx+1/2 and my+1/2 excluded,
 *          lib/2d/GetCFL.cpp:54-58)
 *   stream hipStream_t (NULL: the null stream)
 * Its only other device memory is a fixed 4 KiB array per device, of which
 * it uses 2064 bytes for any grid size; every call on a device shares it, so
 * calls that overlap on one device race.  The call is asynchronous on
 * `stream` and launches kernels only (no memset, copy or allocation). */
int finess_construct_l_planes_noalloc(const finess_params* p, const double* q, double* L, double* smax2,
                                      void* stream);

/* ---- device-resident stepping ------------------------------------------ */
typedef struct finess_stats {
    long long steps_accepted;
    long long steps_rejected;
    double    dt_min, dt_max;
    double    last_cfl;
    double    t;
} finess_stats;

/* Advances the state held by the context (set by finess_set_state) from its
 * current time to tend with the configured stepper, leaving it on the device.
 * dtv as above.  max_steps <= 0 means "nv".  stats may be NULL. */
int finess_advance(finess_ctx* ctx, double tend, double* dtv, finess_stats* stats);

/* Advances exactly nsteps accepted steps with the CFL controller (no frame
 * clipping, no nv limit).  dtv as above. */
int finess_advance_steps(finess_ctx* ctx, long long nsteps, double* dtv, finess_stats* stats);

/* Runs exactly nsteps accepted steps of size dt (no frame clipping, no CFL
 * rejection; the CFL of each step is still computed).  Used for timing. */
int finess_run_fixed_steps(finess_ctx* ctx, double dt, int nsteps, double* last_cfl);

/* Device synchronisation of the context's stream. */
int finess_sync(finess_ctx* ctx);

/* Kernel time of the stage launches of the last finess_run_fixed_steps call
 * (ms, each launch bracketed by HIP events on the context's stream) and the
 * number of launches. */
int finess_last_kernel_time(finess_ctx* ctx, double* ms, long long* launches);

/* ---- troubled-interface indicator (no counterpart in the reference) ------ */

/* Installs the network (finess_mlp parameter vector, FINESS_MLP_NPARAM
 * floats) on a context whose routing is FINESS_ROUTE_LEARNED: faces whose
 * logit exceeds `threshold` take characteristic WENO5, the others the
 * component-wise linear fifth-order flux. */
int finess_set_indicator(finess_ctx* ctx, const float* params, float threshold);
/* Troubled and total face counts accumulated since the last call (and reset). */
int finess_route_stats(finess_ctx* ctx, long long* troubled, long long* faces);

#define FINESS_MLP_NPARAM 1889
typedef struct finess_mlp finess_mlp;

/* A 24-32-32-1 network with FP32 master parameters (He-initialised from
 * `seed`) and a device training set of up to `capacity` interfaces. */
int  finess_mlp_create(unsigned long long seed, long long capacity, finess_mlp** out);
void finess_mlp_destroy(finess_mlp* mlp);
/* Appends every `stride`-th face of the context's current 2D Euler state:
 * features and the label [max weight departure > tau]. *count = set size. */
int  finess_mlp_harvest(finess_mlp* mlp, finess_ctx* ctx, int stride, double tau, long long* count);
/* Copies up to n samples of the set (X: n x 24, y: n, departure: n; any may be NULL). */
int  finess_mlp_get_data(finess_mlp* mlp, long long first, long long n, float* X, float* y, float* departure);
int  finess_mlp_set_data(finess_mlp* mlp, const float* X, const float* y, long long n);
/* Adam on the mean binary cross-entropy; loss[e] receives epoch e's mean loss. */
int  finess_mlp_train(finess_mlp* mlp, int epochs, int batch, float lr, float* loss);
int  finess_mlp_get_params(finess_mlp* mlp, float* params);
int  finess_mlp_set_params(finess_mlp* mlp, const float* params);
/* Gradient of the mean BCE loss over (X, y) (n <= 65536); *loss = mean loss. */
int  finess_mlp_gradient(finess_mlp* mlp, const float* X, const float* y, int n, float* grad, float* loss);
/* Logits of n interfaces: layer GEMMs (path 0) or the in-kernel wave routine (path 1). */
int  finess_mlp_eval(finess_mlp* mlp, const float* X, long long n, int path, float* logits);

/* ---- batched mode (no counterpart in the reference) ------------------------
 * Many independent 2D problems of uneven grid sizes (same physics, scheme and
 * controller settings, each on the base domain with its own mx x my) advanced
 * together: every stage of every problem is one persistent launch over the
 * tiles of all problems, each problem keeps its own a-posteriori CFL
 * controller on the device (accept/reject and dt from its own wave speeds),
 * and the step sequence is captured once in a HIP graph and replayed.  The
 * tile shape is chosen by finess_batch_autotune.  RK time stepping, 2D. */
typedef struct finess_batch finess_batch;

int  finess_batch_create(const finess_params* base, int nprob, const int* mx, const int* my, finess_batch** out);
void finess_batch_destroy(finess_batch* b);
const char* finess_batch_last_error(finess_batch* b);
/* Problem k's state (reference layout of its own mx x my grid), time and first dt. */
int  finess_batch_set_state(finess_batch* b, int k, const double* q, double t, double dt);
int  finess_batch_get_state(finess_batch* b, int k, double* q, double* t, double* dt, long long* accepted,
                            long long* rejected);
/* Times `trial_launches` stage launches for each candidate tile shape and
 * keeps the fastest (the state is not changed). */
int  finess_batch_autotune(finess_batch* b, int trial_launches, int* tile_i, int* tile_j);
/* Advances every problem by `steps` accepted steps (rejected ones are retried
 * with the controller's smaller dt); *seconds = wall time of the run. */
int  finess_batch_run(finess_batch* b, int steps, double* seconds);

/* ---- the device physics, evaluated on the host ----------------------------
 * No counterpart in the reference.  The kernels carry the app callbacks as
 * compiled device forms; these entry points evaluate the very same code on
 * the host, so that the drop-in layer can check an app's own callbacks
 * against them before it runs the app on the device (no GPU needed).
 *   finess_physics_eig         the matrices of ProjectLeftEig / ProjectRightEig
 *                              (apps/2d/euler/lib/ProjectLeftEig.cpp:9-67,
 *                              ProjectRightEig.cpp:9-63; shallow water alike) at the
 *                              mean state qavg, direction ixy (1: x, 2: y):
 *                              Lmat, Rmat are meqn x meqn, row-major, column j the
 *                              projection of the unit vector e_j
 *   finess_physics_wave_speed  max(|s1|, |s2|) of SetWaveSpd
 *                              (apps/2d/euler/lib/SetWaveSpd.cpp:11-69) for the
 *                              left and right states; 1D advection reads aux
 *   finess_physics_source      psi of one point for p->source_kind (SourceTermFunc) */
int finess_physics_eig(const finess_params* p, int ixy, const double* qavg, double* Lmat, double* Rmat);
int finess_physics_wave_speed(const finess_params* p, int ixy, const double* ql, const double* qr,
                              const double* auxl, const double* auxr, double* speed);
int finess_physics_source(const finess_params* p, const double* q, const double* aux, double* psi);

/* ---- library info --------------------------------------------------------- */
const char* finess_version(void);
int finess_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* FINESS_HIP_H */
