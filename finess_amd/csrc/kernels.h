// kernels.h -- host-side launchers of the gfx950 kernels (internal to the
// library; the public interface is include/finess_hip.h).
#pragma once
#include <hip/hip_runtime.h>

namespace finess {

enum : int { PHYS_ADVECTION = 0, PHYS_EULER = 1, PHYS_SHALLOW_WATER = 2 };

struct StageArgs;
struct Stage1DArgs;
struct Stage1DStep;
struct LxwArgs;

// 2D fused stage.  wk: WenoKind; fp32: stencil arithmetic in FP32.
int launch_stage2d(int physics, int fp32, int wk, int lxw, const StageArgs& a, hipStream_t st);
// WENO orders other than 5 (face-list kernels, FP64, Runge-Kutta)
int launch_stage2d_wide(int physics, int wk, const StageArgs& a, hipStream_t st);
// 1D stage (one workgroup up to 8192 cells, else two launches: faces, update).
int launch_stage1d(int wk, const Stage1DArgs& a, hipStream_t st);
// The stages of one 1D time step (one launch up to 8192 cells).
int launch_step1d(int wk, const Stage1DStep& s, hipStream_t st);
// A run of 1D SSP steps with the device controller in one single-workgroup
// launch (grids that fit it: run1d_fits).
struct Stage1DRun;
int launch_run1d(int wk, const Stage1DRun& r, hipStream_t st);
bool run1d_fits(int wk, int mx);
// 1D Lax-Wendroff time-integrated fluxes of the advection flux on [-3, mx+2].
int launch_lxw_flux1d(const Stage1DArgs& a, double dt, int time_order, double* F, hipStream_t st);
// Lax-Wendroff time-integrated fluxes on [-3, mx+2] x [-3, my+2]; fp32:
// computed in FP32 (the FP32 stencil form, Euler).
int launch_lxw_flux(int physics, int fp32, const LxwArgs& a, hipStream_t st);

int launch_stage2d_routed(int physics, int fp32, int wk, const StageArgs& a, hipStream_t st);
// tiles of the routed kernels (records of the per-step routing record)
long long routed_tiles(int mx, int my);

struct HarvestArgs;
struct MlpFrags;
int launch_harvest(const HarvestArgs& a, hipStream_t st);
int launch_mlp_eval(const float* X, long long n, const MlpFrags* fr, float* logits, hipStream_t st);
int mlp_forward(const float* params, const float* X, int B, float* H1, float* H2, float* z, hipStream_t st);
int mlp_backward(const float* params, const float* X, const float* y, int B, float* H1, float* H2, float* z,
                 float* dz, float* dH2, float* dH1, float* grad, float* loss_sum, hipStream_t st);
int mlp_adam(float* p, float* g, float* m, float* v, float lr, int t, hipStream_t st);
int mlp_zero(float* p, long long n, hipStream_t st);

int launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st);

// Conservation sums of a state (step_control.hip): record (t, sum of w*q per
// field) into log; with ctl, only for an accepted attempt not yet recorded.
struct StepCtl;
int launch_con_soln(const double* q, long long plane, int M, double w, double* partial, StepCtl* ctl,
                    long long base, double t, double* log, hipStream_t st);

// Positivity-preserving limiter of the 2D Lax-Wendroff step, Euler
// (mpp2d.hip): from the high-order interface fluxes fh, gh (x faces
// [m][(mx+1) my], y faces [m][mx (my+1)]) to the limited update q -> out.
struct MppLaunch {
    const double* q;
    double* out;
    int mx, my;
    int bc[4];
    double dx, dy, dt, gamma;
    double *fh, *gh, *fl, *gl, *amin;
    unsigned long long *gmax, *smax;
};
int launch_mpp(const MppLaunch& L, hipStream_t st);
// the limiter alone: fh, gh limited in place from fl, gl (q, dt given)
int launch_mpp_limit(const MppLaunch& L, hipStream_t st);

// RK4-PIF stages, 2D Euler (pif2d.hip): the weighted flux sums on the
// extended grid, and one non-characteristic stage q -> out = qold + cdt L_NOC.
struct PifLaunch {
    const double* q;
    const double* qold;
    double* out;
    int mx, my;
    int bc[4];
    double dx, dy, cdt, w;
    int first, last;
    double eps, power, gamma;
    int physics, sw_compat_hl;
    double *F, *G, *fx, *fy;
};
int launch_pif_flux_sum(const PifLaunch& L, hipStream_t st);

// Elementwise parts of the 10-stage fourth-order and 8-stage fifth-order
// Runge-Kutta schemes (rk_low_storage.hip); n = meqn * cells.
int launch_rk4_combine(const double* qn, double* q2, double* q1, long long n, hipStream_t st);
int launch_rk5_update(const double* q1, double* q2, const double* qold, const double* L, double* out, long long n,
                      double g1, double g2, double g3, double delta, double bdt, int first, hipStream_t st);
int launch_pif_noc_stage(int wk, const PifLaunch& L, hipStream_t st);

// Device-side time-step controller (step_control.hip).
struct StepCtl;
struct StepCtlParams;
int launch_step_control(StepCtl* c, unsigned long long* smax, const StepCtlParams& p, hipStream_t st);
int launch_ref_to_planes(const double* src, double* dst, int mx, int my, int meqn, int g, int dg, hipStream_t st);
int launch_planes_to_ref(const double* src, double* dst, int mx, int my, int meqn, int g, const int bc[4],
                         int nx_mom, int ny_mom, hipStream_t st);
// BC_HOST exchange: the interior edge ring of the planes, packed; the
// ghosted planes with BC_HOST edge ghosts from packed strips
int launch_ring_pack(const double* src, double* ring, int mx, int my, int meqn, int g, hipStream_t st);
int launch_ghost_fill(const double* src, const double* ghost, double* dst, int mx, int my, int meqn, int g,
                      const int bc[4], int nx_mom, int ny_mom, hipStream_t st);

} // namespace finess
