// capi.cpp -- host controller and C-ABI (include/finess_hip.h).
//
// The state lives on the device as field planes q[m][i][j] (interior cells
// only; the halo is produced by each stage kernel through the boundary map).
// SSP-RK3 uses three plane sets A, B, C with A = q^n never written during a
// step:
//     stage 1: B = A + dt L(A)
//     stage 2: C = 3/4 A + 1/4 B + 1/4 dt L(B)
//     stage 3: B = 2/3 C + 1/3 A + 2/3 dt L(C)
// (coefficients of lib/2d/SetRKinfo.cpp:50-64, operand order of
// lib/2d/FinSolveRK.cpp:158-211), so a rejected step needs no copy back:
// the controller simply retries from A.  Accepting a step swaps the roles.
//
// The time-step controller is the reference's a-posteriori one
// (lib/2d/FinSolveRK.cpp:77-367, lib/1d/FinSolveRK.cpp:88-363,
// lib/2d/FinSolveLxW.cpp:100-229): frame clipping, CFL from the per-face wave
// speeds of the step just taken, dt = min(max_dt, dt*desired_cfl/cfl),
// reject when cfl > max_cfl.
#include <hip/hip_runtime.h>

#include <math.h>
#include <cmath>
#include <limits>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/finess_hip.h"
#include "kernels.h"
#include "lxw_body.h"
#include "mlp.h"
#include "stage1d_body.h"
#include "stage2d_body.h"
#include "step_control.h"

using namespace finess;

struct finess_ctx {
    finess_params p;
    int device = 0;
    hipStream_t st = nullptr;
    int M = 0;
    long long plane = 0;        // cells of one field plane
    // plane sets: 0-2 the rotating SSP-RK / LxW states; 3, 4 the extra
    // registers of the fourth- and fifth-order schemes (q2; L)
    double* buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double* aux = nullptr;      // 1D advection speed (interior cells)
    double* aux2d = nullptr;    // 2D aux planes (interior cells), for the source term
    double* lxwF = nullptr;
    double* lxwG = nullptr;
    double* staging = nullptr;  // reference layout with ghosts
    double* ghosted = nullptr;  // field planes with ghosts (drop-in operators)
    double* faces_x = nullptr;  // per-face wave speeds (drop-in ConstructL)
    double* faces_y = nullptr;
    double* fhat1d = nullptr;
    unsigned long long* smax = nullptr;    // device CFL slots [NSLOT][2][CFL_SPREAD]
    unsigned long long* smax_h = nullptr;  // pinned copy
    MlpFrags* frags = nullptr;                    // device copy of the installed network
    unsigned long long* route_counts = nullptr;   // device [2][CFL_SPREAD]
    // routing record of the current step: per-tile troubled-face masks,
    // troubled list, its length, and the troubled faces' fluxes of the
    // current stage (in the stencil precision)
    RouteTileRec* route_tile = nullptr;
    int* route_list = nullptr;
    unsigned long long* route_n = nullptr;
    void* route_flux = nullptr;
    // limited Lax-Wendroff: interface fluxes (high-order, then limited in
    // place; Lax-Friedrichs), per-cell fractions, global wave-speed words
    double *mpp_fh = nullptr, *mpp_gh = nullptr, *mpp_fl = nullptr, *mpp_gl = nullptr, *mpp_amin = nullptr;
    // interface fluxes of the face-list stage (WENO orders other than 5)
    double *wide_fx = nullptr, *wide_fy = nullptr;
    // global_alpha: the two global wave speeds and their spread words
    double* galpha = nullptr;
    unsigned long long* gspread = nullptr;
    unsigned long long* mpp_gmax = nullptr;
    double *pif_fx = nullptr, *pif_fy = nullptr;  // RK4-PIF interface fluxes of a non-characteristic stage
    // per-step conservation log (ConSoln after every accepted step)
    bool steplog = false;
    double* slog = nullptr;      // device records (t, sums) of one chunk
    double* spart = nullptr;     // partial sums
    std::vector<double> hlog;    // records not yet read by the caller
    StepCtl* ctl = nullptr;                       // device controller block
    bool devctl = false;                          // the device controller drives the current run
    const StepCtlParams* ctlp = nullptr;          // its parameters while devctl
    bool ctl_enqueued = false;                    // enqueue_step launched the controller with the step
    // the last step's controller, not launched yet: it runs in the next
    // stage's troubled-face pass, or on its own before an indicator launch
    // or the chunk's read-back (flush_ctl)
    bool ctl_pending = false;
    int ctl_pending_slot = 0;
    bool route_fresh = true;                      // next routed stage starts a step
    int rec_age = -1;                             // steps since the record was decided (-1: none)
    int rec_par = 0;                              // counter pair of the current record
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> tev;  // per-launch timing marks (run_fixed_steps)
    size_t tev_used = 0;
    bool timing = false;
    // FINESS_BC_HOST sides: the host SetBndValues, the device ring and ghost
    // strips (packed side after side), their host copies, and the time
    // UpdateSoln gives each plane set (set before every stage)
    bool hostbc = false;
    finess_bnd_fn bnd_fn = nullptr;
    void* bnd_user = nullptr;
    double* edge = nullptr;  // [2][edge_numel]: interior ring, then the ghosts
    std::vector<double> edge_h, bnd_q;
    double buf_t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double step_t0 = 0.0;    // time of the state a step starts from
    double t = 0.0;
    int cur = 0;
    bool have_state = false;
    double last_ms = 0.0;
    long long last_launches = 0;
    std::string err;
};

namespace {

constexpr int NSLOT = 8;

int fail(finess_ctx* c, int code, const char* msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, x)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            if (ctx) (ctx)->err = std::string(#x ": ") + hipGetErrorString(e_);            \
            return FINESS_ERR_HIP;                                                          \
        }                                                                                   \
    } while (0)

int weno_kind(const finess_params& p)
{
    if (p.space_order == 1) return WENO_O1;
    if (p.space_order == 7)
        return p.weno_version == FINESS_WENO_FD ? WENO_FD7 : p.weno_version == FINESS_WENO_Z ? WENO_Z7 : WENO_JS7;
    if (p.space_order == 9)
        return p.weno_version == FINESS_WENO_FD ? WENO_FD9 : p.weno_version == FINESS_WENO_Z ? WENO_Z9 : WENO_JS9;
    if (p.space_order == 11) return WENO_JS11;
    if (p.weno_version == FINESS_WENO_FD) return WENO_FD5;
    if (p.weno_version == FINESS_WENO_Z) return WENO_Z5;
    return p.power_param == 2.0 ? WENO_JS5_P2 : WENO_JS5_POW;
}

size_t ref_numel(const finess_params& p, int fields)
{
    if (p.ndims == 1) return (size_t)(p.mx + 2 * p.mbc) * fields;
    return (size_t)(p.mx + 2 * p.mbc) * (p.my + 2 * p.mbc) * fields;
}

int normal_x(const finess_params& p) { return p.physics == FINESS_PHYS_ADVECTION ? -1 : 1; }
int normal_y(const finess_params& p) { return p.physics == FINESS_PHYS_ADVECTION ? -1 : 2; }

double bits_to_double(unsigned long long b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

StageArgs base_stage_args(finess_ctx* c)
{
    StageArgs a;
    memset(&a, 0, sizeof a);
    const finess_params& p = c->p;
    a.mx = p.mx;
    a.my = p.my;
    a.plane = c->plane;
    for (int k = 0; k < 4; k++) a.bc[k] = p.bc[k];
    a.dx = p.dx;
    a.dy = p.dy;
    a.idx = 1.0 / p.dx;
    a.idy = 1.0 / p.dy;
    a.alpha_scaling = p.alpha_scaling;
    a.eps = p.epsilon;
    a.power = p.power_param;
    a.pp.gamma = p.gamma;
    a.pp.sw_compat_hl = p.sw_compat_hl;
    if (p.space_order != 5) {  // the face-list stage's interface fluxes
        a.face_x = c->wide_fx;
        a.face_y = c->wide_fy;
    }
    if (p.global_alpha) {
        a.galpha = c->galpha;
        a.gspread = c->gspread;
    }
    a.src_kind = source_arg(p.source_term, p.source_kind);
    a.src_coef = p.source_coef;
    a.aux = c->aux2d;
    return a;
}

Stage1DArgs base_stage1d_args(finess_ctx* c)
{
    Stage1DArgs a;
    memset(&a, 0, sizeof a);
    const finess_params& p = c->p;
    a.mx = p.mx;
    a.bc_lo = p.bc[0];
    a.bc_hi = p.bc[1];
    a.dx = p.dx;
    a.alpha_scaling = p.alpha_scaling;
    a.eps = p.epsilon;
    a.power = p.power_param;
    a.aux = c->aux;
    if (p.global_alpha) a.galpha = c->galpha;
    a.fhat = c->fhat1d;
    a.src_kind = source_arg(p.source_term, p.source_kind);
    a.src_coef = p.source_coef;
    return a;
}

// stepping: the stage belongs to a time step (the first stage of a step runs
// the indicator and records its routing for the later ones when
// route_per_step is set); otherwise a standalone ConstructL.
// Records the next timing mark on the stream (stage launches are bracketed
// by two marks while run_fixed_steps measures them).
int tmark(finess_ctx* c)
{
    if (!c->timing) return 0;
    if (c->tev_used == c->tev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return fail(c, FINESS_ERR_HIP, "event creation failed");
        c->tev.push_back(e);
    }
    if (hipEventRecord(c->tev[c->tev_used++], c->st) != hipSuccess) return fail(c, FINESS_ERR_HIP, "event record failed");
    return 0;
}

#ifndef FINESS_FOLD_CTL
#define FINESS_FOLD_CTL 1
#endif

// Launches the pending controller on its own.
int flush_ctl(finess_ctx* c)
{
    if (!c->ctl_pending) return 0;
    c->ctl_pending = false;
    if (launch_step_control(c->ctl, c->smax + 2 * CFL_SPREAD * c->ctl_pending_slot, *c->ctlp, c->st))
        return fail(c, FINESS_ERR_HIP, "controller launch failed");
    return 0;
}

int stage_routed(finess_ctx* c, StageArgs& a, bool stepping)
{
    const finess_params& p = c->p;
    if (p.time_stepping_method != FINESS_TS_RK) return fail(c, FINESS_ERR_CONFIG, "routing needs the RK stepper");
    if (p.routing == FINESS_ROUTE_LEARNED && !c->frags)
        return fail(c, FINESS_ERR_CONFIG, "routing = learned but no indicator installed (finess_set_indicator)");
    int rc0 = 0;
    a.frags = c->frags;
    a.route_exact = p.routing == FINESS_ROUTE_ORACLE;
    a.route_tau = p.route_tau;
    a.route_counts = c->route_counts;
    if (stepping && p.route_per_step && c->route_tile) {
        a.route_tile = c->route_tile;
        a.route_list = c->route_list;
        a.route_flux = c->route_flux;
        bool decide = c->route_fresh && (c->rec_age < 0 || c->rec_age >= p.route_per_step);
        a.route_mask_read = decide ? 0 : 1;
        // two counter pairs: a new record counts into the pair the previous
        // indicator stage cleared and clears the other one
        if (decide) {
            a.route_n = c->route_n + 2 * (1 - c->rec_par);
            a.route_n_clear = c->route_n + 2 * c->rec_par;
            c->rec_par = 1 - c->rec_par;
            c->rec_age = 0;
        } else {
            a.route_n = c->route_n + 2 * c->rec_par;
        }
        c->route_fresh = false;
        // the previous step's controller: with the troubled-face pass of a
        // stage that reuses the record, else before the indicator
        if (c->ctl_pending) {
            if (a.route_mask_read) {
                a.ctl_run = c->ctl;
                a.ctl_smax = c->smax + 2 * CFL_SPREAD * c->ctl_pending_slot;
                a.ctl_p = *c->ctlp;
                c->ctl_pending = false;
            } else if (const int rc = flush_ctl(c)) {
                return rc;
            }
        }
    }
    if (c->ctl_pending && (rc0 = flush_ctl(c))) return rc0;
    const int rc = launch_stage2d_routed(p.physics, p.precision == FINESS_PREC_FP32, weno_kind(p), a, c->st);
    return rc ? fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "routed stage launch failed") : 0;
}

// The arguments of one 1D stage: out = a1*qa + a2*qb + (beta*dt) L(src).
Stage1DArgs stage1d_args(finess_ctx* c, int src, int qa, int qb, int out, double a1, double a2, double beta,
                         double dt, int slot)
{
    Stage1DArgs a = base_stage1d_args(c);
    a.src = c->buf[src];
    a.qa = c->buf[qa];
    a.qb = qb >= 0 ? c->buf[qb] : nullptr;
    a.use_b = (qb >= 0 && a2 != 0.0);
    a.out = c->buf[out];
    a.a1 = a1;
    a.a2 = a2;
    a.bdt = beta * dt;
    // 1D GetCFL reads the last stage's speeds only (lib/1d/ConstructL.cpp:149
    // overwrites smax): the other stages pass slot -1 and reduce nothing
    a.smax_slot = slot >= 0 ? c->smax + 2 * CFL_SPREAD * slot : nullptr;
    if (c->p.time_stepping_method == FINESS_TS_LXW) a.lxwF = c->lxwF;  // the 1D ConstructLxWL
    if (c->devctl) {
        a.dtp = &c->ctl->dt;
        a.beta = beta;
        a.halt = &c->ctl->halt;
    }
    return a;
}

int alloc(finess_ctx* c, void** ptr, size_t bytes);
int ensure_ghosted(finess_ctx* c);

// Edge values of the ring (or of the ghost strips): 2 meqn g (mx + my).
size_t edge_numel(const finess_params& p, int M) { return 2 * (size_t)M * p.mbc * ((size_t)p.mx + p.my); }

// SetBndValues on the host for the FINESS_BC_HOST sides (lib/2d/FinSolveRK.cpp:
// 171-206 calls it before every stage), then the stage's ghosted source in
// c->ghosted: the callback sees the interior cells within mbc of the boundary
// (the rest NaN) in the reference layout, at time t; the edge ghosts it writes
// on those sides are uploaded, and the other sides keep the device rule.
int host_boundary(finess_ctx* c, int src, double t)
{
    const finess_params& p = c->p;
    if (!c->bnd_fn) return fail(c, FINESS_ERR_CONFIG, "a FINESS_BC_HOST side needs finess_set_bnd_callback");
    const int g = p.mbc, M = c->M, mx = p.mx, my = p.my, ld = my + 2 * g;
    const size_t ne = edge_numel(p, M), nx = (size_t)M * g * my, ny = (size_t)M * g * mx;
    int rc;
    if (!c->edge && (rc = alloc(c, (void**)&c->edge, 2 * ne * sizeof(double)))) return rc;
    if ((rc = ensure_ghosted(c))) return rc;
    if (launch_ring_pack(c->buf[src], c->edge, mx, my, M, g, c->st))
        return fail(c, FINESS_ERR_HIP, "edge ring launch failed");
    c->edge_h.resize(2 * ne);
    HIPCHK(c, hipMemcpyAsync(c->edge_h.data(), c->edge, ne * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->bnd_q.assign(ref_numel(p, M), std::numeric_limits<double>::quiet_NaN());
    // the reference-layout index of edge slot (side, m, layer, along): for
    // the ring (ghost = false) the l-th interior cell from the side, for the
    // ghosts the l-th cell beyond it
    auto ref_at = [&](int side, int m, int l, int a, bool ghost) -> size_t {
        const int in = ghost ? -1 - l : l;
        const int i = side == 0 ? in : side == 1 ? mx - 1 - in : a;
        const int j = side == 2 ? in : side == 3 ? my - 1 - in : a;
        return ((size_t)(i + g) * ld + (size_t)(j + g)) * M + m;
    };
    for (int side = 0; side < 4; side++) {
        const int n = side < 2 ? my : mx;
        const size_t off = side < 2 ? side * nx : 2 * nx + (side - 2) * ny;
        for (int m = 0; m < M; m++)
            for (int l = 0; l < g; l++)
                for (int a = 0; a < n; a++)
                    c->bnd_q[ref_at(side, m, l, a, false)] = c->edge_h[off + ((size_t)m * g + l) * n + a];
    }
    if (c->bnd_fn(c->bnd_user, t, c->bnd_q.data()))
        return fail(c, FINESS_ERR_CONFIG, "the boundary callback returned an error");
    double* gh = c->edge_h.data() + ne;
    for (int side = 0; side < 4; side++) {
        const int n = side < 2 ? my : mx;
        const size_t off = side < 2 ? side * nx : 2 * nx + (side - 2) * ny;
        for (int m = 0; m < M; m++)
            for (int l = 0; l < g; l++)
                for (int a = 0; a < n; a++) {
                    double v = 0.0;
                    if (p.bc[side] == FINESS_BC_HOST) {
                        v = c->bnd_q[ref_at(side, m, l, a, true)];
                        if (!std::isfinite(v))
                            return fail(c, FINESS_ERR_CONFIG,
                                        "the boundary callback left a non-finite ghost value (a rule reading cells "
                                        "deeper than mbc from the boundary is not supported)");
                    }
                    gh[off + ((size_t)m * g + l) * n + a] = v;
                }
    }
    HIPCHK(c, hipMemcpyAsync(c->edge + ne, gh, ne * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (launch_ghost_fill(c->buf[src], c->edge + ne, c->ghosted, mx, my, M, g, p.bc, normal_x(p), normal_y(p), c->st))
        return fail(c, FINESS_ERR_HIP, "ghost fill launch failed");
    // the upload reads the host copy: it must land before the next stage
    // rewrites that copy
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

// UpdateSoln's time, alpha1 tstar + alpha2 tnew + beta dt
// (lib/2d/UpdateSoln.cpp:27), rounded after every operation as the
// reference's build does (no fused multiply-add)
double update_time(double a1, double ta, bool use_b, double a2, double tb, double beta, double dt)
{
#pragma clang fp contract(off)
    double tn = a1 * ta;
    if (use_b) tn = tn + a2 * tb;
    return tn + beta * dt;
}

// One stage: out = a1*qa + a2*qb + (beta*dt) L(src).  slot: CFL slot.
int stage(finess_ctx* c, int src, int qa, int qb, int out, double a1, double a2, double beta, double dt, int slot)
{
    const finess_params& p = c->p;
    if (p.ndims == 1) {
        const Stage1DArgs a = stage1d_args(c, src, qa, qb, out, a1, a2, beta, dt, slot);
        if (tmark(c)) return FINESS_ERR_HIP;
        const int rc = launch_stage1d(weno_kind(p), a, c->st);
        if (tmark(c)) return FINESS_ERR_HIP;
        return rc ? fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "1D stage launch failed") : 0;
    }
    StageArgs a = base_stage_args(c);
    a.src = c->buf[src];
    a.qa = c->buf[qa];
    a.a_is_src = (qa == src);
    a.qb = qb >= 0 ? c->buf[qb] : nullptr;
    a.b_is_src = (qb == src);
    a.use_b = (qb >= 0 && a2 != 0.0);
    a.out = c->buf[out];
    a.a1 = a1;
    a.a2 = a2;
    a.bdt = beta * dt;
    a.smax = slot >= 0 ? c->smax + 2 * CFL_SPREAD * slot : nullptr;
    if (c->devctl) {
        a.dtp = &c->ctl->dt;
        a.beta = beta;
        a.halt = &c->ctl->halt;
    }
    // Lax-Wendroff and the final RK4-PIF stage read their stencil fluxes from
    // lxwF / lxwG (time-integrated, resp. time-averaged fluxes)
    const bool lxw = p.time_stepping_method != FINESS_TS_RK;
    if (lxw) {
        a.lxw_F = c->lxwF;
        a.lxw_G = c->lxwG;
    }
    int rc;
    if (c->hostbc) {
        // the stage reads its ghosted source; the source's time is the one
        // UpdateSoln formed (alpha1 tstar + alpha2 tnew + beta dt,
        // lib/2d/UpdateSoln.cpp:20-27)
        if ((rc = host_boundary(c, src, c->buf_t[src]))) return rc;
        a.src = c->ghosted;
        a.src_ghosts = p.mbc;
        c->buf_t[out] = update_time(a1, c->buf_t[qa], qb >= 0, a2, qb >= 0 ? c->buf_t[qb] : 0.0, beta, dt);
    }
    if (tmark(c)) return FINESS_ERR_HIP;
    if (p.routing != FINESS_ROUTE_NONE) {
        rc = stage_routed(c, a, true);
        if (!rc && tmark(c)) return FINESS_ERR_HIP;
        return rc;
    }
    rc = launch_stage2d(p.physics, p.precision == FINESS_PREC_FP32, weno_kind(p), lxw, a, c->st);
    if (tmark(c)) return FINESS_ERR_HIP;
    return rc ? fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "2D stage launch failed") : 0;
}

// Bytes per stored time-integrated flux value: the 2D kernels keep them in
// their arithmetic (FP32 in the FP32 form); the 1D path in FP64.
static size_t lxw_flux_bytes(const finess_params& p)
{
    return p.ndims == 2 && p.precision == FINESS_PREC_FP32 ? sizeof(float) : sizeof(double);
}

int lxw_fluxes(finess_ctx* c, int src, double dt)
{
    const finess_params& p = c->p;
    if (p.ndims == 1) {  // ConstructIntegratedF (lib/1d/ConstructIntegratedF.cpp:54-212)
        Stage1DArgs a = base_stage1d_args(c);
        a.src = c->buf[src];
        const int rc = launch_lxw_flux1d(a, dt, p.time_order, c->lxwF, c->st);
        return rc ? fail(c, FINESS_ERR_HIP, "1D LxW flux launch failed") : 0;
    }
    LxwArgs a;
    memset(&a, 0, sizeof a);
    a.q = c->buf[src];
    a.F = c->lxwF;
    a.G = c->lxwG;
    a.mx = p.mx;
    a.my = p.my;
    a.plane = c->plane;
    for (int k = 0; k < 4; k++) a.bc[k] = p.bc[k];
    a.time_order = p.time_order;
    a.dt = dt;
    a.dx = p.dx;
    a.dy = p.dy;
    a.pp.gamma = p.gamma;
    a.pp.sw_compat_hl = p.sw_compat_hl;
    const int rc = launch_lxw_flux(p.physics, p.precision == FINESS_PREC_FP32, a, c->st);
    return rc ? fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "LxW flux launch failed") : 0;
}

// The limited Lax-Wendroff step (lib/2d/FinSolveLxW.cpp:126-157): the stage
// kernel writes the high-order interface fluxes (no update), then the
// limiter (mpp2d.hip) forms the Lax-Friedrichs fluxes, limits and updates
// buf[src] into buf[dst].
int mpp_step(finess_ctx* c, int src, int dst, double dt)
{
    const finess_params& p = c->p;
    StageArgs a = base_stage_args(c);
    a.src = c->buf[src];
    a.qa = c->buf[src];
    a.a_is_src = 1;
    a.a1 = 1.0;
    a.bdt = dt;
    a.smax = c->smax;
    a.lxw_F = c->lxwF;
    a.lxw_G = c->lxwG;
    a.face_x = c->mpp_fh;
    a.face_y = c->mpp_gh;
    if (tmark(c)) return FINESS_ERR_HIP;
    int rc = launch_stage2d(p.physics, p.precision == FINESS_PREC_FP32, weno_kind(p), 1, a, c->st);
    if (rc) return fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "LxW face-flux launch failed");
    MppLaunch m;
    m.q = c->buf[src];
    m.out = c->buf[dst];
    m.mx = p.mx;
    m.my = p.my;
    for (int k = 0; k < 4; k++) m.bc[k] = p.bc[k];
    m.dx = p.dx;
    m.dy = p.dy;
    m.dt = dt;
    m.gamma = p.gamma;
    m.fh = c->mpp_fh;
    m.gh = c->mpp_gh;
    m.fl = c->mpp_fl;
    m.gl = c->mpp_gl;
    m.amin = c->mpp_amin;
    m.gmax = c->mpp_gmax;
    m.smax = c->smax;
    if (launch_mpp(m, c->st)) return fail(c, FINESS_ERR_HIP, "limiter launch failed");
    if (tmark(c)) return FINESS_ERR_HIP;
    return 0;
}

// One RK4-PIF step (lib/2d/FinSolveRK_PIF.cpp:151-206) from buf[A]: three
// non-characteristic stages A -> B -> C -> B with the flux samples summed into
// lxwF / lxwG, then the characteristic stage on the averaged fluxes into C.
int pif_step(finess_ctx* c, int A, int B, int C, double dt)
{
    const finess_params& p = c->p;
    PifLaunch L;
    memset(&L, 0, sizeof L);
    L.qold = c->buf[A];
    L.mx = p.mx;
    L.my = p.my;
    for (int k = 0; k < 4; k++) L.bc[k] = p.bc[k];
    L.dx = p.dx;
    L.dy = p.dy;
    L.eps = p.epsilon;
    L.power = p.power_param;
    L.gamma = p.gamma;
    L.physics = p.physics;
    L.sw_compat_hl = p.sw_compat_hl;
    L.F = c->lxwF;
    L.G = c->lxwG;
    L.fx = c->pif_fx;
    L.fy = c->pif_fy;
    auto sum = [&](int src, double w, int first, int last) {
        L.q = c->buf[src];
        L.w = w;
        L.first = first;
        L.last = last;
        return launch_pif_flux_sum(L, c->st);
    };
    auto noc = [&](int src, int dst, double cdt) {
        L.q = c->buf[src];
        L.out = c->buf[dst];
        L.cdt = cdt;
        return launch_pif_noc_stage(weno_kind(p), L, c->st);
    };
    if (tmark(c)) return FINESS_ERR_HIP;
    if (sum(A, 1.0, 1, 0) || noc(A, B, 0.5 * dt) || sum(B, 2.0, 0, 0) || noc(B, C, 0.5 * dt) ||
        sum(C, 2.0, 0, 0) || noc(C, B, dt) || sum(B, 1.0, 0, 1))
        return fail(c, FINESS_ERR_HIP, "PIF stage launch failed");
    if (tmark(c)) return FINESS_ERR_HIP;
    return stage(c, A, A, -1, C, 1.0, 0.0, 1.0, dt, 0);
}

// Coefficients of the 8-stage fifth-order scheme as published in
// lib/2d/SetRKinfo.cpp:84-127 (delta, gamma rows 1-3, beta).
const double RK5_DELTA[8] = {1.0e0, 1.528486658778845e00, 4.720094096662784e-02, 8.801244253465348e-01,
                             1.019066090228480e+00, 1.049772291176110e+01, -4.254616508506826e+00, 0.0};
const double RK5_GAMMA[3][8] = {
    {0.0, -1.552288007713033e+01, 4.127286635722417e-01, -1.011819196331377e+00, -2.765748383780848e-01,
     5.075770311217778e-02, 6.999810478513669e+00, -1.114908881433104e+01},
    {1.0, 6.534691420958578e+00, 2.280056542904473e-01, 1.308684311397668e+00, 4.769419552531064e-01,
     -6.368809762042849e-03, 9.339446057238532e-02, 9.556626047962331e-01},
    {0.0, 0.0, 0.0, -2.510747784045939e+00, -8.576822794622042e-01, 1.044599944472272e+00, -7.000810861049136e+00,
     1.906311811144179e+00}};
const double RK5_BETA[8] = {8.653258038183180e-02, 9.544677980851571e-01, 2.651941386774408e-01,
                            2.736914413910379e-01, 5.999778649323600e-01, 4.433177471748104e-03,
                            5.309971130968292e-03, 5.830861806762871e-01};

// L(buf[src]) into Lbuf, with the wave-speed maxima into `slot` (-1: none).
int stage_l(finess_ctx* c, int src, double* Lbuf, int slot)
{
    const finess_params& p = c->p;
    if (p.ndims == 1) {
        Stage1DArgs a = base_stage1d_args(c);
        a.src = c->buf[src];
        a.qa = c->buf[src];
        a.Lout = Lbuf;
        a.smax_slot = slot >= 0 ? c->smax + 2 * CFL_SPREAD * slot : nullptr;
        if (tmark(c)) return FINESS_ERR_HIP;
        const int rc = launch_stage1d(weno_kind(p), a, c->st);
        if (tmark(c)) return FINESS_ERR_HIP;
        return rc ? fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "1D stage launch failed") : 0;
    }
    StageArgs a = base_stage_args(c);
    a.src = c->buf[src];
    a.qa = c->buf[src];
    a.a_is_src = 1;
    a.smax = slot >= 0 ? c->smax + 2 * CFL_SPREAD * slot : nullptr;
    a.Lout = Lbuf;
    if (tmark(c)) return FINESS_ERR_HIP;
    const int rc = launch_stage2d(p.physics, p.precision == FINESS_PREC_FP32, weno_kind(p), 0, a, c->st);
    if (tmark(c)) return FINESS_ERR_HIP;
    return rc ? fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "2D stage launch failed") : 0;
}

// Enqueues one full time step from buf[cur]; returns the index holding the
// result in *res and the slot holding the CFL speeds in *slot.
int enqueue_step(finess_ctx* c, double dt, int* res, int* slot)
{
    const finess_params& p = c->p;
    const int A = c->cur, B = (c->cur + 1) % 3, C = (c->cur + 2) % 3;
    // (under the device controller the maxima are cleared by the controller)
    int rc = c->devctl ? 0 : launch_fill_u64(c->smax, 2 * CFL_SPREAD * NSLOT, 0ull, c->st);
    if (rc) return fail(c, FINESS_ERR_HIP, "slot reset failed");
    if (c->rec_age >= 0 && c->route_fresh == false) c->rec_age++;  // the previous step used the record
    c->route_fresh = true;
    c->buf_t[A] = c->step_t0;
    if (p.time_stepping_method == FINESS_TS_PIF) {
        if ((rc = pif_step(c, A, B, C, dt))) return rc;
        *res = C;
        *slot = 0;
        return 0;
    }
    if (p.time_stepping_method == FINESS_TS_LXW) {
        if ((rc = lxw_fluxes(c, A, dt))) return rc;
        if (p.mpp_limiter) {
            if ((rc = mpp_step(c, A, B, dt))) return rc;
        } else if ((rc = stage(c, A, A, -1, B, 1.0, 0.0, 1.0, dt, 0))) {
            return rc;
        }
        *res = B;
        *slot = 0;
        return 0;
    }
    const bool one_d = p.ndims == 1;  // 1D: smax overwritten per stage
    if (one_d && !c->timing && p.time_order <= 3) {
        // the stages of SetRKinfo's SSP schemes (lib/2d/SetRKinfo.cpp:50-64),
        // as the switch below enqueues them one by one, in one launch
        Stage1DStep s;
        memset(&s, 0, sizeof s);
        s.n = p.time_order;
        // (only the last stage reduces its wave speeds: 1D GetCFL reads those)
        const int n = p.time_order;
        const int role[3] = {A, B, C};
        for (int k = 0; k < n; k++) {
            const SspStage g = ssp_stage(n, k);
            s.stage[k] = stage1d_args(c, role[g.src], role[g.qa], g.qb >= 0 ? role[g.qb] : -1, role[g.out], g.a1, g.a2,
                                      g.beta, dt, k == n - 1 ? n - 1 : -1);
        }
        if (c->devctl && c->ctlp) {  // the controller goes with the step
            s.ctl = c->ctl;
            s.cp = *c->ctlp;
            c->ctl_enqueued = true;
        }
        rc = launch_step1d(weno_kind(p), s, c->st);
        if (rc) return fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "1D step launch failed");
        *res = role[ssp_result(n)];
        *slot = n - 1;
        return 0;
    }
    switch (p.time_order) {
    case 1:
    case 2:
    case 3: {
        // 2D: every stage max-accumulates into slot 0; 1D: only the last
        // stage's speeds count (slot n - 1)
        const int n = p.time_order;
        const int role[3] = {A, B, C};
        for (int k = 0; k < n; k++) {
            const SspStage g = ssp_stage(n, k);
            if ((rc = stage(c, role[g.src], role[g.qa], g.qb >= 0 ? role[g.qb] : -1, role[g.out], g.a1, g.a2, g.beta,
                            dt, one_d ? (k == n - 1 ? n - 1 : -1) : 0)))
                return rc;
        }
        *res = role[ssp_result(n)];
        *slot = one_d ? n - 1 : 0;
        return 0;
    }
    case 4: {
        // SSP(10,4) (lib/2d/FinSolveRK.cpp:213-291): stages ping-pong between
        // B and C, q2 lives in buffer 3
        // (1D: only the last stage's speeds count, in slot 1)
        int x = A, y = B;
        for (int s = 1; s <= 9; s++) {
            if ((rc = stage(c, x, x, -1, y, 1.0, 0.0, 1.0 / 6.0, dt, one_d ? -1 : 0))) return rc;
            x = y;
            y = (y == B) ? C : B;
            if (s == 5 && launch_rk4_combine(c->buf[A], c->buf[3], c->buf[x], (long long)c->M * c->plane, c->st))
                return fail(c, FINESS_ERR_HIP, "RK4 combination launch failed");
        }
        if ((rc = stage(c, x, 3, x, y, 1.0, 3.0 / 5.0, 1.0 / 10.0, dt, one_d ? 1 : 0))) return rc;
        *res = y;
        *slot = one_d ? 1 : 0;
        return 0;
    }
    case 5: {
        // 8-stage fifth-order scheme (lib/2d/FinSolveRK.cpp:292-305): q2 in
        // buffer 3 (zero at the start of the step), L in buffer 4
        const long long n = (long long)c->M * c->plane;
        int x = A, y = B;
        for (int s = 0; s < 8; s++) {
            if ((rc = stage_l(c, x, c->buf[4], one_d ? (s == 7 ? 1 : -1) : 0))) return rc;
            if (launch_rk5_update(c->buf[x], c->buf[3], c->buf[A], c->buf[4], c->buf[y], n, RK5_GAMMA[0][s],
                                  RK5_GAMMA[1][s], RK5_GAMMA[2][s], RK5_DELTA[s], RK5_BETA[s] * dt, s == 0, c->st))
                return fail(c, FINESS_ERR_HIP, "RK5 update launch failed");
            x = y;
            y = (y == B) ? C : B;
        }
        *res = x;
        *slot = one_d ? 1 : 0;
        return 0;
    }
    default:
        return fail(c, FINESS_ERR_CONFIG, "time_order not supported");
    }
}

// GetCFL from the slot's maxima (lib/2d/GetCFL.cpp:31-74, lib/1d/GetCFL.cpp:9-45).
int read_cfl(finess_ctx* c, int slot, double dt, double dtmax, double* cfl_out)
{
    HIPCHK(c, hipMemcpyAsync(c->smax_h, c->smax + 2 * CFL_SPREAD * slot, 2 * CFL_SPREAD * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const finess_params& p = c->p;
    unsigned long long bx = 0, by = 0;  // bit patterns of non-negative doubles (or +inf)
    for (int k = 0; k < CFL_SPREAD; k++) {
        bx = c->smax_h[k] > bx ? c->smax_h[k] : bx;
        by = c->smax_h[CFL_SPREAD + k] > by ? c->smax_h[CFL_SPREAD + k] : by;
    }
    const double sx = bits_to_double(bx), sy = bits_to_double(by);
    if (dt > dtmax) return fail(c, FINESS_ERR_DT, "dt is out of bounds (dt > max_dt)");
    if (!isfinite(sx) || !isfinite(sy)) return fail(c, FINESS_ERR_NONFINITE, "non-finite wave speed");
    double cfl = -100.0;
    if (p.ndims == 1) {
        cfl = rmax(dt * sx / p.dx, cfl);
        *cfl_out = cfl;
        if (cfl > 1.0e10) return fail(c, FINESS_ERR_CFL, "CFL number is out of bounds");
        return 0;
    }
    cfl = rmax(dt * sx / p.dx, cfl);
    cfl = rmax(dt * sy / p.dy, cfl);
    *cfl_out = cfl;
    if (cfl > 1.0e8) return fail(c, FINESS_ERR_CFL, "CFL number is out of bounds");
    return 0;
}

int alloc(finess_ctx* c, void** ptr, size_t bytes)
{
    HIPCHK(c, hipMalloc(ptr, bytes));
    return 0;
}

int ensure_ghosted(finess_ctx* c)
{
    const finess_params& p = c->p;
    if (!c->ghosted) {
        int rc;
        if ((rc = alloc(c, (void**)&c->ghosted, ref_numel(p, c->M) * sizeof(double)))) return rc;
    }
    return 0;
}

// Uploads a reference-layout host array with ghosts into c->ghosted (planes
// with ghosts, dg = mbc).
int upload_ghosted(finess_ctx* c, const double* q)
{
    const finess_params& p = c->p;
    int rc;
    if ((rc = ensure_ghosted(c))) return rc;
    const size_t n = ref_numel(p, c->M);
    if (p.ndims == 1) {
        std::vector<double> tmp(n);
        const int L = p.mx + 2 * p.mbc;
        for (int i = 0; i < L; i++)
            for (int m = 0; m < c->M; m++) tmp[(size_t)m * L + i] = q[(size_t)i * c->M + m];
        HIPCHK(c, hipMemcpyAsync(c->ghosted, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        return 0;
    }
    HIPCHK(c, hipMemcpyAsync(c->staging, q, n * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (launch_ref_to_planes(c->staging, c->ghosted, p.mx, p.my, c->M, p.mbc, p.mbc, c->st))
        return fail(c, FINESS_ERR_HIP, "layout conversion failed");
    return 0;
}

// The 2D aux planes the source term reads: the interior of the caller's aux
// (reference layout with ghosts, maux fields), once per call that hands it over.
int upload_aux2d(finess_ctx* c, const double* aux)
{
    if (!c->aux2d || !aux) return 0;
    const finess_params& p = c->p;
    const int g = p.mbc, na = p.maux;
    std::vector<double> tmp((size_t)na * c->plane);
    for (int m = 0; m < na; m++)
        for (int i = 0; i < p.mx; i++)
            for (int j = 0; j < p.my; j++)
                tmp[(size_t)m * c->plane + (size_t)i * p.my + j] = aux[((size_t)(i + g) * (p.my + 2 * g) + (j + g)) * na + m];
    HIPCHK(c, hipMemcpy(c->aux2d, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

} // namespace

namespace {
// Rotated frame of direction ixy (1: x, 2: y): the two in-plane momenta swap,
// as the reference's (mu, mv) pairs do.
template <class P>
void rotate(const double* q, int dir, double* r)
{
    for (int k = 0; k < P::MEQN; k++) r[P::rot_q(k, dir)] = q[k];
}

template <class P>
int eig_matrices(const finess_params& p, int dir, const double* qavg, double* Lm, double* Rm)
{
    constexpr int M = P::MEQN;
    PhysParams pp{p.gamma, p.sw_compat_hl};
    double qa[M];
    rotate<P>(qavg, dir, qa);
    typename P::template Eig<double> e;
    P::eig(qa, pp, e);
    // The kernels number the characteristic fields in the rotated frame.  The
    // Euler callbacks number the entropy and shear fields of ixy = 2 the other
    // way round (their rows follow (mu, mv)), so Euler's field k is the
    // callbacks' field rot_q(k, dir); shallow water keeps one order.  The
    // order does not change the flux: every field is reconstructed on its own.
    auto fld = [dir](int k) { return std::is_same<P, Euler2D>::value ? P::rot_q(k, dir) : k; };
    for (int j = 0; j < M; j++) {
        // column j of L: L applied to the unit vector e_j of the unrotated frame
        double v[M] = {}, vr[M], w[M];
        v[j] = 1.0;
        rotate<P>(v, dir, vr);
        P::left(e, vr, w);
        for (int k = 0; k < M; k++) Lm[fld(k) * M + j] = w[k];
        // column j of R: the unrotated image of the characteristic unit vector
        double u[M] = {}, qr[M], qo[M];
        u[j] = 1.0;
        P::right(e, u, qr);
        rotate<P>(qr, dir, qo);  // the swap is its own inverse
        for (int k = 0; k < M; k++) Rm[k * M + fld(j)] = qo[k];
    }
    return FINESS_OK;
}

template <class P>
double speed(const finess_params& p, int dir, const double* ql, const double* qr)
{
    constexpr int M = P::MEQN;
    PhysParams pp{p.gamma, p.sw_compat_hl};
    double cl[P::NV], cr[P::NV], rl[M], rr[M];
    P::cell_record(ql, pp, cl);
    P::cell_record(qr, pp, cr);
    P::state(cl, dir, rl);
    P::state(cr, dir, rr);
    return P::wave_speed(rl, rr, cl, cr, pp);
}
} // namespace

// ======================================================================
// C-ABI
// ======================================================================
extern "C" {

void finess_params_default(finess_params* p)
{
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->ndims = 2;
    p->mx = p->my = 64;
    p->mbc = 3;
    p->xlow = p->ylow = 0.0;
    p->xhigh = p->yhigh = 1.0;
    p->meqn = 5;
    p->maux = 0;
    p->space_order = 5;
    p->time_order = 3;
    p->time_stepping_method = FINESS_TS_RK;
    p->initial_dt = 1.0e-3;
    p->max_dt = 1.0e10;
    p->desired_cfl = 0.4;
    p->max_cfl = 0.45;
    p->nv = 1000000;
    p->weno_version = FINESS_WENO_JS;   // python/finess/params/weno.py:15-51
    p->epsilon = 1.0e-6;
    p->power_param = 2.0;
    p->alpha_scaling = 1.1;
    p->physics = FINESS_PHYS_EULER;
    p->gamma = 1.4;                      // python/finess/params/euler.py:13-19
    p->sw_compat_hl = 1;
    for (int k = 0; k < 4; k++) p->bc[k] = FINESS_BC_EXTRAP;
    p->precision = FINESS_PREC_FP64;
    p->routing = FINESS_ROUTE_NONE;
    p->route_tau = 0.05;
    p->route_per_step = 1;
    finess_params_finalize(p);
}

int finess_params_finalize(finess_params* p)
{
    if (!p) return FINESS_ERR_ARG;
    if (p->mx < 1 || (p->ndims == 2 && p->my < 1)) return FINESS_ERR_CONFIG;
    p->dx = (p->xhigh - p->xlow) / p->mx;
    p->dy = p->ndims == 2 ? (p->yhigh - p->ylow) / p->my : 1.0;
    // checks of the reference's parameter schema (python/finess/params/*.py)
    if (!(p->nv > 0) || !(p->initial_dt > 0.0) || !(p->max_dt >= p->initial_dt)) return FINESS_ERR_CONFIG;
    if (!(p->desired_cfl > 0.0) || !(p->max_cfl >= p->desired_cfl)) return FINESS_ERR_CONFIG;
    if (!(p->alpha_scaling >= 1.0) || !(p->epsilon > 0.0)) return FINESS_ERR_CONFIG;
    // what the device path implements (GetWenoReconstruct throws for the rest,
    // lib/WenoReconstruct.cpp:39-73)
    if (p->weno_version != FINESS_WENO_JS && p->weno_version != FINESS_WENO_FD && p->weno_version != FINESS_WENO_Z)
        return FINESS_ERR_CONFIG;
    const int so = p->space_order;
    const bool have = so == 1 || so == 5 || (p->weno_version == FINESS_WENO_JS && (so == 7 || so == 9 || so == 11)) ||
                      (p->weno_version != FINESS_WENO_JS && (so == 7 || so == 9));
    if (!have) return FINESS_ERR_CONFIG;
    // orders other than 5 run the face-tile kernels in 2D and the face
    // kernel in 1D: Runge-Kutta and ConstructL in FP64, mbc >= r
    // (lib/2d/ConstructL.cpp:31-33, lib/1d/ConstructL.cpp:32-34), and every
    // stencil within the boundary map's range
    if (so != 5) {
        const int r = (so + 1) / 2;
        if (p->time_stepping_method != FINESS_TS_RK || p->routing != FINESS_ROUTE_NONE ||
            p->precision != FINESS_PREC_FP64 || p->mbc < r || p->mx < r || (p->ndims == 2 && p->my < r))
            return FINESS_ERR_CONFIG;
    }
    // source terms (SampleFunction + SourceTermFunc in ConstructL): Runge-Kutta
    // and ConstructL; ConstructLxWL exits on them (lib/2d/ConstructLxWL.cpp:376-380)
    if (p->source_term) {
        if (p->time_stepping_method != FINESS_TS_RK || p->routing != FINESS_ROUTE_NONE) return FINESS_ERR_CONFIG;
        if (p->source_kind < FINESS_SRC_ZERO || p->source_kind > FINESS_SRC_DAMPING) return FINESS_ERR_CONFIG;
        if (p->source_kind == FINESS_SRC_SW_BATHYMETRY &&
            (p->ndims != 2 || p->physics != FINESS_PHYS_SHALLOW_WATER || p->maux < 3))
            return FINESS_ERR_CONFIG;
    }
    // global_alpha (GlobalWaveSpd before each operator call) runs in the 1D
    // and 2D stage kernels: Runge-Kutta, Lax-Wendroff (limited or not), RK4-PIF's
    // final stage (ConstructL_NOC computes the speeds but does not use them,
    // lib/2d/FinSolveRK_PIF.cpp:759-764), ConstructL
    if (p->global_alpha && p->routing != FINESS_ROUTE_NONE) return FINESS_ERR_CONFIG;
    if (p->ndims == 1) {
        if (p->physics != FINESS_PHYS_ADVECTION || p->meqn != 1) return FINESS_ERR_CONFIG;
        // Runge-Kutta, or Lax-Wendroff without the limiter (lib/1d/FinSolveLxW.cpp)
        if (p->time_stepping_method == FINESS_TS_PIF || (p->time_stepping_method == FINESS_TS_LXW && p->mpp_limiter))
            return FINESS_ERR_CONFIG;
        if (p->mx < 3) return FINESS_ERR_CONFIG;
    } else if (p->ndims == 2) {
        if (p->physics == FINESS_PHYS_EULER && p->meqn != 5) return FINESS_ERR_CONFIG;
        if (p->physics == FINESS_PHYS_SHALLOW_WATER && p->meqn != 3) return FINESS_ERR_CONFIG;
        if (p->physics == FINESS_PHYS_ADVECTION) return FINESS_ERR_CONFIG;
        if (p->mx < 5 || p->my < 5) return FINESS_ERR_CONFIG;
    } else {
        return FINESS_ERR_CONFIG;
    }
    if (p->mbc < 3) return FINESS_ERR_CONFIG;
    // the boundary map folds ghost indices g in [-n, 2n-1] only
    if (p->mbc > p->mx || (p->ndims == 2 && p->mbc > p->my)) return FINESS_ERR_CONFIG;
    if (p->time_stepping_method < FINESS_TS_RK || p->time_stepping_method > FINESS_TS_PIF) return FINESS_ERR_CONFIG;
    if (p->time_stepping_method == FINESS_TS_LXW) {
        // the 5-point differences of the integrated fluxes need mbc >= 5
        // (lib/2d/ConstructIntegratedR.cpp:87-90, lib/1d/ConstructIntegratedF.cpp)
        if (p->mbc < 5 || (p->ndims == 2 && p->physics == FINESS_PHYS_ADVECTION)) return FINESS_ERR_CONFIG;
        // the limiter (ApplyMPPLimiter2D) and the FP32 stencil form of this
        // step exist for Euler only
        if (p->physics != FINESS_PHYS_EULER && (p->mpp_limiter || p->precision != FINESS_PREC_FP64))
            return FINESS_ERR_CONFIG;
        if (p->time_order < 2 || p->time_order > 3) return FINESS_ERR_CONFIG;
    } else if (p->time_stepping_method == FINESS_TS_PIF) {
        // RK4-PIF exists for time_order 4 only (lib/2d/FinSolveRK_PIF.cpp:151-206)
        if (p->ndims != 2 || p->physics == FINESS_PHYS_ADVECTION || p->time_order != 4) return FINESS_ERR_CONFIG;
        if (p->precision != FINESS_PREC_FP64) return FINESS_ERR_CONFIG;
    } else if (p->time_order < 1 || p->time_order > 5) {
        return FINESS_ERR_CONFIG;
    }
    // the routed stages follow the three-stage schedule
    if (p->routing != FINESS_ROUTE_NONE && p->time_order > 3) return FINESS_ERR_CONFIG;
    if (p->precision == FINESS_PREC_FP32 && weno_kind(*p) != WENO_JS5_P2) return FINESS_ERR_CONFIG;
    if (p->routing != FINESS_ROUTE_NONE &&
        (p->ndims != 2 || p->physics != FINESS_PHYS_EULER || weno_kind(*p) != WENO_JS5_P2 ||
         p->time_stepping_method != FINESS_TS_RK))
        return FINESS_ERR_CONFIG;
    // the routing record packs a face's cell indices in 16 bits each and
    // numbers the faces of both directions with 32-bit integers
    if (p->routing != FINESS_ROUTE_NONE &&
        (p->mx > 65535 || p->my > 65535 ||
         (long long)(p->mx + 1) * p->my + (long long)p->mx * (p->my + 1) >= (1ll << 31)))
        return FINESS_ERR_CONFIG;
    bool host = false;
    for (int k = 0; k < 4; k++) {
        if (p->bc[k] < FINESS_BC_EXTRAP || p->bc[k] > FINESS_BC_HOST) return FINESS_ERR_CONFIG;
        host = host || p->bc[k] == FINESS_BC_HOST;
    }
    // a host SetBndValues runs before every stage of the 2D Runge-Kutta
    // schemes with the SSP schedule (orders 1-3), plain or routed
    if (host && (p->ndims != 2 || p->time_stepping_method != FINESS_TS_RK || p->time_order > 3))
        return FINESS_ERR_CONFIG;
    return FINESS_OK;
}

// ---- the device physics on the host (callback conformance probe) --------


int finess_physics_eig(const finess_params* p, int ixy, const double* qavg, double* Lmat, double* Rmat)
{
    if (!p || !qavg || !Lmat || !Rmat || (ixy != 1 && ixy != 2)) return FINESS_ERR_ARG;
    if (p->physics == FINESS_PHYS_EULER) return eig_matrices<Euler2D>(*p, ixy - 1, qavg, Lmat, Rmat);
    if (p->physics == FINESS_PHYS_SHALLOW_WATER) return eig_matrices<ShallowWater2D>(*p, ixy - 1, qavg, Lmat, Rmat);
    if (p->physics == FINESS_PHYS_ADVECTION) {  // identity projections (apps/1d/advection/lib)
        Lmat[0] = Rmat[0] = 1.0;
        return FINESS_OK;
    }
    return FINESS_ERR_CONFIG;
}

int finess_physics_wave_speed(const finess_params* p, int ixy, const double* ql, const double* qr,
                              const double* auxl, const double* auxr, double* speed_out)
{
    if (!p || !ql || !qr || !speed_out || (ixy != 1 && ixy != 2)) return FINESS_ERR_ARG;
    if (p->physics == FINESS_PHYS_EULER) *speed_out = speed<Euler2D>(*p, ixy - 1, ql, qr);
    else if (p->physics == FINESS_PHYS_SHALLOW_WATER) *speed_out = speed<ShallowWater2D>(*p, ixy - 1, ql, qr);
    else if (p->physics == FINESS_PHYS_ADVECTION) {
        if (!auxl || !auxr) return FINESS_ERR_ARG;
        // s1 = min(a_l, mean), s2 = max(a_r, mean) (apps/1d/advection/lib/SetWaveSpd.cpp)
        const double al = auxl[0], ar = auxr[0];
        const double s1 = rmin(al, 0.5 * (al + ar)), s2 = rmax(ar, 0.5 * (al + ar));
        *speed_out = rmax(fabs(s1), fabs(s2));
    } else {
        return FINESS_ERR_CONFIG;
    }
    return FINESS_OK;
}

int finess_physics_source(const finess_params* p, const double* q, const double* aux, double* psi)
{
    if (!p || !q || !psi) return FINESS_ERR_ARG;
    if (p->source_kind == FINESS_SRC_SW_BATHYMETRY && (p->meqn != 3 || p->maux < 3 || !aux)) return FINESS_ERR_CONFIG;
    const int M = p->meqn;
    const int kind = source_arg(1, p->source_kind);
    if (M == 5) source_psi<5>(kind, p->source_coef, q, aux, 1, psi);
    else if (M == 3) source_psi<3>(kind, p->source_coef, q, aux, 1, psi);
    else if (M == 1) source_psi<1>(kind, p->source_coef, q, aux, 1, psi);
    else return FINESS_ERR_CONFIG;
    return FINESS_OK;
}

const char* finess_version(void) { return "finess-hip 0.1 (gfx950)"; }

int finess_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int finess_create(const finess_params* pin, finess_ctx** out)
{
    if (!pin || !out) return FINESS_ERR_ARG;
    *out = nullptr;
    finess_params p = *pin;
    int rc = finess_params_finalize(&p);
    if (rc) return rc;
    if (finess_device_count() < 1) return FINESS_ERR_NODEVICE;
    finess_ctx* c = new finess_ctx();
    c->p = p;
    c->M = p.meqn;
    c->plane = p.ndims == 1 ? (long long)p.mx : (long long)p.mx * p.my;
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        finess_destroy(c);
        return FINESS_ERR_HIP;
    }
    const size_t sbytes = (size_t)c->M * c->plane * sizeof(double);
    for (int k = 0; k < 3 && !rc; k++) rc = alloc(c, (void**)&c->buf[k], sbytes);
    if (p.time_order >= 4 && p.time_stepping_method == FINESS_TS_RK)
        for (int k = 3; k < 5 && !rc; k++) rc = alloc(c, (void**)&c->buf[k], sbytes);
    if (!rc) rc = alloc(c, (void**)&c->smax, 2 * CFL_SPREAD * NSLOT * sizeof(unsigned long long));
    if (!rc && hipHostMalloc((void**)&c->smax_h, 2 * CFL_SPREAD * NSLOT * sizeof(unsigned long long), hipHostMallocDefault) !=
                   hipSuccess)
        rc = FINESS_ERR_HIP;
    if (!rc) rc = alloc(c, (void**)&c->staging, ref_numel(p, c->M) * sizeof(double));
    if (!rc && p.ndims == 1) {
        rc = alloc(c, (void**)&c->aux, (size_t)p.mx * sizeof(double));
        if (!rc) rc = alloc(c, (void**)&c->fhat1d, (size_t)(p.mx + 1) * sizeof(double));
        if (!rc) {
            std::vector<double> one((size_t)p.mx, 1.0);  // AuxFunc default: speed 1
            if (hipMemcpy(c->aux, one.data(), p.mx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                rc = FINESS_ERR_HIP;
        }
    }
    if (!rc && p.ndims == 2 && p.source_term && p.source_kind == FINESS_SRC_SW_BATHYMETRY) {
        // aux = 0 until the caller provides it (AuxFunc)
        const size_t ab = (size_t)p.maux * c->plane * sizeof(double);
        rc = alloc(c, (void**)&c->aux2d, ab);
        if (!rc && hipMemset(c->aux2d, 0, ab) != hipSuccess) rc = FINESS_ERR_HIP;
    }
    if (!rc && p.time_stepping_method == FINESS_TS_PIF) {
        rc = alloc(c, (void**)&c->pif_fx, (size_t)(p.mx + 1) * p.my * c->M * sizeof(double));
        if (!rc) rc = alloc(c, (void**)&c->pif_fy, (size_t)p.mx * (p.my + 1) * c->M * sizeof(double));
    }
    if (!rc && p.time_stepping_method != FINESS_TS_RK) {
        // (the FP32 form stores the time-integrated fluxes in FP32)
        const size_t lb = (size_t)c->M * (p.mx + 6) * (p.my + 6) * lxw_flux_bytes(p);
        rc = alloc(c, (void**)&c->lxwF, lb);
        if (!rc) rc = alloc(c, (void**)&c->lxwG, lb);
        if (!rc && p.mpp_limiter) {
            const size_t nfx = (size_t)(p.mx + 1) * p.my, nfy = (size_t)p.mx * (p.my + 1);
            rc = alloc(c, (void**)&c->mpp_fh, nfx * c->M * sizeof(double));
            if (!rc) rc = alloc(c, (void**)&c->mpp_fl, nfx * c->M * sizeof(double));
            if (!rc) rc = alloc(c, (void**)&c->mpp_gh, nfy * c->M * sizeof(double));
            if (!rc) rc = alloc(c, (void**)&c->mpp_gl, nfy * c->M * sizeof(double));
            if (!rc) rc = alloc(c, (void**)&c->mpp_amin, (size_t)4 * p.mx * p.my * sizeof(double));
            if (!rc) rc = alloc(c, (void**)&c->mpp_gmax, 2 * CFL_SPREAD * sizeof(unsigned long long));
        }
    }
    for (int k = 0; k < 4; k++) c->hostbc = c->hostbc || p.bc[k] == FINESS_BC_HOST;
    if (!rc && p.global_alpha) {
        rc = alloc(c, (void**)&c->galpha, 2 * sizeof(double));
        if (!rc) rc = alloc(c, (void**)&c->gspread, 2 * CFL_SPREAD * sizeof(unsigned long long));
    }
    if (!rc && p.space_order != 5) {
        rc = alloc(c, (void**)&c->wide_fx, (size_t)(p.mx + 1) * p.my * c->M * sizeof(double));
        if (!rc) rc = alloc(c, (void**)&c->wide_fy, (size_t)p.mx * (p.my + 1) * c->M * sizeof(double));
    }
    if (!rc) rc = alloc(c, (void**)&c->route_counts, 2 * CFL_SPREAD * sizeof(unsigned long long));
    if (!rc) rc = alloc(c, (void**)&c->ctl, sizeof(StepCtl));
    if (!rc && p.routing != FINESS_ROUTE_NONE && p.ndims == 2) {
        const size_t nf = (size_t)(p.mx + 1) * p.my + (size_t)p.mx * (p.my + 1);
        rc = alloc(c, (void**)&c->route_tile, (size_t)routed_tiles(p.mx, p.my) * sizeof(RouteTileRec));
        if (!rc) rc = alloc(c, (void**)&c->route_list, nf * sizeof(int));
        if (!rc) rc = alloc(c, (void**)&c->route_n, 4 * sizeof(unsigned long long));
        if (!rc) rc = launch_fill_u64(c->route_n, 4, 0ull, c->st) ? FINESS_ERR_HIP : 0;
        if (!rc) rc = alloc(c, &c->route_flux, nf * 6 * (p.precision == FINESS_PREC_FP32 ? sizeof(float) : sizeof(double)));
    }
    if (!rc) rc = launch_fill_u64(c->route_counts, 2 * CFL_SPREAD, 0ull, c->st) ? FINESS_ERR_HIP : 0;
    if (!rc) rc = launch_fill_u64(c->smax, 2 * CFL_SPREAD * NSLOT, 0ull, c->st) ? FINESS_ERR_HIP : 0;
    if (rc) {
        finess_destroy(c);
        return rc;
    }
    *out = c;
    return FINESS_OK;
}

void finess_destroy(finess_ctx* c)
{
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    for (int k = 0; k < 5; k++) hipFree(c->buf[k]);
    hipFree(c->aux);
    hipFree(c->aux2d);
    hipFree(c->lxwF);
    hipFree(c->lxwG);
    hipFree(c->staging);
    hipFree(c->ghosted);
    hipFree(c->faces_x);
    hipFree(c->faces_y);
    hipFree(c->fhat1d);
    hipFree(c->smax);
    hipFree(c->frags);
    hipFree(c->route_counts);
    hipFree(c->route_tile);
    hipFree(c->route_list);
    hipFree(c->route_n);
    hipFree(c->ctl);
    hipFree(c->mpp_fh);
    hipFree(c->mpp_gh);
    hipFree(c->wide_fx);
    hipFree(c->wide_fy);
    hipFree(c->galpha);
    hipFree(c->gspread);
    hipFree(c->edge);
    hipFree(c->mpp_fl);
    hipFree(c->mpp_gl);
    hipFree(c->mpp_amin);
    hipFree(c->mpp_gmax);
    hipFree(c->pif_fx);
    hipFree(c->slog);
    hipFree(c->spart);
    hipFree(c->pif_fy);
    hipFree(c->route_flux);
    if (c->smax_h) hipHostFree(c->smax_h);
    for (hipEvent_t e : c->tev) hipEventDestroy(e);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

const char* finess_last_error(const finess_ctx* c) { return c ? c->err.c_str() : "null context"; }

int finess_set_bnd_callback(finess_ctx* c, finess_bnd_fn fn, void* user)
{
    if (!c) return FINESS_ERR_ARG;
    c->bnd_fn = fn;
    c->bnd_user = user;
    return FINESS_OK;
}

int finess_set_step_log(finess_ctx* c, int on)
{
    if (!c) return FINESS_ERR_ARG;
    if (on && !c->slog) {
        int rc;
        if ((rc = alloc(c, (void**)&c->slog, (size_t)64 * (c->M + 1) * sizeof(double)))) return rc;
        if ((rc = alloc(c, (void**)&c->spart, (size_t)c->M * 240 * sizeof(double)))) return rc;
    }
    c->steplog = on != 0;
    if (!on) c->hlog.clear();
    return FINESS_OK;
}

int finess_read_step_log(finess_ctx* c, double* rec, int max_records)
{
    if (!c || (max_records > 0 && !rec) || max_records < 0) return FINESS_ERR_ARG;
    const size_t w = (size_t)c->M + 1, avail = c->hlog.size() / w;
    const size_t n = avail < (size_t)max_records ? avail : (size_t)max_records;
    if (n) {
        memcpy(rec, c->hlog.data(), n * w * sizeof(double));
        c->hlog.erase(c->hlog.begin(), c->hlog.begin() + (long)(n * w));
    }
    return (int)n;
}

int finess_apply_mpp_limiter_2d(finess_ctx* c, double dt, const double* q, const double* fLF, const double* gLF,
                                double* fHat, double* gHat)
{
    if (!c || !q || !fLF || !gLF || !fHat || !gHat) return FINESS_ERR_ARG;
    const finess_params& p = c->p;
    if (!c->mpp_fh) return fail(c, FINESS_ERR_CONFIG, "no limiter on this context (2D Euler, Lax-Wendroff, mpp_limiter)");
    const int mx = p.mx, my = p.my, g = p.mbc, M = c->M;
    const size_t nfx = (size_t)(mx + 1) * my, nfy = (size_t)mx * (my + 1);
    // reference layout (i, j, m), 0-based interior indices, leading dimension ld
    auto ref = [&](int i, int j, int m, int ld) { return ((size_t)(i + g) * (ld + 2 * g) + (j + g)) * M + m; };
    std::vector<double> hx(nfx * M), lx(nfx * M), hy(nfy * M), ly(nfy * M);
    for (int m = 0; m < M; m++)
        for (int i = 0; i <= mx; i++)
            for (int j = 0; j < my; j++) {
                hx[m * nfx + (size_t)i * my + j] = fHat[ref(i, j, m, my)];
                lx[m * nfx + (size_t)i * my + j] = fLF[ref(i, j, m, my)];
            }
    for (int m = 0; m < M; m++)
        for (int i = 0; i < mx; i++)
            for (int j = 0; j <= my; j++) {
                hy[m * nfy + (size_t)i * (my + 1) + j] = gHat[ref(i, j, m, my + 1)];
                ly[m * nfy + (size_t)i * (my + 1) + j] = gLF[ref(i, j, m, my + 1)];
            }
    double* qd = c->buf[(c->cur + 2) % 3];  // not the held state
    HIPCHK(c, hipMemcpyAsync(c->staging, q, ref_numel(p, M) * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (launch_ref_to_planes(c->staging, qd, mx, my, M, g, 0, c->st)) return fail(c, FINESS_ERR_HIP, "layout conversion failed");
    HIPCHK(c, hipMemcpyAsync(c->mpp_fh, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->mpp_fl, lx.data(), lx.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->mpp_gh, hy.data(), hy.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->mpp_gl, ly.data(), ly.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    MppLaunch m;
    m.q = qd;
    m.out = nullptr;
    m.mx = mx;
    m.my = my;
    for (int k = 0; k < 4; k++) m.bc[k] = p.bc[k];
    m.dx = p.dx;
    m.dy = p.dy;
    m.dt = dt;
    m.gamma = p.gamma;
    m.fh = c->mpp_fh;
    m.gh = c->mpp_gh;
    m.fl = c->mpp_fl;
    m.gl = c->mpp_gl;
    m.amin = c->mpp_amin;
    m.gmax = c->mpp_gmax;
    m.smax = c->smax;
    if (launch_mpp_limit(m, c->st)) return fail(c, FINESS_ERR_HIP, "limiter launch failed");
    HIPCHK(c, hipMemcpyAsync(hx.data(), c->mpp_fh, hx.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(hy.data(), c->mpp_gh, hy.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int mm = 0; mm < M; mm++)
        for (int i = 0; i <= mx; i++)
            for (int j = 0; j < my; j++) fHat[ref(i, j, mm, my)] = hx[mm * nfx + (size_t)i * my + j];
    for (int mm = 0; mm < M; mm++)
        for (int i = 0; i < mx; i++)
            for (int j = 0; j <= my; j++) gHat[ref(i, j, mm, my + 1)] = hy[mm * nfy + (size_t)i * (my + 1) + j];
    return FINESS_OK;
}

int finess_set_state(finess_ctx* c, const double* q, const double* aux, double t)
{
    if (!c || !q) return FINESS_ERR_ARG;
    const finess_params& p = c->p;
    const size_t n = ref_numel(p, c->M);
    c->cur = 0;
    if (p.ndims == 1) {
        std::vector<double> tmp((size_t)p.mx);
        for (int i = 0; i < p.mx; i++) tmp[i] = q[(size_t)(i + p.mbc) * c->M];
        HIPCHK(c, hipMemcpy(c->buf[0], tmp.data(), p.mx * sizeof(double), hipMemcpyHostToDevice));
        if (aux) {
            const int na = p.maux > 0 ? p.maux : 1;
            for (int i = 0; i < p.mx; i++) tmp[i] = aux[(size_t)(i + p.mbc) * na];
            HIPCHK(c, hipMemcpy(c->aux, tmp.data(), p.mx * sizeof(double), hipMemcpyHostToDevice));
        }
    } else {
        HIPCHK(c, hipMemcpyAsync(c->staging, q, n * sizeof(double), hipMemcpyHostToDevice, c->st));
        if (launch_ref_to_planes(c->staging, c->buf[0], p.mx, p.my, c->M, p.mbc, 0, c->st))
            return fail(c, FINESS_ERR_HIP, "layout conversion failed");
        HIPCHK(c, hipStreamSynchronize(c->st));
        int rc;
        if ((rc = upload_aux2d(c, aux))) return rc;
    }
    c->t = t;
    c->rec_age = -1;  // a new state has no routing record yet
    c->have_state = true;
    return FINESS_OK;
}

int finess_get_state(finess_ctx* c, double* q, double* t)
{
    if (!c || !q) return FINESS_ERR_ARG;
    if (!c->have_state) return fail(c, FINESS_ERR_ARG, "no state");
    const finess_params& p = c->p;
    const size_t n = ref_numel(p, c->M);
    if (p.ndims == 1) {
        std::vector<double> tmp((size_t)p.mx);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), c->buf[c->cur], p.mx * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (int i = -p.mbc; i < p.mx + p.mbc; i++) {
            int fl;
            q[(size_t)(i + p.mbc) * c->M] = tmp[bc_map(p.bc[0], p.bc[1], i, p.mx, fl)];
        }
    } else {
        if (launch_planes_to_ref(c->buf[c->cur], c->staging, p.mx, p.my, c->M, p.mbc, p.bc, normal_x(p), normal_y(p),
                                 c->st))
            return fail(c, FINESS_ERR_HIP, "layout conversion failed");
        HIPCHK(c, hipMemcpyAsync(q, c->staging, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    if (t) *t = c->t;
    return FINESS_OK;
}

int finess_construct_l(finess_ctx* c, const double* q, const double* aux, double* Lstar, double* smax)
{
    if (!c || !q || !Lstar || !smax) return FINESS_ERR_ARG;
    const finess_params& p = c->p;
    int rc;
    if ((rc = upload_ghosted(c, q))) return rc;
    double* Ld = c->buf[(c->cur + 2) % 3];  // scratch: never the current state
    if (p.ndims == 1) {
        if (!c->faces_x && (rc = alloc(c, (void**)&c->faces_x, (size_t)(p.mx + 1) * sizeof(double)))) return rc;
        double* auxg = nullptr;
        if ((rc = alloc(c, (void**)&auxg, (size_t)(p.mx + 2 * p.mbc) * sizeof(double)))) return rc;
        std::vector<double> at((size_t)(p.mx + 2 * p.mbc), 1.0);
        const int na = p.maux > 0 ? p.maux : 1;
        if (aux)
            for (int i = 0; i < p.mx + 2 * p.mbc; i++) at[i] = aux[(size_t)i * na];
        hipError_t e = hipMemcpy(auxg, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice);
        Stage1DArgs a = base_stage1d_args(c);
        a.src = c->ghosted;
        a.aux = auxg;
        a.src_ghosts = p.mbc;
        a.smax_faces = c->faces_x;
        a.Lout = Ld;
        if (e == hipSuccess && launch_stage1d(weno_kind(p), a, c->st)) e = hipErrorLaunchFailure;
        std::vector<double> L((size_t)p.mx), al((size_t)p.mx + 1);
        if (e == hipSuccess) e = hipMemcpyAsync(L.data(), Ld, p.mx * sizeof(double), hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(al.data(), c->faces_x, (p.mx + 1) * sizeof(double), hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        hipFree(auxg);
        if (e != hipSuccess) return fail(c, FINESS_ERR_HIP, hipGetErrorString(e));
        for (int i = 0; i < p.mx; i++) Lstar[(size_t)(i + p.mbc) * c->M] = L[i];
        for (int f = 0; f <= p.mx; f++) smax[f + p.mbc] = al[f];  // smax.set(i, alpha): overwritten
        return FINESS_OK;
    }
    if (!c->faces_x && (rc = alloc(c, (void**)&c->faces_x, (size_t)(p.mx + 1) * p.my * sizeof(double)))) return rc;
    if (!c->faces_y && (rc = alloc(c, (void**)&c->faces_y, (size_t)p.mx * (p.my + 1) * sizeof(double)))) return rc;
    if ((rc = upload_aux2d(c, aux))) return rc;
    StageArgs a = base_stage_args(c);
    a.src = c->ghosted;
    a.src_ghosts = p.mbc;
    a.Lout = Ld;
    a.smax_faces_x = c->faces_x;
    a.smax_faces_y = c->faces_y;
    if (p.routing != FINESS_ROUTE_NONE) {
        if ((rc = stage_routed(c, a, false))) return rc;
    } else {
        rc = launch_stage2d(p.physics, p.precision == FINESS_PREC_FP32, weno_kind(p), 0, a, c->st);
        if (rc) return fail(c, rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP, "stage launch failed");
    }
    const size_t nL = (size_t)c->M * c->plane;
    std::vector<double> L(nL), ax((size_t)(p.mx + 1) * p.my), ay((size_t)p.mx * (p.my + 1));
    HIPCHK(c, hipMemcpyAsync(L.data(), Ld, nL * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(ax.data(), c->faces_x, ax.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(ay.data(), c->faces_y, ay.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const int g = p.mbc, M = c->M;
    const long long ldq = p.my + 2 * g, lds = p.my + 1 + 2 * g;
    for (int i = 0; i < p.mx; i++)
        for (int j = 0; j < p.my; j++)
            for (int m = 0; m < M; m++)
                Lstar[((i + g) * ldq + (j + g)) * M + m] = L[(size_t)m * c->plane + (size_t)i * p.my + j];
    // smax(i,j,1) for x-face i-1/2 (i = 1..mx+1, j = 1..my), smax(i,j,2) for y
    for (int i = 0; i <= p.mx; i++)
        for (int j = 0; j < p.my; j++) {
            double& s = smax[((i + g) * lds + (j + g)) * 2 + 0];
            s = rmax(s, ax[(size_t)i * p.my + j]);
        }
    for (int i = 0; i < p.mx; i++)
        for (int j = 0; j <= p.my; j++) {
            double& s = smax[((i + g) * lds + (j + g)) * 2 + 1];
            s = rmax(s, ay[(size_t)i * (p.my + 1) + j]);
        }
    return FINESS_OK;
}

int finess_construct_integrated_r(finess_ctx* c, double dt, const double* q, const double* aux, double* F, double* G)
{
    if (!c || !q || !F) return FINESS_ERR_ARG;
    const finess_params& p = c->p;
    if (p.mbc < 5) return fail(c, FINESS_ERR_CONFIG, "the integrated fluxes need mbc >= 5");
    int rc;
    if (p.ndims == 1) {
        // ConstructIntegratedF (lib/1d/ConstructIntegratedF.cpp:54-212) on the
        // caller's ghosted q and aux; F receives cells [1-3, mx+3]
        if ((rc = upload_ghosted(c, q))) return rc;
        const int n = p.mx + 2 * p.mbc, na = p.maux > 0 ? p.maux : 1;
        std::vector<double> at((size_t)n, 1.0), hF((size_t)p.mx + 6);
        if (aux)
            for (int i = 0; i < n; i++) at[i] = aux[(size_t)i * na];
        double *auxg = nullptr, *Fd = nullptr;
        if ((rc = alloc(c, (void**)&auxg, (size_t)n * sizeof(double)))) return rc;
        if ((rc = alloc(c, (void**)&Fd, hF.size() * sizeof(double)))) { hipFree(auxg); return rc; }
        Stage1DArgs a = base_stage1d_args(c);
        a.src = c->ghosted;
        a.aux = auxg;
        a.src_ghosts = p.mbc;
        hipError_t e = hipMemcpy(auxg, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess && launch_lxw_flux1d(a, dt, p.time_order, Fd, c->st)) e = hipErrorLaunchFailure;
        if (e == hipSuccess) e = hipMemcpyAsync(hF.data(), Fd, hF.size() * sizeof(double), hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        hipFree(auxg);
        hipFree(Fd);
        if (e != hipSuccess) return fail(c, FINESS_ERR_HIP, hipGetErrorString(e));
        for (int i = -3; i < p.mx + 3; i++) F[(size_t)(i + p.mbc) * c->M] = hF[i + 3];
        return FINESS_OK;
    }
    if (!G) return FINESS_ERR_ARG;
    if (p.physics == FINESS_PHYS_ADVECTION) return fail(c, FINESS_ERR_CONFIG, "2D LxW needs Euler or shallow water");
    if ((rc = upload_ghosted(c, q))) return rc;
    double *Fd = c->lxwF, *Gd = c->lxwG;
    bool own = false;
    const size_t lb = (size_t)c->M * (p.mx + 6) * (p.my + 6);
    const size_t vb = lxw_flux_bytes(p);  // the kernel's precision
    if (!Fd) {
        own = true;
        if ((rc = alloc(c, (void**)&Fd, lb * vb))) return rc;
        if ((rc = alloc(c, (void**)&Gd, lb * vb))) { hipFree(Fd); return rc; }
    }
    LxwArgs a;
    memset(&a, 0, sizeof a);
    a.q = c->ghosted;
    a.src_ghosts = p.mbc;
    a.F = Fd;
    a.G = Gd;
    a.mx = p.mx;
    a.my = p.my;
    a.plane = c->plane;
    for (int k = 0; k < 4; k++) a.bc[k] = p.bc[k];
    a.time_order = p.time_order;
    a.dt = dt;
    a.dx = p.dx;
    a.dy = p.dy;
    a.pp.gamma = p.gamma;
    std::vector<double> hF(lb), hG(lb);
    std::vector<float> hF32(vb == sizeof(float) ? lb : 0), hG32(hF32.size());
    void* dstF = vb == sizeof(float) ? (void*)hF32.data() : (void*)hF.data();
    void* dstG = vb == sizeof(float) ? (void*)hG32.data() : (void*)hG.data();
    hipError_t e = launch_lxw_flux(p.physics, p.precision == FINESS_PREC_FP32, a, c->st) ? hipErrorLaunchFailure : hipSuccess;
    if (e == hipSuccess) e = hipMemcpyAsync(dstF, Fd, lb * vb, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dstG, Gd, lb * vb, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (own) { hipFree(Fd); hipFree(Gd); }
    if (e != hipSuccess) return fail(c, FINESS_ERR_HIP, hipGetErrorString(e));
    for (size_t k = 0; k < hF32.size(); k++) { hF[k] = hF32[k]; hG[k] = hG32[k]; }
    const int g = p.mbc, M = c->M;
    const long long ldq = p.my + 2 * g, le = p.my + 6, lpl = (long long)(p.mx + 6) * le;
    for (int i = -3; i < p.mx + 3; i++)
        for (int j = -3; j < p.my + 3; j++)
            for (int m = 0; m < M; m++) {
                const long long o = (long long)m * lpl + (long long)(i + 3) * le + (j + 3);
                F[((i + g) * ldq + (j + g)) * M + m] = hF[o];
                G[((i + g) * ldq + (j + g)) * M + m] = hG[o];
            }
    return FINESS_OK;
}

// Weight of a cell in ConSoln's sums: dx dy (2D), dx (1D).
static double step_log_weight(const finess_params& p)
{
    return p.ndims == 1 ? p.dx : p.dx * p.dy;
}

// Appends the conservation record of the held state at time t (host
// controller: after an accepted step).
static int log_step(finess_ctx* c, double t)
{
    const finess_params& p = c->p;
    if (launch_con_soln(c->buf[c->cur], c->plane, c->M, step_log_weight(p), c->spart, nullptr, 0, t, c->slog, c->st))
        return fail(c, FINESS_ERR_HIP, "conservation sum launch failed");
    const size_t at = c->hlog.size();
    c->hlog.resize(at + c->M + 1);
    HIPCHK(c, hipMemcpyAsync(c->hlog.data() + at, c->slog, (c->M + 1) * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

static const char* ctl_error_message(long long code)
{
    switch (code) {
    case FINESS_ERR_DT: return "dt is out of bounds (dt > max_dt)";
    case FINESS_ERR_NONFINITE: return "non-finite wave speed";
    case FINESS_ERR_CFL: return "CFL number is out of bounds";
    case FINESS_ERR_NV: return "exceeded allowed # of time steps";
    default: return "device controller error";
    }
}

// advance_impl's controller run on the device (k_step_control after every
// attempt).  Steps are enqueued in chunks with the buffer rotation of
// accepted steps and the control block is read back once per chunk.  A
// rejected attempt halts the chunk's later kernels; the host then resumes
// from that attempt's input, with the routing-record state it had after the
// attempt, exactly as the host controller retries.  Decisions, dt sequence
// and results are those of the host controller.
static int advance_dev(finess_ctx* c, double tend, double* dtv, long long fixed_steps, finess_stats* stats)
{
    const finess_params& p = c->p;
    const bool by_steps = fixed_steps > 0;
    StepCtl h;
    memset(&h, 0, sizeof h);
    h.t = c->t;
    h.dt_next = dtv[1];
    h.dt = (!by_steps && h.t + dtv[1] > tend) ? tend - h.t : dtv[1];
    h.dtmin = h.dtmax = dtv[1];
    h.n_step = 1;
    StepCtlParams cp;
    cp.tend = tend;
    cp.max_dt = dtv[2];
    cp.cfl_max = p.max_cfl;
    cp.cfl_target = p.desired_cfl;
    cp.dx = p.dx;
    cp.dy = p.dy;
    cp.one_d = p.ndims == 1;
    cp.nv = p.nv;
    cp.target = by_steps ? fixed_steps : 0;
    int rc = 0;
    bool run = by_steps || c->t < tend;
    if (run && !by_steps && h.n_step > p.nv) {
        rc = fail(c, FINESS_ERR_NV, ctl_error_message(FINESS_ERR_NV));
        run = false;
    }
    if (run) {
        HIPCHK(c, hipMemcpyAsync(c->ctl, &h, sizeof h, hipMemcpyHostToDevice, c->st));
        if (launch_fill_u64(c->smax, 2 * CFL_SPREAD * NSLOT, 0ull, c->st)) return fail(c, FINESS_ERR_HIP, "slot reset failed");
        c->devctl = true;
        c->ctlp = &cp;
    }
    struct Snap {
        int in, out, rec_age, rec_par;
        bool fresh;
    };
    // 1D grids of one workgroup without per-step records: the whole run in
    // launches of up to 65536 attempts each, the controller deciding on the
    // device and the kernel rotating the plane sets itself
    if (run && !rc && p.ndims == 1 && !c->steplog && !c->timing && p.time_order <= 3 && run1d_fits(weno_kind(p), p.mx)) {
        Stage1DRun r;
        memset(&r, 0, sizeof r);
        r.base = base_stage1d_args(c);
        r.base.dtp = &c->ctl->dt;
        r.base.halt = &c->ctl->halt;
        for (int k = 0; k < 3; k++) r.buf[k] = c->buf[k];
        r.n = p.time_order;
        r.smax_slot = c->smax;
        r.ctl = c->ctl;
        r.cp = cp;
        r.max_attempts = 65536;
        while (run) {
            r.cur = c->cur;
            if (launch_run1d(weno_kind(p), r, c->st)) { rc = fail(c, FINESS_ERR_HIP, "1D run launch failed"); break; }
            if (hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
                hipStreamSynchronize(c->st) != hipSuccess) {
                rc = fail(c, FINESS_ERR_HIP, "device controller: stream failure");
                break;
            }
            c->cur = (int)h.buf;
            if (h.halt == HALT_NONE) continue;  // the launch's attempts ran out
            if (h.halt != HALT_DONE) rc = fail(c, (int)h.err, ctl_error_message(h.err));
            break;
        }
        run = false;
    }
    std::vector<Snap> snap;
    // the controller of a routed run with per-step records (and no per-step
    // conservation records, which follow the controller) goes with the next
    // step's troubled-face pass where it can (stage_routed)
    const bool fold_ctl = FINESS_FOLD_CTL && p.ndims == 2 && p.time_stepping_method == FINESS_TS_RK &&
                          p.time_order <= 3 && p.routing != FINESS_ROUTE_NONE && p.route_per_step > 0 &&
                          c->route_tile && !c->steplog;
    long long seen = 0;   // attempts processed before the current chunk
    long long logged = 0; // conservation records read back
    long long kmax = 64;  // chunk length: 64, or shorter after a rejection
    while (run) {
        long long K;
        if (by_steps) {
            K = fixed_steps - h.acc;
        } else {
            const double est = (tend - h.t) / h.dt_next;
            K = est < 63.0 ? (long long)ceil(est) + 1 : 64;
        }
        K = K < 1 ? 1 : (K > kmax ? kmax : K);
        snap.clear();
        for (long long k = 0; k < K && !rc; k++) {
            Snap sn;
            sn.in = c->cur;
            int res = 0, slot = 0;
            c->ctl_enqueued = false;
            if ((rc = enqueue_step(c, 0.0, &res, &slot))) break;
            if (!c->ctl_enqueued) {
                if (fold_ctl) {
                    c->ctl_pending = true;
                    c->ctl_pending_slot = slot;
                } else if (launch_step_control(c->ctl, c->smax + 2 * CFL_SPREAD * slot, cp, c->st)) {
                    rc = fail(c, FINESS_ERR_HIP, "controller launch failed");
                    break;
                }
            }
            // the attempt's conservation record, written only if the controller just accepted it
            if (c->steplog && launch_con_soln(c->buf[res], c->plane, c->M, step_log_weight(p), c->spart, c->ctl,
                                              h.nrec, 0.0, c->slog, c->st)) {
                rc = fail(c, FINESS_ERR_HIP, "conservation sum launch failed");
                break;
            }
            sn.out = res;
            sn.rec_age = c->rec_age;
            sn.rec_par = c->rec_par;
            sn.fresh = c->route_fresh;
            snap.push_back(sn);
            c->cur = res;
        }
        if (!rc) rc = flush_ctl(c);
        if (!rc && hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->st) != hipSuccess) rc = FINESS_ERR_HIP;
        if (!rc && hipStreamSynchronize(c->st) != hipSuccess) rc = FINESS_ERR_HIP;
        if (rc) {
            if (rc == FINESS_ERR_HIP && c->err.empty()) c->err = "device controller: stream failure";
            break;
        }
        const long long m = h.done - seen - 1;  // the chunk's last processed attempt
        seen = h.done;
        if (c->steplog && h.nrec > logged) {
            const size_t n = (size_t)(h.nrec - logged) * (c->M + 1), at = c->hlog.size();
            c->hlog.resize(at + n);
            if (hipMemcpy(c->hlog.data() + at, c->slog, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                rc = fail(c, FINESS_ERR_HIP, "conservation log read failed");
                break;
            }
            logged = h.nrec;
        }
        if (h.halt == HALT_NONE) {
            if (kmax < 64) kmax *= 2;
            continue;
        }
        if (m < 0 || m >= (long long)snap.size()) {
            rc = fail(c, FINESS_ERR_HIP, "device controller: inconsistent control block");
            break;
        }
        const Snap& sn = snap[(size_t)m];
        c->rec_age = sn.rec_age;
        c->rec_par = sn.rec_par;
        c->route_fresh = sn.fresh;
        if (h.halt == HALT_DONE) {
            c->cur = sn.out;
            break;
        }
        if (h.halt == HALT_REJECT) {
            c->cur = sn.in;
            kmax = 4;
            if (launch_fill_u64(&c->ctl->halt, 1, (unsigned long long)HALT_NONE, c->st)) {
                rc = fail(c, FINESS_ERR_HIP, "controller restart failed");
                break;
            }
            continue;
        }
        c->cur = h.err == FINESS_ERR_NV ? sn.out : sn.in;
        rc = fail(c, (int)h.err, ctl_error_message(h.err));
        break;
    }
    c->devctl = false;
    c->ctlp = nullptr;
    c->ctl_enqueued = false;
    c->ctl_pending = false;
    if (!rc) dtv[1] = h.dt_next;
    c->t = h.t;
    if (stats) {
        stats->steps_accepted = h.acc;
        stats->steps_rejected = h.rej;
        stats->dt_min = h.dtmin;
        stats->dt_max = h.dtmax;
        stats->last_cfl = h.cfl;
        stats->t = h.t;
    }
    return rc;
}

static int advance_impl(finess_ctx* c, double tend, double* dtv, long long fixed_steps, finess_stats* stats)
{
    if (c && dtv && c->have_state && c->ctl && c->p.time_stepping_method == FINESS_TS_RK &&
        c->p.time_order <= 3 && !c->p.verbosity && !c->p.host_controller && !c->hostbc)
        return advance_dev(c, tend, dtv, fixed_steps, stats);
    if (!c || !dtv) return FINESS_ERR_ARG;
    if (!c->have_state) return fail(c, FINESS_ERR_ARG, "no state");
    const finess_params& p = c->p;
    const double CFL_max = p.max_cfl, CFL_target = p.desired_cfl;
    double t = c->t, dt = dtv[1], cfl = 0.0;
    double dtmin = dt, dtmax = dt;
    long long n_step = 0, acc = 0, rej = 0;
    int rc = 0;
    const bool by_steps = fixed_steps > 0;
    while (by_steps ? (acc < fixed_steps) : (t < tend)) {
        int m_accept = 0;
        n_step = n_step + 1;
        if (!by_steps && n_step > p.nv) { rc = fail(c, FINESS_ERR_NV, "exceeded allowed # of time steps"); break; }
        while (m_accept == 0) {
            const double told = t;
            if (!by_steps && told + dt > tend) dt = tend - told;
            t = told + dt;
            int res = 0, slot = 0;
            c->step_t0 = told;
            if ((rc = enqueue_step(c, dt, &res, &slot))) break;
            if ((rc = read_cfl(c, slot, dt, dtv[2], &cfl))) break;
            if (cfl > 0.0) {
                dt = rmin(dtv[2], dt * CFL_target / cfl);
                dtmin = rmin(dt, dtmin);
                dtmax = rmax(dt, dtmax);
            } else {
                dt = dtv[2];
            }
            if (cfl <= CFL_max) {
                m_accept = 1;
                c->cur = res;
                acc++;
                if (c->steplog && (rc = log_step(c, t))) break;
            } else {
                t = told;
                rej++;
            }
            if (p.verbosity)
                printf("finess-hip ... Step %5lld   CFL =%6.3f   dt =%11.3e   t =%11.3e%s\n", n_step, cfl, dt, t,
                       m_accept ? "" : "  (rejected)");
        }
        if (rc) break;
    }
    if (!rc) dtv[1] = dt;
    c->t = t;
    if (stats) {
        stats->steps_accepted = acc;
        stats->steps_rejected = rej;
        stats->dt_min = dtmin;
        stats->dt_max = dtmax;
        stats->last_cfl = cfl;
        stats->t = t;
    }
    return rc;
}

int finess_advance(finess_ctx* c, double tend, double* dtv, finess_stats* stats)
{
    return advance_impl(c, tend, dtv, 0, stats);
}

int finess_advance_steps(finess_ctx* c, long long nsteps, double* dtv, finess_stats* stats)
{
    if (nsteps < 1) return FINESS_ERR_ARG;
    return advance_impl(c, 0.0, dtv, nsteps, stats);
}

int finess_fin_solve_rk(finess_ctx* c, double* q, const double* aux, double* t, double tend, double* dtv)
{
    if (!c || !q || !t || !dtv) return FINESS_ERR_ARG;
    if (c->p.time_stepping_method != FINESS_TS_RK) return fail(c, FINESS_ERR_CONFIG, "context is not RK");
    int rc = finess_set_state(c, q, aux, *t);
    if (rc) return rc;
    rc = finess_advance(c, tend, dtv, nullptr);
    const int rc2 = finess_get_state(c, q, t);
    return rc ? rc : rc2;
}

int finess_fin_solve_rk_pif(finess_ctx* c, double* q, const double* aux, double* t, double tend, double* dtv)
{
    if (!c || !q || !t || !dtv) return FINESS_ERR_ARG;
    if (c->p.time_stepping_method != FINESS_TS_PIF) return fail(c, FINESS_ERR_CONFIG, "context is not RK4-PIF");
    int rc = finess_set_state(c, q, aux, *t);
    if (rc) return rc;
    rc = finess_advance(c, tend, dtv, nullptr);
    const int rc2 = finess_get_state(c, q, t);
    return rc ? rc : rc2;
}

int finess_fin_solve_lxw(finess_ctx* c, double* q, const double* aux, double* t, double tend, double* dtv)
{
    if (!c || !q || !t || !dtv) return FINESS_ERR_ARG;
    if (c->p.time_stepping_method != FINESS_TS_LXW) return fail(c, FINESS_ERR_CONFIG, "context is not LxW");
    int rc = finess_set_state(c, q, aux, *t);
    if (rc) return rc;
    rc = finess_advance(c, tend, dtv, nullptr);
    const int rc2 = finess_get_state(c, q, t);
    return rc ? rc : rc2;
}

int finess_run_fixed_steps(finess_ctx* c, double dt, int nsteps, double* last_cfl)
{
    if (!c || nsteps < 0) return FINESS_ERR_ARG;
    if (!c->have_state) return fail(c, FINESS_ERR_ARG, "no state");
    int res = c->cur, slot = 0, rc;
    // every stage launch is bracketed by two events: the figure is kernel
    // time only, without the launch gaps and the slot resets between stages
    c->timing = true;
    c->tev_used = 0;
    for (int n = 0; n < nsteps; n++) {
        c->step_t0 = c->t;
        if ((rc = enqueue_step(c, dt, &res, &slot))) { c->timing = false; return rc; }
        c->cur = res;
        c->t += dt;
    }
    c->timing = false;
    HIPCHK(c, hipStreamSynchronize(c->st));
    double ms = 0.0;
    for (size_t k = 0; k + 1 < c->tev_used; k += 2) {
        float e = 0.f;
        HIPCHK(c, hipEventElapsedTime(&e, c->tev[k], c->tev[k + 1]));
        ms += e;
    }
    c->last_ms = ms;
    c->last_launches = (long long)(c->tev_used / 2);
    if (last_cfl && nsteps > 0) {
        double cfl = 0.0;
        if ((rc = read_cfl(c, slot, dt, 1e300, &cfl))) return rc;
        *last_cfl = cfl;
    }
    return FINESS_OK;
}

int finess_set_indicator(finess_ctx* c, const float* params, float threshold)
{
    if (!c || !params) return FINESS_ERR_ARG;
    MlpFrags h;
    mlp_make_frags(params, threshold, &h);
    if (!c->frags) {
        int rc;
        if ((rc = alloc(c, (void**)&c->frags, sizeof(MlpFrags)))) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(c->frags, &h, sizeof h, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FINESS_OK;
}

int finess_route_stats(finess_ctx* c, long long* troubled, long long* faces)
{
    if (!c) return FINESS_ERR_ARG;
    unsigned long long h[2 * CFL_SPREAD];
    HIPCHK(c, hipMemcpyAsync(h, c->route_counts, sizeof h, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (launch_fill_u64(c->route_counts, 2 * CFL_SPREAD, 0ull, c->st)) return fail(c, FINESS_ERR_HIP, "reset failed");
    unsigned long long nt = 0, nf = 0;
    for (int k = 0; k < CFL_SPREAD; k++) {
        nt += h[k];
        nf += h[CFL_SPREAD + k];
    }
    if (troubled) *troubled = (long long)nt;
    if (faces) *faces = (long long)nf;
    return FINESS_OK;
}

// Accessors for the network object (mlp_api.cpp).
hipStream_t finess_ctx_stream(finess_ctx* c) { return c->st; }
const double* finess_ctx_state(finess_ctx* c) { return c->have_state ? c->buf[c->cur] : nullptr; }
const finess_params* finess_ctx_params(finess_ctx* c) { return &c->p; }

int finess_sync(finess_ctx* c)
{
    if (!c) return FINESS_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FINESS_OK;
}

int finess_last_kernel_time(finess_ctx* c, double* ms, long long* launches)
{
    if (!c) return FINESS_ERR_ARG;
    if (ms) *ms = c->last_ms;
    if (launches) *launches = c->last_launches;
    return FINESS_OK;
}

} // extern "C"
