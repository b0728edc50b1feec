// emu_host.cpp -- runs the device kernels' phase functions on the CPU, one
// workgroup and one thread at a time (test infrastructure: lets the CPU test
// suite check the kernels' tiling, halo mapping and physics against the
// oracle without a GPU).
#include <hip/hip_runtime.h>
#include <string.h>
#include <vector>

#include "../../include/finess_hip.h"
#include "../../finess_amd/csrc/lxw_body.h"
#include "../../finess_amd/csrc/stage1d_body.h"
#include "../../finess_amd/csrc/stage2d_body.h"
#include "../../finess_amd/csrc/indicator.h"

using namespace finess;

namespace {
constexpr int TI = 7, TJ = 32, NT = 256;
constexpr int LTI = 8, LTJ = 32;  // tile of the Lax-Wendroff flux kernel

int wk_of(const finess_params* p)
{
    if (p->weno_version == FINESS_WENO_FD) return WENO_FD5;
    if (p->weno_version == FINESS_WENO_Z) return WENO_Z5;
    return p->power_param == 2.0 ? WENO_JS5_P2 : WENO_JS5_POW;
}

std::vector<double> to_planes_ghosted(const finess_params* p, const double* q, int M)
{
    const int g = p->mbc, ld = p->my + 2 * g, n = (p->mx + 2 * g) * ld;
    std::vector<double> out((size_t)M * n);
    for (int k = 0; k < n; k++)
        for (int m = 0; m < M; m++) out[(size_t)m * n + k] = q[(size_t)k * M + m];
    return out;
}

template <class P, typename R, int WK, bool LXW = false>
void run_tile(const StageArgs& a, int i0, int j0, StageSmem<P, R, TI, TJ>& sm)
{
    memset(&sm, 0, sizeof sm);
    std::vector<double> Lx((size_t)NT * P::MEQN, 0.0);
    for (int t = 0; t < NT; t++) phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, sm);
    double ax = 0, ay = 0;
    for (int t = 0; t < NT; t++) phase_faces<P, R, WK, LXW, TI, TJ>(t, 0, i0, j0, a, sm, ax);
    for (int t = 0; t < NT; t++) phase_lx<P, R, TI, TJ>(t, a, sm, &Lx[(size_t)t * P::MEQN]);
    for (int t = 0; t < NT; t++) phase_faces<P, R, WK, LXW, TI, TJ>(t, 1, i0, j0, a, sm, ay);
    for (int t = 0; t < NT; t++) {
        double pa[P::MEQN], pb[P::MEQN];
        phase_prefetch<P, R, TI, TJ>(t, i0, j0, a, pa, pb);
        phase_update<P, R, TI, TJ>(t, i0, j0, a, sm, &Lx[(size_t)t * P::MEQN], pa, pb);
    }
}

template <class P, typename R, int WK, bool LXW = false>
void run_tiles(const StageArgs& a0)
{
    StageArgs a = a0;
    const int nti = (a.mx + TI - 1) / TI, ntj = (a.my + TJ - 1) / TJ;
    a.ntj = ntj;
    std::vector<StageSmem<P, R, TI, TJ>> smv(1);
    for (int t = 0; t < nti * ntj; t++) run_tile<P, R, WK, LXW>(a, (t / ntj) * TI, (t % ntj) * TJ, smv[0]);
}

template <class P, typename R>
void dispatch(int wk, const StageArgs& a)
{
    switch (wk) {
    case WENO_FD5: run_tiles<P, R, WENO_FD5>(a); break;
    case WENO_Z5: run_tiles<P, R, WENO_Z5>(a); break;
    case WENO_JS5_POW: run_tiles<P, R, WENO_JS5_POW>(a); break;
    default: run_tiles<P, R, WENO_JS5_P2>(a); break;
    }
}
} // namespace

extern "C" {

// ConstructL through the kernel's phases (drop-in mode: ghosts read from q).
int emu_construct_l(const finess_params* p, const double* q, double* Lstar, double* smax)
{
    const int M = p->meqn, g = p->mbc;
    if (p->ndims == 1) {
        const int L = p->mx + 2 * g;
        std::vector<double> qs(L), aux(L, 1.0), fh(p->mx + 1), al(p->mx + 1), Lo(p->mx);
        for (int i = 0; i < L; i++) qs[i] = q[(size_t)i * M];
        Stage1DArgs a;
        memset(&a, 0, sizeof a);
        a.src = qs.data(); a.aux = aux.data(); a.fhat = fh.data(); a.mx = p->mx;
        a.src_ghosts = g; a.dx = p->dx; a.alpha_scaling = p->alpha_scaling; a.eps = p->epsilon;
        a.power = p->power_param; a.smax_faces = al.data(); a.Lout = Lo.data();
        const int wk = wk_of(p);
        for (int f = 0; f <= p->mx; f++) {
            if (wk == WENO_FD5) face1d<WENO_FD5>(f, a);
            else if (wk == WENO_Z5) face1d<WENO_Z5>(f, a);
            else if (wk == WENO_JS5_POW) face1d<WENO_JS5_POW>(f, a);
            else face1d<WENO_JS5_P2>(f, a);
        }
        for (int i = 0; i < p->mx; i++) update1d(i, a);
        for (int i = 0; i < p->mx; i++) Lstar[(size_t)(i + g) * M] = Lo[i];
        for (int f = 0; f <= p->mx; f++) smax[f + g] = al[f];
        return 0;
    }
    std::vector<double> qg = to_planes_ghosted(p, q, M);
    const long long plane = (long long)p->mx * p->my;
    std::vector<double> Lo((size_t)M * plane), ax((size_t)(p->mx + 1) * p->my), ay((size_t)p->mx * (p->my + 1));
    StageArgs a;
    memset(&a, 0, sizeof a);
    a.src = qg.data(); a.src_ghosts = g; a.mx = p->mx; a.my = p->my; a.plane = plane;
    for (int k = 0; k < 4; k++) a.bc[k] = p->bc[k];
    a.dx = p->dx; a.dy = p->dy; a.idx = 1.0 / p->dx; a.idy = 1.0 / p->dy;
        a.alpha_scaling = p->alpha_scaling; a.eps = p->epsilon; a.power = p->power_param;
    a.pp.gamma = p->gamma; a.pp.sw_compat_hl = p->sw_compat_hl;
    a.Lout = Lo.data(); a.smax_faces_x = ax.data(); a.smax_faces_y = ay.data();
    const int wk = wk_of(p);
    const bool f32 = p->precision == FINESS_PREC_FP32;
    if (p->physics == FINESS_PHYS_EULER) {
        if (f32) run_tiles<Euler2D, float, WENO_JS5_P2>(a); else dispatch<Euler2D, double>(wk, a);
    } else {
        if (f32) run_tiles<ShallowWater2D, float, WENO_JS5_P2>(a); else dispatch<ShallowWater2D, double>(wk, a);
    }
    const long long ldq = p->my + 2 * g, lds = p->my + 1 + 2 * g;
    for (int i = 0; i < p->mx; i++)
        for (int j = 0; j < p->my; j++)
            for (int m = 0; m < M; m++)
                Lstar[((i + g) * ldq + (j + g)) * M + m] = Lo[(size_t)m * plane + (size_t)i * p->my + j];
    for (int i = 0; i <= p->mx; i++)
        for (int j = 0; j < p->my; j++) {
            double& s = smax[((i + g) * lds + (j + g)) * 2];
            s = rmax(s, ax[(size_t)i * p->my + j]);
        }
    for (int i = 0; i < p->mx; i++)
        for (int j = 0; j <= p->my; j++) {
            double& s = smax[((i + g) * lds + (j + g)) * 2 + 1];
            s = rmax(s, ay[(size_t)i * (p->my + 1) + j]);
        }
    return 0;
}

// ConstructIntegratedR through the LxW kernel's phases.
int emu_construct_integrated_r(const finess_params* p, double dt, const double* q, double* F, double* G)
{
    const int M = p->meqn, g = p->mbc;
    std::vector<double> qg = to_planes_ghosted(p, q, M);
    const long long le = p->my + 6, lpl = (long long)(p->mx + 6) * le;
    std::vector<double> Fo((size_t)M * lpl), Go((size_t)M * lpl);
    LxwArgs a;
    memset(&a, 0, sizeof a);
    a.q = qg.data(); a.src_ghosts = g; a.F = Fo.data(); a.G = Go.data(); a.mx = p->mx; a.my = p->my;
    a.plane = (long long)p->mx * p->my; a.time_order = p->time_order; a.dt = dt; a.dx = p->dx; a.dy = p->dy;
    a.pp.gamma = p->gamma;
    for (int k = 0; k < 4; k++) a.bc[k] = p->bc[k];
    const int nti = (p->mx + 6 + LTI - 1) / LTI, ntj = (p->my + 6 + LTJ - 1) / LTJ;
    std::vector<LxwSmem<LTI, LTJ, 5>> smv(1);
    for (int t = 0; t < nti * ntj; t++) {
        const int e0i = (t / ntj) * LTI, e0j = (t % ntj) * LTJ;
        for (int tid = 0; tid < LTI * LTJ; tid++) lxw_phase_load<Euler2D, LTI, LTJ>(tid, LTI * LTJ, e0i, e0j, a, smv[0]);
        for (int tid = 0; tid < LTI * LTJ; tid++) lxw_phase_cell<Euler2D, LTI, LTJ>(tid, e0i, e0j, a, smv[0]);
    }
    const long long ldq = p->my + 2 * g;
    for (int i = -3; i < p->mx + 3; i++)
        for (int j = -3; j < p->my + 3; j++)
            for (int m = 0; m < M; m++) {
                const long long o = (long long)m * lpl + (long long)(i + 3) * le + (j + 3);
                F[((i + g) * ldq + (j + g)) * M + m] = Fo[o];
                G[((i + g) * ldq + (j + g)) * M + m] = Go[o];
            }
    return 0;
}

// Fixed steps (no controller) through the stepping path: interior planes,
// halo produced by the kernels' boundary map, 3-buffer SSP-RK / LxW update.
int emu_run_fixed_steps(const finess_params* p, double* q, double dt, int nsteps)
{
    if (p->ndims != 2) return -1;
    const int M = p->meqn, g = p->mbc;
    const long long plane = (long long)p->mx * p->my, ldq = p->my + 2 * g;
    std::vector<double> buf[3];
    for (int k = 0; k < 3; k++) buf[k].assign((size_t)M * plane, 0.0);
    for (int i = 0; i < p->mx; i++)
        for (int j = 0; j < p->my; j++)
            for (int m = 0; m < M; m++)
                buf[0][(size_t)m * plane + (size_t)i * p->my + j] = q[((i + g) * ldq + (j + g)) * M + m];
    const bool lxw = p->time_stepping_method == FINESS_TS_LXW;
    const long long le = p->my + 6, lpl = (long long)(p->mx + 6) * le;
    std::vector<double> Fo, Go;
    if (lxw) { Fo.assign((size_t)M * lpl, 0.0); Go.assign((size_t)M * lpl, 0.0); }
    int cur = 0;
    auto stage = [&](int src, int qa, int qb, int out, double a1, double a2, double beta) {
        StageArgs a;
        memset(&a, 0, sizeof a);
        a.src = buf[src].data(); a.qa = buf[qa].data(); a.a_is_src = qa == src;
        a.qb = qb >= 0 ? buf[qb].data() : nullptr; a.b_is_src = qb == src; a.use_b = qb >= 0 && a2 != 0.0;
        a.out = buf[out].data(); a.a1 = a1; a.a2 = a2; a.bdt = beta * dt;
        a.mx = p->mx; a.my = p->my; a.plane = plane;
        for (int k = 0; k < 4; k++) a.bc[k] = p->bc[k];
        a.dx = p->dx; a.dy = p->dy; a.idx = 1.0 / p->dx; a.idy = 1.0 / p->dy;
        a.alpha_scaling = p->alpha_scaling; a.eps = p->epsilon; a.power = p->power_param;
        a.pp.gamma = p->gamma; a.pp.sw_compat_hl = p->sw_compat_hl;
        if (lxw) { a.lxw_F = Fo.data(); a.lxw_G = Go.data(); }
        if (p->physics == FINESS_PHYS_EULER) {
            if (lxw) run_tiles<Euler2D, double, WENO_JS5_P2, true>(a);
            else run_tiles<Euler2D, double, WENO_JS5_P2, false>(a);
        } else {
            run_tiles<ShallowWater2D, double, WENO_JS5_P2, false>(a);
        }
    };
    for (int n = 0; n < nsteps; n++) {
        const int A = cur, B = (cur + 1) % 3, C = (cur + 2) % 3;
        if (lxw) {
            LxwArgs la;
            memset(&la, 0, sizeof la);
            la.q = buf[A].data(); la.F = Fo.data(); la.G = Go.data(); la.mx = p->mx; la.my = p->my; la.plane = plane;
            for (int k = 0; k < 4; k++) la.bc[k] = p->bc[k];
            la.time_order = p->time_order; la.dt = dt; la.dx = p->dx; la.dy = p->dy; la.pp.gamma = p->gamma;
            const int nti = (p->mx + 6 + LTI - 1) / LTI, ntj = (p->my + 6 + LTJ - 1) / LTJ;
            std::vector<LxwSmem<LTI, LTJ, 5>> smv(1);
            for (int t = 0; t < nti * ntj; t++) {
                const int e0i = (t / ntj) * LTI, e0j = (t % ntj) * LTJ;
                for (int tid = 0; tid < LTI * LTJ; tid++) lxw_phase_load<Euler2D, LTI, LTJ>(tid, LTI * LTJ, e0i, e0j, la, smv[0]);
                for (int tid = 0; tid < LTI * LTJ; tid++) lxw_phase_cell<Euler2D, LTI, LTJ>(tid, e0i, e0j, la, smv[0]);
            }
            stage(A, A, -1, B, 1.0, 0.0, 1.0);
            cur = B;
            continue;
        }
        stage(A, A, -1, B, 1.0, 0.0, 1.0);
        stage(B, A, B, C, 0.75, 0.25, 0.25);
        stage(C, C, A, B, 2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0);
        cur = B;
    }
    for (int i = 0; i < p->mx; i++)
        for (int j = 0; j < p->my; j++)
            for (int m = 0; m < M; m++)
                q[((i + g) * ldq + (j + g)) * M + m] = buf[cur][(size_t)m * plane + (size_t)i * p->my + j];
    return 0;
}

// ---- troubled-interface indicator: the device functions on one 6-cell
// stencil qs[6][5] (states q(i-3..i+2) across the face), dir 1 = x, 2 = y.
static void emu_records(const finess_params* p, const double* qs, double cv[8][6])
{
    PhysParams pp{};
    pp.gamma = p->gamma;
    for (int s = 0; s < 6; s++) {
        double rec[8];
        Euler2D::cell_record(qs + 5 * s, pp, rec);
        for (int v = 0; v < 8; v++) cv[v][s] = rec[v];
    }
}

void emu_face_features(const finess_params* p, int dir, const double* qs, int fp32, float* feat)
{
    PhysParams pp{};
    pp.gamma = p->gamma;
    double cv[8][6];
    emu_records(p, qs, cv);
    if (fp32) {
        float rec[6][8];
        for (int s = 0; s < 6; s++) for (int v = 0; v < 8; v++) rec[s][v] = (float)cv[v][s];
        face_features<Euler2D, float>(rec, dir - 1, pp, feat);
    } else {
        double rec[6][8];
        for (int s = 0; s < 6; s++) for (int v = 0; v < 8; v++) rec[s][v] = cv[v][s];
        face_features<Euler2D, double>(rec, dir - 1, pp, feat);
    }
}

double emu_face_departure(const finess_params* p, int dir, const double* qs)
{
    PhysParams pp{};
    pp.gamma = p->gamma;
    double cv[8][6], rec[6][8];
    emu_records(p, qs, cv);
    for (int s = 0; s < 6; s++) for (int v = 0; v < 8; v++) rec[s][v] = cv[v][s];
    return face_departure<Euler2D, double>(rec, dir - 1, pp, p->epsilon, p->alpha_scaling);
}

// fh in the unrotated component order, as the stage kernel stores it
double emu_face_flux_linear(const finess_params* p, int dir, const double* qs, double* fh)
{
    PhysParams pp{};
    pp.gamma = p->gamma;
    double cv[8][6];
    emu_records(p, qs, cv);
    return face_flux_linear<Euler2D, double, 6>(cv, 0, 1, dir - 1, pp, p->alpha_scaling, fh);
}

} // extern "C"
