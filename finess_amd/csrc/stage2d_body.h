// stage2d_body.h -- one fused 2D Runge-Kutta / Lax-Wendroff stage.
//
// Replaces, per stage, the reference's sequence
//   SetBndValues(q)                    (app SetBndValues.cpp)
//   ConstructL(q, Lstar, smax)         (lib/2d/ConstructL.cpp:13-408)
//   UpdateSoln(a1, a2, b, dt, ...)     (lib/2d/UpdateSoln.cpp:14-38)
// and the wave-speed part of GetCFL (lib/2d/GetCFL.cpp:54-58).
//
// A workgroup owns a TI x TJ tile of cells (j, the contiguous index, along
// TJ).  Phase 1 stages the tile plus a 3-cell cross-shaped halo in LDS,
// mapping halo cells through the boundary condition on the fly (the halo
// fill) and computing each cell's record (state + primitives) once.
// Phase 2 computes every x-face and y-face flux of the tile: characteristic
// projection at the mean state, local Lax-Friedrichs splitting, WENO5 on the
// 5 characteristic fields for both split parts, projection back.  Faces are
// dealt to threads as one list, x-faces first, so each wave runs the same
// instructions whatever the face direction.  Phase 3 differences the face
// fluxes into L and writes  out = a1*qa + a2*qb + (b dt)*L  in FP64.
// The per-face wave speeds feed a workgroup max that the caller reduces into
// the step's CFL slot.
//
// The phases are plain functions of the thread index so that a CPU harness
// can run the very same code one thread at a time.
#pragma once
#include "physics.h"
#include "weno.h"

namespace finess {

// Boundary condition kinds (values equal the FINESS_BC_* macros).
enum : int { BC_EXTRAP = 0, BC_PERIODIC = 1, BC_REFLECT = 2 };

// Source index along an axis of n cells for a (possibly ghost) index g,
// 0-based; *flip is set for a reflecting wall.  Valid for g in [-n, 2n-1].
__host__ __device__ __forceinline__ int bc_map(int bc_lo, int bc_hi, int g, int n, int& flip)
{
    flip = 0;
    if (g < 0) {
        if (bc_lo == BC_PERIODIC) return g + n;
        if (bc_lo == BC_REFLECT) { flip = 1; return -1 - g; }
        return 0;
    }
    if (g >= n) {
        if (bc_hi == BC_PERIODIC) return g - n;
        if (bc_hi == BC_REFLECT) { flip = 1; return 2 * n - 1 - g; }
        return n - 1;
    }
    return g;
}

struct StageArgs {
    // state planes, field-major: q[m][i][j] at m*plane + i*ld + j (ld = my)
    const double* src;     // stencil operand
    const double* qa;      // UpdateSoln's "qstar" operand (a1)
    const double* qb;      // UpdateSoln's "qnew" operand (a2)
    double* out;
    // Lax-Wendroff: time-integrated fluxes on the extended grid
    // [-3, mx+2] x [-3, my+2], plane (mx+6)*(my+6), F at m*lplane + (i+3)*(my+6) + (j+3)
    const double* lxw_F;
    const double* lxw_G;
    // ghosts stored in src instead of mapped (drop-in ConstructL mode):
    // src planes are (mx+2*gh) x (my+2*gh)
    int src_ghosts;
    int mx, my;
    long long plane;       // mx*my
    int ntj;               // tiles along j
    int bc[4];             // left, right, bottom, top
    int a_is_src, b_is_src, use_b;
    double a1, a2, bdt;
    double dx, dy;
    double alpha_scaling;
    double eps, power;
    PhysParams pp;
    // wave-speed maxima (bit patterns of non-negative doubles), [2] per slot
    unsigned long long* smax;
    // optional face-flux / L output for the drop-in ConstructL (may be null):
    double* Lout;          // L planes (mx*my)
    double* smax_faces_x;  // (mx+1) x my alphas of x-faces, row-major i*(my)+j
    double* smax_faces_y;  // mx x (my+1)
};

template <class P, typename R, int TI, int TJ>
struct StageSmem {
    static constexpr int HI = TI + 6, HJ = TJ + 6, NC = HI * HJ;
    static constexpr int NFX = (TI + 1) * TJ, NFY = TI * (TJ + 1);
    R cv[P::NV][NC];
    R fx[P::MEQN][NFX];
    R gy[P::MEQN][NFY];
};

// Cell load with the boundary mapping (phase 1).
template <class P>
__host__ __device__ __forceinline__ void load_state(const StageArgs& a, int gi, int gj, double q[P::MEQN])
{
    if (a.src_ghosts) {
        const int g = a.src_ghosts;
        const long long ld = a.my + 2 * g;
        const long long pl = (long long)(a.mx + 2 * g) * ld;
        const long long idx = (long long)(gi + g) * ld + (gj + g);
        for (int m = 0; m < P::MEQN; m++) q[m] = a.src[m * pl + idx];
        return;
    }
    int fi, fj;
    const int si = bc_map(a.bc[0], a.bc[1], gi, a.mx, fi);
    const int sj = bc_map(a.bc[2], a.bc[3], gj, a.my, fj);
    const long long idx = (long long)si * a.my + sj;
    for (int m = 0; m < P::MEQN; m++) q[m] = a.src[m * a.plane + idx];
    if (fi) q[P::NORMAL_X] = -q[P::NORMAL_X];
    if (fj) q[P::NORMAL_Y] = -q[P::NORMAL_Y];
}

template <class P, typename R, int TI, int TJ>
__host__ __device__ __forceinline__ void phase_load(int tid, int nthreads, int i0, int j0, const StageArgs& a,
                                                    StageSmem<P, R, TI, TJ>& sm)
{
    using S = StageSmem<P, R, TI, TJ>;
    for (int c = tid; c < S::NC; c += nthreads) {
        const int ci = c / S::HJ, cj = c - ci * S::HJ;
        const bool ri = ci < 3 || ci >= TI + 3, rj = cj < 3 || cj >= TJ + 3;
        if (ri && rj) continue;                       // corners are never read
        const int gi = i0 + ci - 3, gj = j0 + cj - 3;
        if (gi > a.mx + 2 || gj > a.my + 2) continue; // beyond every face's stencil
        double q[P::MEQN], rec[P::NV];
        load_state<P>(a, gi, gj, q);
        P::cell_record(q, a.pp, rec);
        for (int v = 0; v < P::NV; v++) sm.cv[v][c] = R(rec[v]);
    }
}

// One face: characteristic WENO-LLF flux (lib/2d/ConstructL.cpp:63-208 for
// x, :215-358 for y).  base/stride locate the 6 stencil records i-3..i+2.
template <class P, typename R, int WK, bool LXW, int NC>
__host__ __device__ __forceinline__ R face_flux(const R (*cv)[NC], int base, int stride, int dir,
                                                const double* lflux, long long lbase, long long lstride,
                                                long long lplane, const PhysParams& pp,
                                                const WenoParams<R>& wp, R ascale, R fh[P::MEQN])
{
    constexpr int M = P::MEQN, NV = P::NV;
    R qs[6][M], fs[6][M];
    R cl[NV], cr[NV];
    for (int s = 0; s < 6; s++) {
        R rec[NV];
        for (int v = 0; v < NV; v++) rec[v] = cv[v][base + s * stride];
        P::state_flux(rec, dir, qs[s], fs[s]);
        if (s == 2) for (int v = 0; v < NV; v++) cl[v] = rec[v];
        if (s == 3) for (int v = 0; v < NV; v++) cr[v] = rec[v];
        if (LXW) {
            for (int k = 0; k < M; k++)
                fs[s][k] = R(lflux[P::rot_q(k, dir) * lplane + lbase + s * lstride]);
        }
    }
    R qa[M];
    for (int k = 0; k < M; k++) qa[k] = R(0.5) * (qs[3][k] + qs[2][k]);
    typename P::template Eig<R> e;
    P::eig(qa, pp, e);
    const R alpha = P::wave_speed(cl, cr, dir, pp);
    const R la = ascale * alpha;
    // Lax-Friedrichs split projected in one go: the reference forms
    // gp = (L f + la L q)/2 and gm = (L f - la L q)/2 (ConstructL.cpp:177-185);
    // L is linear, so L(f + la q) and L(f - la q) are projected instead, and
    // the 1/2 is applied after the reconstruction: every WENO5 variant
    // satisfies W(2g; 4 eps) = 2 W(g; eps) exactly in binary arithmetic.
    R gp[M][5], gm[M][5];
    for (int s = 0; s < 6; s++) {
        R vp[M], vm[M], t[M];
        for (int k = 0; k < M; k++) {
            vp[k] = fs[s][k] + la * qs[s][k];
            vm[k] = fs[s][k] - la * qs[s][k];
        }
        if (s < 5) {
            P::left(e, vp, t);
            for (int k = 0; k < M; k++) gp[k][s] = t[k];
        }
        if (s > 0) {
            P::left(e, vm, t);
            for (int k = 0; k < M; k++) gm[k][5 - s] = t[k];
        }
    }
    const WenoParams<R> wp4{R(4) * wp.eps, wp.power};
    R gh[M];
    for (int k = 0; k < M; k++) gh[k] = R(0.5) * (weno5<WK>(gp[k], wp4) + weno5<WK>(gm[k], wp4));
    R fr[M];
    P::right(e, gh, fr);
    for (int k = 0; k < M; k++) fh[P::rot_q(k, dir)] = fr[k];
    return alpha;
}

template <class P, typename R, int WK, bool LXW, int TI, int TJ>
__host__ __device__ __forceinline__ void phase_faces(int tid, int nthreads, int i0, int j0, const StageArgs& a,
                                                     StageSmem<P, R, TI, TJ>& sm, double& ax, double& ay)
{
    using S = StageSmem<P, R, TI, TJ>;
    constexpr int M = P::MEQN;
    const WenoParams<R> wp{R(a.eps), R(a.power)};
    const R ascale = R(a.alpha_scaling);
    const long long lld = a.my + 6, lpl = (long long)(a.mx + 6) * lld;
    for (int f = tid; f < S::NFX + S::NFY; f += nthreads) {
        int dir, fi, fj, base, stride;
        long long lbase = 0, lstride = 0;
        bool valid, count;
        if (f < S::NFX) {
            dir = 0; fi = f / TJ; fj = f - fi * TJ;
            base = fi * S::HJ + (fj + 3); stride = S::HJ;
            const int gi = i0 + fi, gj = j0 + fj;
            valid = gi <= a.mx && gj < a.my;
            count = gi < a.mx;                         // GetCFL skips face mx+1/2
            lbase = (long long)gi * lld + (gj + 3); lstride = lld;
        } else {
            const int h = f - S::NFX;
            dir = 1; fi = h / (TJ + 1); fj = h - fi * (TJ + 1);
            base = (fi + 3) * S::HJ + fj; stride = 1;
            const int gi = i0 + fi, gj = j0 + fj;
            valid = gi < a.mx && gj <= a.my;
            count = gj < a.my;
            lbase = (long long)(gi + 3) * lld + gj; lstride = 1;
        }
        if (!valid) continue;
        R fh[M];
        const R alpha = face_flux<P, R, WK, LXW, S::NC>(sm.cv, base, stride, dir, dir ? a.lxw_G : a.lxw_F,
                                                        lbase, lstride, lpl, a.pp, wp, ascale, fh);
        if (dir == 0) {
            for (int m = 0; m < M; m++) sm.fx[m][f] = fh[m];
            if (count) ax = rmax(ax, double(alpha));
            if (a.smax_faces_x) a.smax_faces_x[(long long)(i0 + fi) * a.my + (j0 + fj)] = double(alpha);
        } else {
            for (int m = 0; m < M; m++) sm.gy[m][f - S::NFX] = fh[m];
            if (count) ay = rmax(ay, double(alpha));
            if (a.smax_faces_y) a.smax_faces_y[(long long)(i0 + fi) * (a.my + 1) + (j0 + fj)] = double(alpha);
        }
    }
}

template <class P, typename R, int TI, int TJ>
__host__ __device__ __forceinline__ void phase_update(int tid, int i0, int j0, const StageArgs& a,
                                                      StageSmem<P, R, TI, TJ>& sm)
{
    using S = StageSmem<P, R, TI, TJ>;
    constexpr int M = P::MEQN;
    const int ci = tid / TJ, cj = tid - ci * TJ;
    const int gi = i0 + ci, gj = j0 + cj;
    if (ci >= TI || gi >= a.mx || gj >= a.my) return;
    const long long idx = (long long)gi * a.my + gj;
    const int c = (ci + 3) * S::HJ + (cj + 3);
    const bool exact_lds = sizeof(R) == sizeof(double) && !a.src_ghosts;
    const double idx_ = 1.0 / a.dx, idy_ = 1.0 / a.dy;
    for (int m = 0; m < M; m++) {
        // Lstar = -(Fhat(i+1)-Fhat(i))/dx - (Ghat(j+1)-Ghat(j))/dy  (ConstructL.cpp:388-399)
        double L = -(double(sm.fx[m][(ci + 1) * TJ + cj]) - double(sm.fx[m][ci * TJ + cj])) * idx_;
        L = L - (double(sm.gy[m][ci * (TJ + 1) + cj + 1]) - double(sm.gy[m][ci * (TJ + 1) + cj])) * idy_;
        if (a.Lout) a.Lout[m * a.plane + idx] = L;
        if (!a.out) continue;
        const double qa = (exact_lds && a.a_is_src) ? double(sm.cv[m][c]) : a.qa[m * a.plane + idx];
        double v = a.a1 * qa;
        if (a.use_b) {
            const double qb = (exact_lds && a.b_is_src) ? double(sm.cv[m][c]) : a.qb[m * a.plane + idx];
            v = v + a.a2 * qb;
        }
        a.out[m * a.plane + idx] = v + a.bdt * L;
    }
}

} // namespace finess
