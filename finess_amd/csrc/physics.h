// physics.h -- device restatement of the app callback libraries the kernels
// are compiled with.  Every 2D function works in the frame rotated to the
// face normal: for an x-face the conserved vector is used as stored, for a
// y-face the two in-plane momentum components are swapped.  The reference's
// callbacks do the same with their (mu, mv) index pairs (ixy = 1: mu=2, mv=3;
// ixy = 2: mu=3, mv=2), so the rotated x-formulas below are theirs.
//
//   Euler:         FluxFunc        apps/2d/euler/lib/FluxFunc.cpp:9-42
//                  ProjectLeftEig  apps/2d/euler/lib/ProjectLeftEig.cpp:9-67
//                  ProjectRightEig apps/2d/euler/lib/ProjectRightEig.cpp:9-63
//                  SetWaveSpd      apps/2d/euler/lib/SetWaveSpd.cpp:11-69
//   Shallow water: apps/2d/shallow_water/lib/{FluxFunc,ProjectLeftEig,
//                  ProjectRightEig,SetWaveSpd}.cpp  (g = 1)
//
// Per-cell values kept in LDS ("cell record"): the conserved state followed
// by the primitive values the flux needs, so each cell's divisions are done
// once per stage and not once per stencil point.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "fastmath.h"

namespace finess {

template <typename R>
__host__ __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }  // lib/dog_math.h:18-22
template <typename R>
__host__ __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }

struct PhysParams {
    double gamma;
    int sw_compat_hl;
};

// ------------------------------------------------------------------------
// 2D Euler, meqn = 5: (rho, rho u, rho v, rho w, E)
// cell record: q0..q4, u, v, w, p
// ------------------------------------------------------------------------
struct Euler2D {
    static constexpr int MEQN = 5;
    static constexpr int NV = 8;
    static constexpr int NORMAL_X = 1;  // momentum component normal to x-walls
    static constexpr int NORMAL_Y = 2;

    // cell record: q0..q4, 1/rho, p, c  with c = sqrt(|gamma p / rho|), the
    // cell sound speed SetWaveSpd uses for the left and right states
    // (apps/2d/euler/lib/SetWaveSpd.cpp:47-49).  Formed in FP64.
    // LP (the FP32 stencil form, whose records are rounded to FP32): 1/rho,
    // the kinetic energy |m|^2 / (2 rho) from the momenta, p and c in FP32,
    // except in near-vacuum cells, where E - |m|^2/(2 rho) cancels more than
    // two bits (p < E/4): there the pressure is recovered in FP64.
    template <bool LP = false>
    __host__ __device__ __forceinline__ static void cell_record(const double q[5], const PhysParams& pp, double v[8])
    {
        const double gamma = pp.gamma;
        v[0] = q[0]; v[1] = q[1]; v[2] = q[2]; v[3] = q[3]; v[4] = q[4];
        if constexpr (LP) {
            const float r = float(q[0]), m1 = float(q[1]), m2 = float(q[2]), m3 = float(q[3]), e = float(q[4]);
            const float ir = frcp(r);
            const float ke = 0.5f * ((m1 * m1 + m2 * m2 + m3 * m3) * ir);
            const float gm1 = float(gamma - 1.0);
            float press;
            if (e - ke >= 0.25f * e) {
                press = gm1 * (e - ke);
            } else {  // near vacuum (or a non-finite state): FP64
                const double irho = rcp_nr(q[0]);
                press = float((gamma - 1.0) * (q[4] - 0.5 * ((q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * irho)));
            }
            v[5] = ir; v[6] = press;
            v[7] = sqrt_lp(fabsf(float(gamma) * press * ir));
            return;
        }
        const double rho = q[0];
        const double irho = 1.0 / rho;
        const double u1 = q[1] * irho, u2 = q[2] * irho, u3 = q[3] * irho;
        const double press = (gamma - 1.0) * (q[4] - 0.5 * rho * (u1 * u1 + u2 * u2 + u3 * u3));
        v[5] = irho; v[6] = press;
        v[7] = sqrt(fabs(gamma * press * irho));
    }

    // LDS variable index of rotated component k (dir 0: x, 1: y)
    __host__ __device__ __forceinline__ static int rot_q(int k, int dir)
    {
        return (dir && (k == 1 || k == 2)) ? 3 - k : k;
    }

    // Rotated state of one cell record.
    template <typename R>
    __host__ __device__ __forceinline__ static void state(const R c[8], int dir, R q[5])
    {
        q[0] = c[0]; q[1] = dir ? c[2] : c[1]; q[2] = dir ? c[1] : c[2]; q[3] = c[3]; q[4] = c[4];
    }

    // Rotated flux of a rotated state with its record (FluxFunc.cpp:20-40).
    template <typename R>
    __host__ __device__ __forceinline__ static void flux(const R q[5], const R c[8], R f[5])
    {
        const R irho = c[5], p = c[6];
        const R un = q[1] * irho, ut = q[2] * irho, uw = q[3] * irho;
        const R run = q[0] * un;
        f[0] = run;
        f[1] = run * un + p;
        f[2] = run * ut;
        f[3] = run * uw;
        f[4] = un * (q[4] + p);
    }

    // Eigen-system at the arithmetic-mean state (lib/2d/ConstructL.cpp:72-77)
    // with the left-eigenvector rows of ProjectLeftEig.cpp:49-65 pre-scaled by
    // their denominators:
    //   w1 = (a1 q_mu + b1 q_1 + cc s),  w5 = (a5 q_mu + b5 q_1 + cc s),
    //   s  = q_5 - u2 q_mv - u3 q_mw,
    //   w_mu = k2 q_1 + e1 q_mu + e2 q_mv + e3 q_mw + m5 q_5,
    //   w_mv = q_mv - u2 q_1,  w_mw = q_mw - u3 q_1.
    template <typename R>
    struct Eig {
        R u1, u2, u3, hu, H, cR;
        R a1, b1, a5, b5, cc, k2, e1, e2, e3, m5;
    };

    template <typename R>
    __host__ __device__ __forceinline__ static void eig(const R qa[5], const PhysParams& pp, Eig<R>& e)
    {
        const R gamma = R(pp.gamma);
        const R rho = qa[0];
        const R irho = div_lp(R(1), rho);
        e.u1 = qa[1] * irho; e.u2 = qa[2] * irho; e.u3 = qa[3] * irho;
        const R energy = qa[4];
        const R umag2 = e.u1 * e.u1 + e.u2 * e.u2 + e.u3 * e.u3;
        const R press = (gamma - R(1)) * (energy - R(0.5) * rho * umag2);
        const R c2 = gamma * press * irho;
        e.cR = sqrt_lp(fabs(c2));                                // ProjectRightEig.cpp:42
        const R c = c2 >= R(0) ? e.cR : R(__builtin_nan(""));   // ProjectLeftEig.cpp:43
        e.H = (energy + press) * irho;
        e.hu = umag2 * R(0.5);
        const R den = R(2) * e.H - umag2;
        const R r = div_lp(R(1), c * den);
        const R inv2 = R(2) * c * r;                             // 2 / den
        e.a1 = (e.hu - e.H - e.u1 * c) * r;
        e.b1 = (e.hu * (c - e.u1) + e.H * e.u1) * r;
        e.a5 = (e.H - e.hu - e.u1 * c) * r;
        e.b5 = (e.hu * (c + e.u1) - e.H * e.u1) * r;
        e.cc = c * r;
        e.k2 = (e.H - umag2) * inv2;
        e.e1 = e.u1 * inv2; e.e2 = e.u2 * inv2; e.e3 = e.u3 * inv2;
        e.m5 = -inv2;
    }

    template <typename R>
    __host__ __device__ __forceinline__ static void left(const Eig<R>& e, const R q[5], R w[5])
    {
        const R s = q[4] - e.u2 * q[2] - e.u3 * q[3];
        const R t = e.cc * s;
        w[0] = e.a1 * q[1] + e.b1 * q[0] + t;
        w[1] = e.k2 * q[0] + e.e1 * q[1] + e.e2 * q[2] + e.e3 * q[3] + e.m5 * q[4];
        w[2] = q[2] - e.u2 * q[0];
        w[3] = q[3] - e.u3 * q[0];
        w[4] = e.a5 * q[1] + e.b5 * q[0] + t;
    }

    // Row k of left() alone (the same expression), for a field index that
    // is uniform over the wave.
    template <typename R>
    __host__ __device__ __forceinline__ static R left_row(const Eig<R>& e, const R q[5], int k)
    {
        switch (k) {
        case 0: return e.a1 * q[1] + e.b1 * q[0] + e.cc * (q[4] - e.u2 * q[2] - e.u3 * q[3]);
        case 1: return e.k2 * q[0] + e.e1 * q[1] + e.e2 * q[2] + e.e3 * q[3] + e.m5 * q[4];
        case 2: return q[2] - e.u2 * q[0];
        case 3: return q[3] - e.u3 * q[0];
        default: return e.a5 * q[1] + e.b5 * q[0] + e.cc * (q[4] - e.u2 * q[2] - e.u3 * q[3]);
        }
    }

    template <typename R>
    __host__ __device__ __forceinline__ static void right(const Eig<R>& e, const R w[5], R q[5])
    {
        const R s = w[0] + w[1] + w[4];
        const R uc = e.u1 * e.cR;
        q[0] = s;
        q[1] = (e.u1 - e.cR) * w[0] + e.u1 * w[1] + (e.u1 + e.cR) * w[4];
        q[2] = e.u2 * s + w[2];
        q[3] = e.u3 * s + w[3];
        q[4] = (e.H - uc) * w[0] + e.hu * w[1] + e.u2 * w[2] + e.u3 * w[3] + (e.H + uc) * w[4];
    }

    // SetWaveSpd with nvec = the face normal; ql, qr the rotated states and
    // cl, cr the records of the left and right cells.  Returns max(|s1|,|s2|).
    template <typename R>
    __host__ __device__ __forceinline__ static R wave_speed(const R ql[5], const R qr[5], const R cl[8], const R cr[8],
                                                            const PhysParams& pp)
    {
        const R gamma = R(pp.gamma);
        const R unl = ql[1] * cl[5], unr = qr[1] * cr[5];
        const R rho = R(0.5) * (ql[0] + qr[0]);
        const R un = R(0.5) * (unl + unr);
        const R press = R(0.5) * (cl[6] + cr[6]);
        const R c = sqrt_lp(fabs(div_lp(gamma * press, rho)));
        const R s1 = rmin(unl - cl[7], un - c);
        const R s2 = rmax(unr + cr[7], un + c);
        return rmax(fabs(s1), fabs(s2));
    }
};

// ------------------------------------------------------------------------
// 2D shallow water (g = 1), meqn = 3: (h, hu, hv)
// cell record: q0..q2, u, v
// ------------------------------------------------------------------------
struct ShallowWater2D {
    static constexpr int MEQN = 3;
    static constexpr int NV = 5;
    static constexpr int NORMAL_X = 1;
    static constexpr int NORMAL_Y = 2;

    // LP (FP32 stencil form): the velocities in FP32.
    template <bool LP = false>
    __host__ __device__ __forceinline__ static void cell_record(const double q[3], const PhysParams&, double v[5])
    {
        v[0] = q[0]; v[1] = q[1]; v[2] = q[2];
        if constexpr (LP) {
            const float ih = frcp(float(q[0]));
            v[3] = float(q[1]) * ih; v[4] = float(q[2]) * ih;
            return;
        }
        const double ih = 1.0 / q[0];
        v[3] = q[1] * ih; v[4] = q[2] * ih;
    }

    __host__ __device__ __forceinline__ static int rot_q(int k, int dir)
    {
        return (dir && (k == 1 || k == 2)) ? 3 - k : k;
    }

    template <typename R>
    __host__ __device__ __forceinline__ static void state(const R c[5], int dir, R q[3])
    {
        q[0] = c[0]; q[1] = dir ? c[2] : c[1]; q[2] = dir ? c[1] : c[2];
    }

    template <typename R>
    __host__ __device__ __forceinline__ static void flux(const R q[3], const R c[5], R f[3])
    {
        (void)c;
        const R h = q[0];
        const R un = div_lp(q[1], h), ut = div_lp(q[2], h);
        const R hun = h * un;
        f[0] = hun;
        f[1] = hun * un + R(0.5) * h * h;
        f[2] = hun * ut;
    }

    template <typename R>
    struct Eig {
        R u1, u2, sh, a1, a3, b;
    };

    template <typename R>
    __host__ __device__ __forceinline__ static void eig(const R qa[3], const PhysParams&, Eig<R>& e)
    {
        const R h = qa[0];
        e.u1 = div_lp(qa[1], h);
        e.u2 = div_lp(qa[2], h);
        e.sh = sqrt_lp(h);
        const R i2s = div_lp(R(1), R(2) * e.sh);
        e.a1 = (e.sh + e.u1) * i2s;
        e.a3 = (e.sh - e.u1) * i2s;
        e.b = div_lp(R(0.5), e.sh);
    }

    template <typename R>
    __host__ __device__ __forceinline__ static void left(const Eig<R>& e, const R q[3], R w[3])
    {
        w[0] = e.a1 * q[0] - e.b * q[1];
        w[1] = -e.u2 * q[0] + q[2];
        w[2] = e.a3 * q[0] + e.b * q[1];
    }

    template <typename R>
    __host__ __device__ __forceinline__ static R left_row(const Eig<R>& e, const R q[3], int k)
    {
        switch (k) {
        case 0: return e.a1 * q[0] - e.b * q[1];
        case 1: return -e.u2 * q[0] + q[2];
        default: return e.a3 * q[0] + e.b * q[1];
        }
    }

    template <typename R>
    __host__ __device__ __forceinline__ static void right(const Eig<R>& e, const R w[3], R q[3])
    {
        q[0] = w[0] + w[2];
        q[1] = (e.u1 - e.sh) * w[0] + (e.u1 + e.sh) * w[2];
        q[2] = e.u2 * w[0] + w[1] + e.u2 * w[2];
    }

    template <typename R>
    __host__ __device__ __forceinline__ static R wave_speed(const R ql[3], const R qr[3], const R cl[5], const R cr[5],
                                                            const PhysParams& pp)
    {
        (void)cl; (void)cr;
        const R hl = ql[0], hr = qr[0];
        const R unl = div_lp(ql[1], hl), utl = div_lp(ql[2], hl);
        // apps/2d/shallow_water/lib/SetWaveSpd.cpp:27-28 divides the right
        // momenta by h_l; kept when sw_compat_hl is set.
        const R hden = pp.sw_compat_hl ? hl : hr;
        const R unr = div_lp(qr[1], hden), utr = div_lp(qr[2], hden);
        const R h = R(0.5) * (hl + hr);
        const R shl = sqrt_lp(hl), shr = sqrt_lp(hr);
        const R idn = div_lp(R(1), shl + shr);
        const R un = (shl * unl + shr * unr) * idn;
        const R sh = sqrt_lp(h);
        (void)utl; (void)utr;
        const R s1 = rmin(un - sh, unl - shl);
        const R s2 = rmax(un + sh, unr + shr);
        return rmax(fabs(s1), fabs(s2));
    }
};

} // namespace finess
