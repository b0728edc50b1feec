// lxw_body.h -- time-integrated fluxes of the third-order Lax-Wendroff
// (Picard integral) step, lib/2d/ConstructIntegratedR.cpp:59-296:
//
//   F = f + dt/2 (f_t + dt/3 f_tt),   G = g + dt/2 (g_t + dt/3 g_tt)
//   S = f_x + g_y,   f_t = -A1 S,   g_t = -A2 S
//   q_tt = H1(q_x, S) + H2(q_y, S) + A1 (f_xx + g_xy) + A2 (f_xy + g_yy)
//   f_tt = A1 q_tt + H1(S, S),   g_tt = A2 q_tt + H2(S, S)
//
// with A_d the flux Jacobians and H_d(a, b) their second derivatives applied
// to two vectors.  The reference forms A and H as explicit tensors from the
// app's DFluxFunc / D2FluxFunc and contracts them; here every product is a
// directional derivative of the flux itself, evaluated exactly with
// hyper-dual numbers (x + a e1 + b e2 + ab e1e2, e1^2 = e2^2 = 0): one flux
// evaluation at q + a e1 + b e2 yields A a (e1 part), A b (e2 part) and
// H(a, b) (e1e2 part).  Derivatives: 5-point central differences Diff1/Diff2
// (lib/WenoReconstruct.cpp:702-716, 791-806) and the 4-point cross
// derivative of ConstructIntegratedR.cpp:202-213.
#pragma once
#include "physics.h"
#include "stage2d_body.h"

namespace finess {

template <typename T>
struct HD {
    T v, a, b, ab;
};
template <typename T>
__host__ __device__ __forceinline__ HD<T> operator+(HD<T> x, HD<T> y) { return {x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab}; }
template <typename T>
__host__ __device__ __forceinline__ HD<T> operator-(HD<T> x, HD<T> y) { return {x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab}; }
template <typename T>
__host__ __device__ __forceinline__ HD<T> operator*(HD<T> x, HD<T> y)
{
    return {x.v * y.v, x.v * y.a + x.a * y.v, x.v * y.b + x.b * y.v, x.v * y.ab + x.a * y.b + x.b * y.a + x.ab * y.v};
}
template <typename T>
__host__ __device__ __forceinline__ HD<T> operator*(T s, HD<T> x) { return {s * x.v, s * x.a, s * x.b, s * x.ab}; }
template <typename T>
__host__ __device__ __forceinline__ HD<T> hinv(HD<T> x)
{
    const T i = T(1) / x.v, i2 = i * i;
    return {i, -x.a * i2, -x.b * i2, (T(2) * x.a * x.b * i - x.ab) * i2};
}

// Flux of each physics in any number type (same expressions as FluxFunc).
template <typename T>
__host__ __device__ __forceinline__ void euler_flux(const T q[5], T gm1, T F[5], T G[5])
{
    const T rho = q[0];
    const T ir = hinv(rho);
    const T u1 = q[1] * ir, u2 = q[2] * ir, u3 = q[3] * ir;
    const T e = q[4];
    const T press = gm1 * (e - typename T::scalar(0.5) * (rho * (u1 * u1 + u2 * u2 + u3 * u3)));
    const T ru1 = rho * u1, ru2 = rho * u2, ep = e + press;
    F[0] = ru1; F[1] = ru1 * u1 + press; F[2] = ru1 * u2; F[3] = ru1 * u3; F[4] = u1 * ep;
    G[0] = ru2; G[1] = ru2 * u1; G[2] = ru2 * u2 + press; G[3] = ru2 * u3; G[4] = u2 * ep;
}

// Shallow water, g = 1 (apps/2d/shallow_water/lib/FluxFunc.cpp:8-35).
template <typename T>
__host__ __device__ __forceinline__ void sw_flux(const T q[3], T F[3], T G[3])
{
    const T h = q[0];
    const T ih = hinv(h);
    const T u1 = q[1] * ih, u2 = q[2] * ih;
    const T hh = typename T::scalar(0.5) * (h * h);
    F[0] = h * u1; F[1] = h * u1 * u1 + hh; F[2] = h * u1 * u2;
    G[0] = h * u2; G[1] = h * u2 * u1; G[2] = h * u2 * u2 + hh;
}

// Hyper-dual number of scalar type T (double for the plain form, float for
// the FP32 stencil form).
template <typename T>
struct HDs : HD<T> {
    using scalar = T;
    __host__ __device__ HDs() {}
    __host__ __device__ HDs(const HD<T>& x) : HD<T>(x) {}
    __host__ __device__ HDs(T v_, T a_, T b_, T ab_) : HD<T>{v_, a_, b_, ab_} {}
};
template <typename T>
__host__ __device__ __forceinline__ HDs<T> operator+(HDs<T> x, HDs<T> y) { return HDs<T>(HD<T>(x) + HD<T>(y)); }
template <typename T>
__host__ __device__ __forceinline__ HDs<T> operator-(HDs<T> x, HDs<T> y) { return HDs<T>(HD<T>(x) - HD<T>(y)); }
template <typename T>
__host__ __device__ __forceinline__ HDs<T> operator*(HDs<T> x, HDs<T> y) { return HDs<T>(HD<T>(x) * HD<T>(y)); }
template <typename T>
__host__ __device__ __forceinline__ HDs<T> operator*(T s, HDs<T> x) { return HDs<T>(s * HD<T>(x)); }
template <typename T>
__host__ __device__ __forceinline__ HDs<T> hinv(HDs<T> x) { return HDs<T>(hinv(HD<T>(x))); }
using HDd = HDs<double>;

template <class P>
struct LxwPhys;

template <>
struct LxwPhys<Euler2D> {
    static constexpr int M = 5;
    template <typename R>
    __host__ __device__ __forceinline__ static void flux(const R q[5], const PhysParams& pp, R F[5], R G[5])
    {
        const R rho = q[0];
        const R u1 = div_lp(q[1], rho), u2 = div_lp(q[2], rho), u3 = div_lp(q[3], rho);
        const R e = q[4];
        const R press = R(pp.gamma - 1.0) * (e - R(0.5) * rho * (u1 * u1 + u2 * u2 + u3 * u3));
        F[0] = rho * u1; F[1] = rho * u1 * u1 + press; F[2] = rho * u1 * u2; F[3] = rho * u1 * u3; F[4] = u1 * (e + press);
        G[0] = rho * u2; G[1] = rho * u2 * u1; G[2] = rho * u2 * u2 + press; G[3] = rho * u2 * u3; G[4] = u2 * (e + press);
    }
    // hyper-dual flux at q + a e1 + b e2
    template <typename R>
    __host__ __device__ __forceinline__ static void flux_hd(const R q[5], const R a[5], const R b[5],
                                                            const PhysParams& pp, HDs<R> F[5], HDs<R> G[5])
    {
        HDs<R> x[5];
        for (int m = 0; m < 5; m++) x[m] = HDs<R>(q[m], a[m], b[m], R(0));
        euler_flux<HDs<R>>(x, HDs<R>(R(pp.gamma - 1.0), R(0), R(0), R(0)), F, G);
    }
};

// Shallow water: DFluxFunc / D2FluxFunc (apps/2d/shallow_water/lib/
// DFluxFunc.cpp, D2FluxFunc.cpp) are the exact derivatives of the flux, so
// the hyper-dual evaluation reproduces them.
template <>
struct LxwPhys<ShallowWater2D> {
    static constexpr int M = 3;
    template <typename R>
    __host__ __device__ __forceinline__ static void flux(const R q[3], const PhysParams&, R F[3], R G[3])
    {
        const R h = q[0];
        const R u1 = div_lp(q[1], h), u2 = div_lp(q[2], h);
        F[0] = h * u1; F[1] = h * u1 * u1 + R(0.5) * h * h; F[2] = h * u1 * u2;
        G[0] = h * u2; G[1] = h * u2 * u1; G[2] = h * u2 * u2 + R(0.5) * h * h;
    }
    template <typename R>
    __host__ __device__ __forceinline__ static void flux_hd(const R q[3], const R a[3], const R b[3],
                                                            const PhysParams&, HDs<R> F[3], HDs<R> G[3])
    {
        HDs<R> x[3];
        for (int m = 0; m < 3; m++) x[m] = HDs<R>(q[m], a[m], b[m], R(0));
        sw_flux<HDs<R>>(x, F, G);
    }
};

struct LxwArgs {
    const double* q;     // state planes (mx*my), boundary-mapped on load
    void* F;             // out: planes (mx+6)*(my+6), extended index (i+3)*(my+6)+(j+3),
    void* G;             //      of R (the kernel's arithmetic: FP64, or FP32 for the FP32 form)
    int mx, my;
    long long plane;
    int bc[4];
    int time_order;
    double dt, dx, dy;
    PhysParams pp;
    int src_ghosts;      // >0: q planes include that many ghosts per side (drop-in mode)
};

// R: the arithmetic of the differences and Taylor terms (double in the plain
// form, float in the FP32 stencil form, whose stage reads these fluxes as FP32).
template <int TI, int TJ, int M, typename R = double>
struct LxwSmem {
    static constexpr int HI = TI + 4, HJ = TJ + 4, NC = HI * HJ;
    R q[M][NC];
    R f[M][NC];
    R g[M][NC];
};

template <class P>
__host__ __device__ __forceinline__ void lxw_load_state(const LxwArgs& a, int gi, int gj, double q[P::MEQN])
{
    if (a.src_ghosts) {
        const int g = a.src_ghosts;
        const long long ld = a.my + 2 * g;
        const long long pl = (long long)(a.mx + 2 * g) * ld;
        const long long idx = (long long)(gi + g) * ld + (gj + g);
        for (int m = 0; m < P::MEQN; m++) q[m] = a.q[m * pl + idx];
        return;
    }
    int fi, fj;
    const int si = bc_map(a.bc[0], a.bc[1], gi, a.mx, fi);
    const int sj = bc_map(a.bc[2], a.bc[3], gj, a.my, fj);
    const long long idx = (long long)si * a.my + sj;
    for (int m = 0; m < P::MEQN; m++) q[m] = a.q[m * a.plane + idx];
    if (fi) q[P::NORMAL_X] = -q[P::NORMAL_X];
    if (fj) q[P::NORMAL_Y] = -q[P::NORMAL_Y];
}

// Phase 1: states and point fluxes of the tile plus a 2-cell halo (corners
// included: the cross derivatives read the diagonal neighbours).
template <class P, int TI, int TJ, typename R = double>
__host__ __device__ __forceinline__ void lxw_phase_load(int tid, int nthreads, int e0i, int e0j, const LxwArgs& a,
                                                        LxwSmem<TI, TJ, P::MEQN, R>& sm)
{
    using S = LxwSmem<TI, TJ, P::MEQN, R>;
    for (int c = tid; c < S::NC; c += nthreads) {
        const int ci = c / S::HJ, cj = c - ci * S::HJ;
        const int gi = e0i + ci - 2 - 3, gj = e0j + cj - 2 - 3;
        if (gi > a.mx + 4 || gj > a.my + 4) continue;
        double qd[P::MEQN];
        R q[P::MEQN], f[P::MEQN], g[P::MEQN];
        lxw_load_state<P>(a, gi, gj, qd);
        for (int m = 0; m < P::MEQN; m++) q[m] = R(qd[m]);
        LxwPhys<P>::flux(q, a.pp, f, g);
        for (int m = 0; m < P::MEQN; m++) { sm.q[m][c] = q[m]; sm.f[m][c] = f[m]; sm.g[m][c] = g[m]; }
    }
}

template <typename R>
__host__ __device__ __forceinline__ R d1(R h, R f0, R f1, R f3, R f4)
{
    R t = (f0 - f4) * R(1.0 / 12.0);
    t += (-f1 + f3) * R(2.0 / 3.0);
    return div_lp(t, h);
}
template <typename R>
__host__ __device__ __forceinline__ R d2(R h, R f0, R f1, R f2, R f3, R f4)
{
    R t = (-f0 - f4) * R(1.0 / 12.0);
    t += (f1 + f3) * R(4.0 / 3.0);
    t += (-f2) * R(5.0 / 2.0);
    return div_lp(t, h * h);
}

// Phase 2: one output cell per thread.
template <class P, int TI, int TJ, typename R = double>
__host__ __device__ __forceinline__ void lxw_phase_cell(int tid, int e0i, int e0j, const LxwArgs& a,
                                                        const LxwSmem<TI, TJ, P::MEQN, R>& sm)
{
    using S = LxwSmem<TI, TJ, P::MEQN, R>;
    constexpr int M = P::MEQN;
    const int ci = tid / TJ, cj = tid - ci * TJ;
    const int ei = e0i + ci, ej = e0j + cj;              // extended index, gi = ei - 3
    if (ci >= TI || ei >= a.mx + 6 || ej >= a.my + 6) return;
    const int c = (ci + 2) * S::HJ + (cj + 2);
    const int X = S::HJ, Y = 1;
    const R dx = R(a.dx), dy = R(a.dy), dt = R(a.dt);
    R q[M], fx[M], gy[M], qx[M], qy[M], S_[M];
    for (int m = 0; m < M; m++) {
        q[m] = sm.q[m][c];
        fx[m] = d1(dx, sm.f[m][c - 2 * X], sm.f[m][c - X], sm.f[m][c + X], sm.f[m][c + 2 * X]);
        qx[m] = d1(dx, sm.q[m][c - 2 * X], sm.q[m][c - X], sm.q[m][c + X], sm.q[m][c + 2 * X]);
        gy[m] = d1(dy, sm.g[m][c - 2 * Y], sm.g[m][c - Y], sm.g[m][c + Y], sm.g[m][c + 2 * Y]);
        qy[m] = d1(dy, sm.q[m][c - 2 * Y], sm.q[m][c - Y], sm.q[m][c + Y], sm.q[m][c + 2 * Y]);
        S_[m] = fx[m] + gy[m];
    }
    HDs<R> F1[M], G1[M];
    LxwPhys<P>::flux_hd(q, qx, S_, a.pp, F1, G1);      // A S (e2), H1(qx, S) (e1e2)
    R f_t[M], g_t[M], f_tt[M], g_tt[M];
    for (int m = 0; m < M; m++) { f_t[m] = -F1[m].b; g_t[m] = -G1[m].b; f_tt[m] = R(0); g_tt[m] = R(0); }
    if (a.time_order > 2) {
        HDs<R> F2[M], G2[M], F3[M], G3[M];
        LxwPhys<P>::flux_hd(q, qy, S_, a.pp, F2, G2);  // H2(qy, S)
        R sx[M], sy[M];
        for (int m = 0; m < M; m++) {
            const R fxx = d2(dx, sm.f[m][c - 2 * X], sm.f[m][c - X], sm.f[m][c], sm.f[m][c + X], sm.f[m][c + 2 * X]);
            const R gyy = d2(dy, sm.g[m][c - 2 * Y], sm.g[m][c - Y], sm.g[m][c], sm.g[m][c + Y], sm.g[m][c + 2 * Y]);
            R t1 = div_lp(R(0.5) * (sm.f[m][c + X + Y] - sm.f[m][c - X + Y]), dx);
            t1 -= div_lp(R(0.5) * (sm.f[m][c + X - Y] - sm.f[m][c - X - Y]), dx);
            t1 *= div_lp(R(0.5), dy);
            R t2 = div_lp(R(0.5) * (sm.g[m][c + X + Y] - sm.g[m][c - X + Y]), dx);
            t2 -= div_lp(R(0.5) * (sm.g[m][c + X - Y] - sm.g[m][c - X - Y]), dx);
            t2 *= div_lp(R(0.5), dy);
            sx[m] = fxx + t2;  // f_xx + g_xy
            sy[m] = t1 + gyy;  // f_xy + g_yy
        }
        LxwPhys<P>::flux_hd(q, sx, sy, a.pp, F3, G3);  // A1 sx (e1 of F), A2 sy (e2 of G)
        R qtt[M];
        for (int m = 0; m < M; m++) qtt[m] = F1[m].ab + G2[m].ab + F3[m].a + G3[m].b;
        HDs<R> F4[M], G4[M], F5[M], G5[M];
        LxwPhys<P>::flux_hd(q, qtt, S_, a.pp, F4, G4); // A q_tt (e1)
        LxwPhys<P>::flux_hd(q, S_, S_, a.pp, F5, G5);  // H(S, S) (e1e2)
        for (int m = 0; m < M; m++) {
            f_tt[m] = F4[m].a + F5[m].ab;
            g_tt[m] = G4[m].a + G5[m].ab;
        }
    }
    const long long lld = a.my + 6, lpl = (long long)(a.mx + 6) * lld;
    const long long o = (long long)ei * lld + ej;
    for (int m = 0; m < M; m++) {
        static_cast<R*>(a.F)[m * lpl + o] = sm.f[m][c] + R(0.5) * dt * (f_t[m] + div_lp(dt, R(3)) * f_tt[m]);
        static_cast<R*>(a.G)[m * lpl + o] = sm.g[m][c] + R(0.5) * dt * (g_t[m] + div_lp(dt, R(3)) * g_tt[m]);
    }
}

} // namespace finess
