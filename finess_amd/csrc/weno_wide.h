// weno_wide.h -- the WENO reconstructions of orders 7, 9 and 11: Jiang-Shu
// (lib/WenoReconstruct.cpp:150-375), WENO-Z 7 and 9 (:440-572) and the
// linear weights 7 and 9 (:618-700), for the 2D face-tile kernels
// (stage2d_wide.hip) and the 1D stage (stage1d_body.h).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "weno.h"

namespace finess {

enum WideVersion : int { WV_JS = 0, WV_Z = 1, WV_FD = 2 };

// Index of entry (a, b), a <= b, in a row-major upper triangle of size r.
__host__ __device__ constexpr int tri_index(int r, int a, int b) { return a * r - a * (a - 1) / 2 + (b - a); }

// Candidate values u_k = sum_j c[k][j] g[k+j], smoothness indicators
// beta_k = sum_{a<=b} B_k[a][b] g[k+a] g[k+b] (scaled as the reference lists
// them: 240, 5040, 120960 times the Jiang-Shu integral) and linear weights
// d_k.  Stencils k <= (r-1)/2 are listed; stencil r-1-k is their mirror
// image, B_{r-1-k}[a][b] = B_k[r-1-b][r-1-a].
template <int R>
struct WideTab;

template <>
struct WideTab<4> {
    static constexpr int NH = 2;
    static __host__ __device__ __forceinline__ double c(int k, int j)
    {
        constexpr double t[4][4] = {{-1. / 4., 13. / 12., -23. / 12., 25. / 12.},
                                    {1. / 12., -5. / 12., 13. / 12., 1. / 4.},
                                    {-1. / 12., 7. / 12., 7. / 12., -1. / 12.},
                                    {1. / 4., 13. / 12., -5. / 12., 1. / 12.}};
        return t[k][j];
    }
    static __host__ __device__ __forceinline__ double d(int k)
    {
        constexpr double t[4] = {1. / 35., 12. / 35., 18. / 35., 4. / 35.};
        return t[k];
    }
    static __host__ __device__ __forceinline__ double bh(int k, int e)
    {
        constexpr double t[2][10] = {{547, -3882, 4642, -1854, 7043, -17246, 7042, 11003, -9402, 2107},
                                     {267, -1642, 1602, -494, 2843, -5966, 1922, 3443, -2522, 547}};
        return t[k][e];
    }
};

template <>
struct WideTab<5> {
    static constexpr int NH = 3;
    static __host__ __device__ __forceinline__ double c(int k, int j)
    {
        constexpr double t[5][5] = {{1. / 5., -21. / 20., 137. / 60., -163. / 60., 137. / 60.},
                                    {-1. / 20., 17. / 60., -43. / 60., 77. / 60., 1. / 5.},
                                    {1. / 30., -13. / 60., 47. / 60., 9. / 20., -1. / 20.},
                                    {-1. / 20., 9. / 20., 47. / 60., -13. / 60., 1. / 30.},
                                    {1. / 5., 77. / 60., -43. / 60., 17. / 60., -1. / 20.}};
        return t[k][j];
    }
    static __host__ __device__ __forceinline__ double d(int k)
    {
        constexpr double t[5] = {1. / 126., 10. / 63., 10. / 21., 20. / 63., 5. / 126.};
        return t[k];
    }
    static __host__ __device__ __forceinline__ double bh(int k, int e)
    {
        constexpr double t[3][15] = {
            {22658, -208501, 364863, -288007, 86329, 482963, -1704396, 1358458, -411487, 1521393, -2462076, 758823,
             1020563, -649501, 107918},
            {6908, -60871, 99213, -70237, 18079, 138563, -464976, 337018, -88297, 406293, -611976, 165153, 242723,
             -140251, 22658},
            {6908, -51001, 67923, -38947, 8209, 104963, -299076, 179098, -38947, 231153, -299076, 67923, 104963,
             -51001, 6908}};
        return t[k][e];
    }
};

template <>
struct WideTab<6> {
    static constexpr int NH = 3;
    static __host__ __device__ __forceinline__ double c(int k, int j)
    {
        constexpr double t[6][6] = {{-1. / 6., 31. / 30., -163. / 60., 79. / 20., -71. / 20., 49. / 20.},
                                    {1. / 30., -13. / 60., 37. / 60., -21. / 20., 29. / 20., 1. / 6.},
                                    {-1. / 60., 7. / 60., -23. / 60., 19. / 20., 11. / 30., -1. / 30.},
                                    {1. / 60., -2. / 15., 37. / 60., 37. / 60., -2. / 15., 1. / 60.},
                                    {-1. / 30., 11. / 30., 19. / 20., -23. / 60., 7. / 60., -1. / 60.},
                                    {1. / 6., 29. / 20., -21. / 20., 37. / 60., -13. / 60., 1. / 30.}};
        return t[k][j];
    }
    static __host__ __device__ __forceinline__ double d(int k)
    {
        constexpr double t[6] = {1. / 462., 5. / 77., 25. / 77., 100. / 231., 25. / 154., 1. / 77.};
        return t[k];
    }
    static __host__ __device__ __forceinline__ double bh(int k, int e)
    {
        constexpr double t[3][21] = {
            {1152561, -12950184, 29442256, -33918804, 19834350, -4712740, 36480687, -166461044, 192596472,
             -113206788, 27060170, 190757572, -444003904, 262901672, -63394124, 260445372, -311771244, 76206736,
             94851237, -47460464, 6150211},
            {271779, -3015728, 6694608, -7408908, 4067018, -880548, 8449957, -37913324, 42405032, -23510468,
             5134574, 43093692, -97838784, 55053752, -12183636, 56662212, -65224244, 14742480, 19365967, -9117992,
             1152561},
            {139633, -1429976, 2863984, -2792660, 1325006, -245620, 3824847, -15880404, 15929912, -7727988, 1458762,
             17195652, -35817664, 17905032, -3462252, 19510972, -20427884, 4086352, 5653317, -2380800, 271779}};
        return t[k][e];
    }
};

// Smoothness indicator of stencil k (indices are compile-time constants
// after unrolling, so every coefficient is an immediate).
template <int R>
__host__ __device__ __forceinline__ double wide_beta(const double* g, int k)
{
    using T = WideTab<R>;
    double b = 0.0;
#pragma unroll
    for (int a = 0; a < R; a++) {
        double row = 0.0;
#pragma unroll
        for (int c = a; c < R; c++) {
            const double coef = k < T::NH ? T::bh(k, tri_index(R, a, c))
                                          : T::bh(R - 1 - k, tri_index(R, R - 1 - c, R - 1 - a));
            row = fma(coef, g[k + c], row);
        }
        b = fma(g[k + a], row, b);
    }
    return b;
}

template <int R>
__host__ __device__ __forceinline__ double wide_cand(const double* g, int k)
{
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < R; j++) s = fma(WideTab<R>::c(k, j), g[k + j], s);
    return s;
}

// Reconstruction at x_{i+1/2} from g[0..2R-2] (g[R-1] = u_i).  JS and the
// linear weights run one candidate at a time (nothing but the running sums
// stays live); Z needs every indicator for tau first.
template <int R, int VER>
__host__ __device__ __forceinline__ double weno_wide(const double* g, double eps, double pw)
{
    using T = WideTab<R>;
    double num = 0.0, den = 0.0;
    if (VER == WV_FD) {
#pragma unroll
        for (int k = 0; k < R; k++) num = fma(T::d(k), wide_cand<R>(g, k), num);
        return num;
    }
    if (VER == WV_JS) {
#pragma unroll
        for (int k = 0; k < R; k++) {
            const double a = eps + wide_beta<R>(g, k);
            const double omt = T::d(k) * (pw == 2.0 ? 1.0 / (a * a) : pow(a, -pw));
            num = fma(omt, wide_cand<R>(g, k), num);
            den += omt;
        }
        return num / den;
    }
    double beta[R];
#pragma unroll
    for (int k = 0; k < R; k++) beta[k] = wide_beta<R>(g, k);
    // lib/WenoReconstruct.cpp:470 (order 7) and :549 (order 9)
    const double tau = R == 4 ? fabs(beta[0] + 3. * (beta[1] - beta[2]) - beta[3]) : fabs(beta[R - 1] - beta[0]);
#pragma unroll
    for (int k = 0; k < R; k++) {
        const double z = tau / (eps + beta[k]);
        const double omt = T::d(k) * (1.0 + (pw == 2.0 ? z * z : pow(z, pw)));
        num = fma(omt, wide_cand<R>(g, k), num);
        den += omt;
    }
    return num / den;
}

// The (radius, version) of a wide WenoKind (weno.h); radius 1 is the
// first-order value.
__host__ __device__ constexpr int wide_radius(int wk)
{
    return wk == WENO_O1 ? 1 : (wk == WENO_JS7 || wk == WENO_Z7 || wk == WENO_FD7) ? 4 : (wk == WENO_JS11) ? 6 : 5;
}
__host__ __device__ constexpr int wide_version(int wk)
{
    return (wk == WENO_Z7 || wk == WENO_Z9) ? WV_Z : (wk == WENO_FD7 || wk == WENO_FD9 || wk == WENO_O1) ? WV_FD : WV_JS;
}

} // namespace finess
