// stage2d_wide.hip -- the 2D Runge-Kutta stage and ConstructL for the WENO
// orders other than five: Jiang-Shu 7, 9, 11 (lib/WenoReconstruct.cpp:
// 150-375), WENO-Z 7 and 9 (:440-572), linear weights 7 and 9 (:618-700) and
// the first-order value (:702-713), in FP64.
//
// The stencil of an interface grows to 2r cells (r = (order + 1) / 2, up to
// 12) and the fused fifth-order kernel's register budget does not stretch
// that far, so these orders run in two steps: a face kernel per direction
// (one thread per interface of a 16 x 16 tile, the tile's stencil cells
// staged in LDS in FP64) forms the interface flux of lib/2d/ConstructL.cpp:
// 63-208 (x faces) / :215-358 (y faces) into a face array, and an update
// kernel takes the flux differences and does the stage update.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "stage2d_body.h"

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
    static __device__ __forceinline__ double c(int k, int j)
    {
        constexpr double t[4][4] = {{-1. / 4., 13. / 12., -23. / 12., 25. / 12.},
                                    {1. / 12., -5. / 12., 13. / 12., 1. / 4.},
                                    {-1. / 12., 7. / 12., 7. / 12., -1. / 12.},
                                    {1. / 4., 13. / 12., -5. / 12., 1. / 12.}};
        return t[k][j];
    }
    static __device__ __forceinline__ double d(int k)
    {
        constexpr double t[4] = {1. / 35., 12. / 35., 18. / 35., 4. / 35.};
        return t[k];
    }
    static __device__ __forceinline__ double bh(int k, int e)
    {
        constexpr double t[2][10] = {{547, -3882, 4642, -1854, 7043, -17246, 7042, 11003, -9402, 2107},
                                     {267, -1642, 1602, -494, 2843, -5966, 1922, 3443, -2522, 547}};
        return t[k][e];
    }
};

template <>
struct WideTab<5> {
    static constexpr int NH = 3;
    static __device__ __forceinline__ double c(int k, int j)
    {
        constexpr double t[5][5] = {{1. / 5., -21. / 20., 137. / 60., -163. / 60., 137. / 60.},
                                    {-1. / 20., 17. / 60., -43. / 60., 77. / 60., 1. / 5.},
                                    {1. / 30., -13. / 60., 47. / 60., 9. / 20., -1. / 20.},
                                    {-1. / 20., 9. / 20., 47. / 60., -13. / 60., 1. / 30.},
                                    {1. / 5., 77. / 60., -43. / 60., 17. / 60., -1. / 20.}};
        return t[k][j];
    }
    static __device__ __forceinline__ double d(int k)
    {
        constexpr double t[5] = {1. / 126., 10. / 63., 10. / 21., 20. / 63., 5. / 126.};
        return t[k];
    }
    static __device__ __forceinline__ double bh(int k, int e)
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
    static __device__ __forceinline__ double c(int k, int j)
    {
        constexpr double t[6][6] = {{-1. / 6., 31. / 30., -163. / 60., 79. / 20., -71. / 20., 49. / 20.},
                                    {1. / 30., -13. / 60., 37. / 60., -21. / 20., 29. / 20., 1. / 6.},
                                    {-1. / 60., 7. / 60., -23. / 60., 19. / 20., 11. / 30., -1. / 30.},
                                    {1. / 60., -2. / 15., 37. / 60., 37. / 60., -2. / 15., 1. / 60.},
                                    {-1. / 30., 11. / 30., 19. / 20., -23. / 60., 7. / 60., -1. / 60.},
                                    {1. / 6., 29. / 20., -21. / 20., 37. / 60., -13. / 60., 1. / 30.}};
        return t[k][j];
    }
    static __device__ __forceinline__ double d(int k)
    {
        constexpr double t[6] = {1. / 462., 5. / 77., 25. / 77., 100. / 231., 25. / 154., 1. / 77.};
        return t[k];
    }
    static __device__ __forceinline__ double bh(int k, int e)
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
__device__ __forceinline__ double wide_beta(const double* g, int k)
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
__device__ __forceinline__ double wide_cand(const double* g, int k)
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
__device__ __forceinline__ double weno_wide(const double* g, double eps, double pw)
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

// A workgroup takes a tile of WT x WT faces of one direction and stages
// every cell their stencils reach in LDS once: the cell record and the flux
// of the state rotated to the face normal, in FP64 (lib/2d/ConstructL.cpp:
// 88-137 samples q and f over the stencil the same way, per face).
constexpr int WT = 16;

template <class P, int R, int DIR>
struct WideSmem {
    static constexpr int HI = DIR == 0 ? WT + 2 * R - 1 : WT;  // cells along i
    static constexpr int HJ = DIR == 0 ? WT : WT + 2 * R - 1;  // cells along j
    static constexpr int NC = HI * HJ;
    double rec[P::NV][NC];
    double f[P::MEQN][NC];
};

// Interface flux of the face whose stencil cell s sits at LDS index
// base + s * stride (cells i-R..i+R-1 along the normal); returns alpha.  The
// characteristic fields are reconstructed one per pass (a loop the compiler
// keeps), each pass projecting the stencil with one row of the left
// eigenvectors, so a thread holds the 2 (2R - 1) split values of one field
// at a time.  R = 1 (first order): the split parts are the upwind values.
template <class P, int R, int VER, int DIR, class S>
__device__ __forceinline__ double wide_face_flux(const S& sm, int base, int stride, const StageArgs& a,
                                                 double fh[P::MEQN])
{
    constexpr int M = P::MEQN, NV = P::NV, WS = 2 * R - 1, NS = 2 * R;
    double cl[NV], cr[NV], ql[M], qr[M];
    for (int v = 0; v < NV; v++) {
        cl[v] = sm.rec[v][base + (R - 1) * stride];
        cr[v] = sm.rec[v][base + R * stride];
    }
    P::state(cl, DIR, ql);
    P::state(cr, DIR, qr);
    double qa[M];
    for (int m = 0; m < M; m++) qa[m] = 0.5 * (qr[m] + ql[m]);
    typename P::template Eig<double> e;
    P::eig(qa, a.pp, e);
    const double ws = P::wave_speed(ql, qr, cl, cr, a.pp);
    const double amin = a.galpha ? a.galpha[DIR] : 0.0;  // global_alpha
    const double alpha = amin > ws ? amin : ws;
    const double la = a.alpha_scaling * alpha;
    double ghat[M];
#pragma unroll 1
    for (int m = 0; m < M; m++) {
        double gp[WS], gm[WS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int c = base + s * stride;
            double rec[NV], q[M], f[M];
            for (int v = 0; v < M; v++) rec[v] = sm.rec[v][c];  // state() reads the conserved part only
            P::state(rec, DIR, q);
            for (int k = 0; k < M; k++) f[k] = sm.f[k][c];
            const double w = P::left_row(e, q, m), gw = P::left_row(e, f, m);
            // lib/2d/ConstructL.cpp:179-185: gp from points 1..ws, gm from
            // points ws+1..2 (reversed)
            if (s < WS) gp[s] = 0.5 * (gw + la * w);
            if (s >= 1) gm[WS - s] = 0.5 * (gw - la * w);
            // keep each cell's LDS reads in their iteration: hoisting all of
            // them ahead of the arithmetic costs 10 registers per cell
            asm volatile("" ::: "memory");
        }
        double gh;
        if constexpr (R == 1) gh = gp[0] + gm[0];
        else gh = weno_wide<R, VER>(gp, a.eps, a.power) + weno_wide<R, VER>(gm, a.eps, a.power);
#pragma unroll
        for (int k = 0; k < M; k++)
            if (m == k) ghat[k] = gh;
    }
    double fr_[M];
    P::right(e, ghat, fr_);
    for (int k = 0; k < M; k++) fh[P::rot_q(k, DIR)] = fr_[k];
    return alpha;
}

// Tiles of faces along each axis for direction DIR.
__host__ __device__ __forceinline__ int wide_tiles_j(int dir, int my) { return (my + (dir == 1) + WT - 1) / WT; }
__host__ __device__ __forceinline__ int wide_tiles_i(int dir, int mx) { return (mx + (dir == 0) + WT - 1) / WT; }

// One workgroup per WT x WT tile of faces of direction DIR: the face flux
// into a.face_x / a.face_y ([m][nf], x face (i, j) at i*my + j, y face at
// i*(my+1) + j), the face speed into smax_faces_* when set, and the
// workgroup maximum over the counted faces into
// a.smax[DIR * CFL_SPREAD + block % CFL_SPREAD].
template <class P, int R, int VER, int DIR>
__global__ __launch_bounds__(256) void k_wide_faces(StageArgs a)
{
    if (stage_halted(a)) return;
    constexpr int M = P::MEQN, NV = P::NV;
    using S = WideSmem<P, R, DIR>;
    __shared__ S sm;
    const int ntj = wide_tiles_j(DIR, a.my);
    const int ti = blockIdx.x / ntj, tj = blockIdx.x - ti * ntj;
    const int i0 = ti * WT, j0 = tj * WT;
    const int t = threadIdx.x;
    // LDS cell (ci, cj) is cell (i0 + ci - R, j0 + cj) for x faces and
    // (i0 + ci, j0 + cj - R) for y faces; cells beyond the last face's
    // stencil (index n + R - 1 along the normal) are not loaded
    for (int c = t; c < S::NC; c += 256) {
        const int ci = c / S::HJ, cj = c - ci * S::HJ;
        const int gi = i0 + ci - (DIR == 0 ? R : 0), gj = j0 + cj - (DIR == 1 ? R : 0);
        const bool ok = DIR == 0 ? (gi <= a.mx + R - 1 && gj < a.my) : (gi < a.mx && gj <= a.my + R - 1);
        if (!ok) continue;
        double q[M], rec[NV], qr[M], f[M];
        load_state<P>(a, gi, gj, q);
        P::cell_record(q, a.pp, rec);
        P::state(rec, DIR, qr);
        P::flux(qr, rec, f);
        for (int v = 0; v < NV; v++) sm.rec[v][c] = rec[v];
        for (int k = 0; k < M; k++) sm.f[k][c] = f[k];
    }
    __syncthreads();
    const int fi = t / WT, fj = t - fi * WT;
    const int i = i0 + fi, j = j0 + fj;
    const bool valid = DIR == 0 ? (i <= a.mx && j < a.my) : (i < a.mx && j <= a.my);
    double amax = 0.0;
    if (valid) {
        const int nj = DIR == 0 ? a.my : a.my + 1;
        const long long nf = DIR == 0 ? (long long)(a.mx + 1) * a.my : (long long)a.mx * (a.my + 1);
        const long long f = (long long)i * nj + j;
        double fh[M];
        const double alpha = wide_face_flux<P, R, VER, DIR>(sm, fi * S::HJ + fj, DIR == 0 ? S::HJ : 1, a, fh);
        double* dst = DIR == 0 ? a.face_x : a.face_y;
        for (int m = 0; m < M; m++) dst[m * nf + f] = fh[m];
        double* sf = DIR == 0 ? a.smax_faces_x : a.smax_faces_y;
        if (sf) sf[f] = alpha;
        const bool count = DIR == 0 ? i < a.mx : j < a.my;
        if (count) amax = alpha;
        if (amax != amax) amax = __builtin_inf();  // a NaN speed must not vanish in the max
    }
    if (!a.smax) return;
    __shared__ double red[4];
    amax = wave_max64(amax);
    const int lane = t & 63, w = t >> 6;
    if (lane == 0) red[w] = amax;
    __syncthreads();
    if (t == 0) {
        double v = red[0];
        for (int k = 1; k < 4; k++) v = red[k] > v ? red[k] : v;
        atomicMax(&a.smax[DIR * CFL_SPREAD + blockIdx.x % CFL_SPREAD], (unsigned long long)__double_as_longlong(v));
    }
}

// L = -(F(i+1) - F(i))/dx - (G(j+1) - G(j))/dy (lib/2d/ConstructL.cpp:
// 388-399) and the stage update out = a1 qa + a2 qb + beta dt L.
template <int M>
__global__ __launch_bounds__(256) void k_wide_update(StageArgs a)
{
    if (stage_halted(a)) return;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= a.plane) return;
    const int i = (int)(c / a.my), j = (int)(c - (long long)i * a.my);
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long fx0 = (long long)i * a.my + j, fy0 = (long long)i * (a.my + 1) + j;
    const double bdt = a.out ? stage_bdt(a) : 0.0;
    for (int m = 0; m < M; m++) {
        const double Lx = -(a.face_x[m * nfx + fx0 + a.my] - a.face_x[m * nfx + fx0]) * a.idx;
        const double L = Lx - (a.face_y[m * nfy + fy0 + 1] - a.face_y[m * nfy + fy0]) * a.idy;
        if (a.Lout) a.Lout[m * a.plane + c] = L;
        if (!a.out) continue;
        double v = a.a1 * a.qa[m * a.plane + c];
        if (a.use_b) v = v + a.a2 * a.qb[m * a.plane + c];
        a.out[m * a.plane + c] = v + bdt * L;
    }
}

template <class P, int R, int VER>
static int launch_wide_t(const StageArgs& a, hipStream_t st)
{
    const unsigned bx = (unsigned)wide_tiles_i(0, a.mx) * wide_tiles_j(0, a.my);
    const unsigned by = (unsigned)wide_tiles_i(1, a.mx) * wide_tiles_j(1, a.my);
    hipLaunchKernelGGL((k_wide_faces<P, R, VER, 0>), dim3(bx), dim3(256), 0, st, a);
    hipLaunchKernelGGL((k_wide_faces<P, R, VER, 1>), dim3(by), dim3(256), 0, st, a);
    hipLaunchKernelGGL((k_wide_update<P::MEQN>), dim3((unsigned)((a.plane + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <class P>
static int dispatch_wide(int wk, const StageArgs& a, hipStream_t st)
{
    switch (wk) {
    case WENO_JS7: return launch_wide_t<P, 4, WV_JS>(a, st);
    case WENO_JS9: return launch_wide_t<P, 5, WV_JS>(a, st);
    case WENO_JS11: return launch_wide_t<P, 6, WV_JS>(a, st);
    case WENO_Z7: return launch_wide_t<P, 4, WV_Z>(a, st);
    case WENO_Z9: return launch_wide_t<P, 5, WV_Z>(a, st);
    case WENO_FD7: return launch_wide_t<P, 4, WV_FD>(a, st);
    case WENO_FD9: return launch_wide_t<P, 5, WV_FD>(a, st);
    case WENO_O1: return launch_wide_t<P, 1, WV_FD>(a, st);
    default: return -1;
    }
}

// The face arrays travel in a.face_x / a.face_y.
int launch_stage2d_wide(int physics, int wk, const StageArgs& a, hipStream_t st)
{
    if (!a.face_x || !a.face_y || a.lxw_F) return -1;
    if (physics == PHYS_EULER) return dispatch_wide<Euler2D>(wk, a, st);
    if (physics == PHYS_SHALLOW_WATER) return dispatch_wide<ShallowWater2D>(wk, a, st);
    return -1;
}

} // namespace finess
