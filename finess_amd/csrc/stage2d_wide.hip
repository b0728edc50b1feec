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
#include "weno_wide.h"

namespace finess {

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
// skip3 (5-field systems; set when every cell of the tile has rho w = 0):
// field 3's flux, the mean state's w and so its characteristic values are
// zero, and so is its reconstruction, which is not computed.  Only from
// stencil radius SKIP3_MIN_R on: for the seventh order the branch in the
// field loop made the face kernel slower (1,185 against 1,282
// Mcell-updates/s); the ninth and eleventh orders gain 14% and 15% (1,004
// against 881, 761 against 661, 1024^2 Euler).
#ifndef FINESS_SKIP3_MIN_R
#define FINESS_SKIP3_MIN_R 5
#endif
constexpr int SKIP3_MIN_R = FINESS_SKIP3_MIN_R;
template <class P, int R, int VER, int DIR, class S>
__device__ __forceinline__ double wide_face_flux(const S& sm, int base, int stride, const StageArgs& a,
                                                 double fh[P::MEQN], bool skip3)
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
        if (M == 5 && R >= SKIP3_MIN_R && m == 3 && skip3) {
            ghat[3] = 0.0;
            continue;
        }
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
    // neighbouring tiles on one XCD: the halo lines they share stay in its L2
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int ti = tile / ntj, tj = tile - ti * ntj;
    const int i0 = ti * WT, j0 = tj * WT;
    const int t = threadIdx.x;
    // LDS cell (ci, cj) is cell (i0 + ci - R, j0 + cj) for x faces and
    // (i0 + ci, j0 + cj - R) for y faces; cells beyond the last face's
    // stencil (index n + R - 1 along the normal) are not loaded
    bool nz3 = false;  // a loaded cell with rho w != 0
    for (int c = t; c < S::NC; c += 256) {
        const int ci = c / S::HJ, cj = c - ci * S::HJ;
        const int gi = i0 + ci - (DIR == 0 ? R : 0), gj = j0 + cj - (DIR == 1 ? R : 0);
        const bool ok = DIR == 0 ? (gi <= a.mx + R - 1 && gj < a.my) : (gi < a.mx && gj <= a.my + R - 1);
        if (!ok) continue;
        double q[M], rec[NV], qr[M], f[M];
        load_state<P>(a, gi, gj, q);
        if constexpr (M == 5 && R >= SKIP3_MIN_R) nz3 = nz3 || q[3] != 0.0;
        P::cell_record(q, a.pp, rec);
        P::state(rec, DIR, qr);
        P::flux(qr, rec, f);
        for (int v = 0; v < NV; v++) sm.rec[v][c] = rec[v];
        for (int k = 0; k < M; k++) sm.f[k][c] = f[k];
    }
    bool skip3 = false;
    if constexpr (M == 5 && R >= SKIP3_MIN_R) skip3 = !__syncthreads_or(nz3);
    else __syncthreads();
    const int fi = t / WT, fj = t - fi * WT;
    const int i = i0 + fi, j = j0 + fj;
    const bool valid = DIR == 0 ? (i <= a.mx && j < a.my) : (i < a.mx && j <= a.my);
    double amax = 0.0;
    if (valid) {
        const int nj = DIR == 0 ? a.my : a.my + 1;
        const long long nf = DIR == 0 ? (long long)(a.mx + 1) * a.my : (long long)a.mx * (a.my + 1);
        const long long f = (long long)i * nj + j;
        double fh[M];
        const double alpha = wide_face_flux<P, R, VER, DIR>(sm, fi * S::HJ + fj, DIR == 0 ? S::HJ : 1, a, fh, skip3);
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
template <class P>
__global__ __launch_bounds__(256) void k_wide_update(StageArgs a)
{
    constexpr int M = P::MEQN;
    if (stage_halted(a)) return;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= a.plane) return;
    const int i = (int)(c / a.my), j = (int)(c - (long long)i * a.my);
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long fx0 = (long long)i * a.my + j, fy0 = (long long)i * (a.my + 1) + j;
    const double bdt = a.out ? stage_bdt(a) : 0.0;
    double psi[M];
    if (a.src_kind != SRC_OFF) {
        double qs[M];
        load_state<P>(a, i, j, qs);
        source_psi<M>(a.src_kind, a.src_coef, qs, a.aux ? a.aux + c : nullptr, a.plane, psi);
    }
    for (int m = 0; m < M; m++) {
        const double Lx = -(a.face_x[m * nfx + fx0 + a.my] - a.face_x[m * nfx + fx0]) * a.idx;
        double L = Lx - (a.face_y[m * nfy + fy0 + 1] - a.face_y[m * nfy + fy0]) * a.idy;
        if (a.src_kind != SRC_OFF) L = psi[m] + L;  // Lstar = psi + tmp (lib/2d/ConstructL.cpp:378-386)
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
    hipLaunchKernelGGL((k_wide_update<P>), dim3((unsigned)((a.plane + 255) / 256)), dim3(256), 0, st, a);
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
