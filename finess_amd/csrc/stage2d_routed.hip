// stage2d_routed.hip -- the fused stage kernel with troubled-interface
// routing (no counterpart in the reference).  Same tile, halo, update and CFL
// reduction as k_stage2d; per face direction:
//   1. every thread forms its face's 24 features from the LDS records;
//   2. each wave evaluates the MLP for its 64 faces on the matrix cores
//      (or, with route_exact, the exact weight-departure test);
//   3. the workgroup compacts troubled faces to the front of a list and the
//      others to the back, so threads 0..nt-1 run full characteristic
//      WENO5 and the rest the component-wise linear flux: at most one wave
//      runs both paths;
//   4. results land in the shared face buffer by face index.
#include <hip/hip_runtime.h>

#include "indicator.h"
#include "kernels.h"
#include "mlp.h"
#include "mlp_infer.h"
#include "stage2d_body.h"

namespace finess {

__device__ __forceinline__ int xcd_remap_r(int b, int n)
{
    constexpr int NXCD = 8;
    if (n % NXCD) return b;
    return (b % NXCD) * (n / NXCD) + b / NXCD;
}

__device__ __forceinline__ double wave_max_r(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

struct RouteSmem {
    MlpFrags fr;  // the network, staged once per workgroup
    uint16_t feat[4][64][MLP_FEAT_LD];
    float lo[4][64];
    int list[256];
    int nt[2], ns[2];  // troubled / smooth counts per face direction
};

template <class P, typename R, int WK, int TI, int TJ, bool EXACT>
__device__ __forceinline__ void routed_faces(int t, int dir, int i0, int j0, const StageArgs& a,
                                             StageSmem<P, R, TI, TJ>& sm, RouteSmem& rs, double& amax)
{
    using S = StageSmem<P, R, TI, TJ>;
    constexpr int M = P::MEQN, NV = P::NV;
    const int lane = t & 63, w = t >> 6;
    const FaceGeom g = face_geom<TI, TJ>(t, dir, i0, j0, a);
    const long long nfx = (long long)(a.mx + 1) * a.my;
    const long long mi = dir == 0 ? (long long)(i0 + g.fi) * a.my + (j0 + g.fj)
                                  : nfx + (long long)(i0 + g.fi) * (a.my + 1) + (j0 + g.fj);
    bool troubled = false;
    if (a.route_mask_read) {  // routing recorded by this step's first stage
        if (g.valid) troubled = a.route_mask[mi] != 0;
    } else if (EXACT) {
        if (g.valid) {
            R rec[6][NV];
            for (int s = 0; s < 6; s++)
                for (int v = 0; v < NV; v++) rec[s][v] = sm.cv[v][g.base + s * g.stride];
            troubled = face_departure<P, R>(rec, dir, a.pp, R(a.eps), R(a.alpha_scaling)) > R(a.route_tau);
        }
    } else {
        float f[MLP_IN];
        for (int k = 0; k < MLP_IN; k++) f[k] = 0.f;
        if (g.valid) {
            R rec[6][NV];
            for (int s = 0; s < 6; s++)
                for (int v = 0; v < NV; v++) rec[s][v] = sm.cv[v][g.base + s * g.stride];
            face_features<P, R>(rec, dir, a.pp, f);
        }
        const float z = mlp_wave_logit(f, rs.fr, rs.feat[w], rs.lo[w], lane);
        troubled = z > rs.fr.threshold;
    }
    if (a.route_mask && !a.route_mask_read && g.valid) a.route_mask[mi] = troubled ? 1 : 0;
    if (g.valid) {
        if (troubled) rs.list[atomicAdd(&rs.nt[dir], 1)] = t;
        else rs.list[255 - atomicAdd(&rs.ns[dir], 1)] = t;
    }
    __syncthreads();
    const int nt = rs.nt[dir], ns = rs.ns[dir];
    if (t == 0 && a.route_counts) {
        const int b = blockIdx.x % CFL_SPREAD;
        atomicAdd(&a.route_counts[b], (unsigned long long)nt);
        atomicAdd(&a.route_counts[CFL_SPREAD + b], (unsigned long long)(nt + ns));
    }
    int face = -1;
    bool full = true;
    if (t < nt) face = rs.list[t];
    else if (t < nt + ns) { face = rs.list[255 - (t - nt)]; full = false; }
    if (face < 0) return;
    const FaceGeom h = face_geom<TI, TJ>(face, dir, i0, j0, a);
    R fh[M];
    R alpha;
    if (full) {
        const WenoParams<R> wp{R(a.eps), R(a.power)};
        alpha = face_flux<P, R, WK, false, S::NC>(sm.cv, h.base, h.stride, dir, nullptr, 0, 0, 0, a.pp, wp,
                                                  R(a.alpha_scaling), fh);
    } else {
        alpha = face_flux_linear<P, R, S::NC>(sm.cv, h.base, h.stride, dir, a.pp, R(a.alpha_scaling), fh);
    }
    for (int m = 0; m < M; m++) sm.face[m][face] = fh[m];
    if (h.count) amax = rmax(amax, double(alpha));
    if (dir == 0 && a.smax_faces_x) a.smax_faces_x[(long long)(i0 + h.fi) * a.my + (j0 + h.fj)] = double(alpha);
    if (dir == 1 && a.smax_faces_y) a.smax_faces_y[(long long)(i0 + h.fi) * (a.my + 1) + (j0 + h.fj)] = double(alpha);
}

template <class P, typename R, int WK, int TI, int TJ, bool EXACT>
__global__ __launch_bounds__(256) void k_stage2d_routed(StageArgs a)
{
    constexpr int NT = 256;
    __shared__ StageSmem<P, R, TI, TJ> sm;
    __shared__ RouteSmem rs;
    __shared__ double red[2][NT / 64];
    const int tile = xcd_remap_r(blockIdx.x, gridDim.x);
    const int ti = tile / a.ntj, tj = tile - ti * a.ntj;
    const int i0 = ti * TI, j0 = tj * TJ;
    const int t = threadIdx.x;
    if (!EXACT) mlp_stage_frags(a.frags, &rs.fr, t, NT);
    if (t < 2) { rs.nt[t] = 0; rs.ns[t] = 0; }  // ordered before the appends by the load barrier
    double pa[P::MEQN], pb[P::MEQN];
    phase_prefetch<P, R, TI, TJ>(t, i0, j0, a, pa, pb);
    phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, sm);
    __syncthreads();
    double ax = 0.0, ay = 0.0;
    double Lx[P::MEQN];
    routed_faces<P, R, WK, TI, TJ, EXACT>(t, 0, i0, j0, a, sm, rs, ax);
    __syncthreads();
    phase_lx<P, R, TI, TJ>(t, a, sm, Lx);
    __syncthreads();
    routed_faces<P, R, WK, TI, TJ, EXACT>(t, 1, i0, j0, a, sm, rs, ay);
    __syncthreads();
    phase_update<P, R, TI, TJ>(t, i0, j0, a, sm, Lx, pa, pb);
    if (a.smax) {
        if (ax != ax) ax = __builtin_inf();
        if (ay != ay) ay = __builtin_inf();
        ax = wave_max_r(ax);
        ay = wave_max_r(ay);
        if ((t & 63) == 0) { red[0][t >> 6] = ax; red[1][t >> 6] = ay; }
        __syncthreads();
        if (t == 0) {
            double mx_ = red[0][0], my_ = red[1][0];
            for (int k = 1; k < NT / 64; k++) {
                mx_ = red[0][k] > mx_ ? red[0][k] : mx_;
                my_ = red[1][k] > my_ ? red[1][k] : my_;
            }
            const int b = blockIdx.x % CFL_SPREAD;
            atomicMax(&a.smax[b], (unsigned long long)__double_as_longlong(mx_));
            atomicMax(&a.smax[CFL_SPREAD + b], (unsigned long long)__double_as_longlong(my_));
        }
    }
}

template <typename R>
static int launch_routed(const StageArgs& a0, hipStream_t st)
{
    constexpr int TI = 7, TJ = 32;
    StageArgs a = a0;
    const int nti = (a.mx + TI - 1) / TI, ntj = (a.my + TJ - 1) / TJ;
    a.ntj = ntj;
    if (a.route_exact)
        hipLaunchKernelGGL((k_stage2d_routed<Euler2D, R, WENO_JS5_P2, TI, TJ, true>), dim3(nti * ntj), dim3(256), 0,
                           st, a);
    else
        hipLaunchKernelGGL((k_stage2d_routed<Euler2D, R, WENO_JS5_P2, TI, TJ, false>), dim3(nti * ntj), dim3(256), 0,
                           st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_stage2d_routed(int physics, int fp32, int wk, const StageArgs& a, hipStream_t st)
{
    if (physics != PHYS_EULER || wk != WENO_JS5_P2) return -1;
    return fp32 ? launch_routed<float>(a, st) : launch_routed<double>(a, st);
}

} // namespace finess
