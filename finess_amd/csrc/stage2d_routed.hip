// stage2d_routed.hip -- the fused stage kernel with troubled-interface
// routing (no counterpart in the reference).  Same tile, halo, stage update
// and CFL reduction as k_stage2d; the face work of both directions is done
// in one pass:
//   1. every face of the tile (x faces then y faces, 487 for a 7 x 32 tile)
//      gets a routing decision: the MLP on the matrix cores (each wave
//      evaluates its 64 x faces, then its y faces), the exact weight-departure
//      test, or the decision recorded by the first stage of the time step;
//   2. the workgroup compacts the troubled faces of both directions to the
//      front of one list and the smooth ones to the back, so the threads run
//      full characteristic WENO5 on list entries 0..nt-1 and the
//      component-wise linear flux on the rest (each thread takes two
//      entries): at most one wave per tile runs both paths;
//   3. the fluxes land in per-direction face buffers, and the update forms
//      L from both.
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

template <class P, typename R, int TI, int TJ>
struct RouteTile {
    static constexpr int NFX = (TI + 1) * TJ, NFY = TI * (TJ + 1), NF = NFX + NFY;
    StageSmem<P, R, TI, TJ> s;  // records; s.face holds the x-face fluxes
    R fy[P::MEQN][NFY];         // y-face fluxes
    int list[NF];               // face ids: x face f -> f, y face f -> NFX + f
    int nt, ns;                 // troubled / smooth faces of the tile
    double red[2][4];
};

struct MlpTile {
    MlpFrags fr;  // the network, staged once per workgroup
    uint16_t feat[4][64][MLP_FEAT_LD];
    float lo[4][64];
};

// Index of a face in the per-step routing record.
__device__ __forceinline__ long long mask_index(int dir, const FaceGeom& g, int i0, int j0, const StageArgs& a)
{
    const long long nfx = (long long)(a.mx + 1) * a.my;
    return dir == 0 ? (long long)(i0 + g.fi) * a.my + (j0 + g.fj)
                    : nfx + (long long)(i0 + g.fi) * (a.my + 1) + (j0 + g.fj);
}

template <class P, typename R, int TI, int TJ>
__device__ __forceinline__ void stencil_records(const StageSmem<P, R, TI, TJ>& s, const FaceGeom& g,
                                                R rec[6][P::NV])
{
    for (int k = 0; k < 6; k++)
        for (int v = 0; v < P::NV; v++) rec[k][v] = s.cv[v][g.base + k * g.stride];
}

// EXACT: route by the weight-departure label itself; EVAL: decide (and
// record when a.route_mask is set), else read this step's record.
template <class P, typename R, int WK, int TI, int TJ, bool EXACT, bool EVAL>
__global__ __launch_bounds__(256, (sizeof(R) == 4 && !EVAL) ? 4 : 1) void k_stage2d_routed(StageArgs a)
{
    constexpr int NT = 256, M = P::MEQN, NV = P::NV;
    using T = RouteTile<P, R, TI, TJ>;
    using S = StageSmem<P, R, TI, TJ>;
    static_assert(T::NFX <= NT && T::NFY <= NT, "one face per thread and direction");
    __shared__ T sm;
    const int tile = xcd_remap_r(blockIdx.x, gridDim.x);
    const int ti = tile / a.ntj, tj = tile - ti * a.ntj;
    const int i0 = ti * TI, j0 = tj * TJ;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    const FaceGeom gx = face_geom<TI, TJ>(t, 0, i0, j0, a);
    const FaceGeom gy = face_geom<TI, TJ>(t, 1, i0, j0, a);
    bool tx = false, ty = false;
    if (!EVAL) {  // this step's routing, read before the tile so the latencies overlap
        if (gx.valid) tx = a.route_mask[mask_index(0, gx, i0, j0, a)] != 0;
        if (gy.valid) ty = a.route_mask[mask_index(1, gy, i0, j0, a)] != 0;
    }
    double pa[M], pb[M];
    phase_prefetch<P, R, TI, TJ>(t, i0, j0, a, pa, pb);
    if (t == 0) { sm.nt = 0; sm.ns = 0; }
    if constexpr (EVAL && !EXACT) {
        __shared__ MlpTile ml;
        mlp_stage_frags(a.frags, &ml.fr, t, NT);
        phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, sm.s);
        __syncthreads();
        for (int dir = 0; dir < 2; dir++) {
            const FaceGeom& g = dir ? gy : gx;
            float f[MLP_IN];
            for (int k = 0; k < MLP_IN; k++) f[k] = 0.f;
            if (g.valid) {
                R rec[6][NV];
                stencil_records<P, R, TI, TJ>(sm.s, g, rec);
                face_features<P, R>(rec, dir, a.pp, f);
            }
            const bool tr = mlp_wave_logit(f, ml.fr, ml.feat[w], ml.lo[w], lane) > ml.fr.threshold;
            if (dir) ty = tr; else tx = tr;
        }
    } else {
        phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, sm.s);
        __syncthreads();
        if (EVAL) {  // EXACT
            if (gx.valid) {
                R rec[6][NV];
                stencil_records<P, R, TI, TJ>(sm.s, gx, rec);
                tx = face_departure<P, R>(rec, 0, a.pp, R(a.eps), R(a.alpha_scaling)) > R(a.route_tau);
            }
            if (gy.valid) {
                R rec[6][NV];
                stencil_records<P, R, TI, TJ>(sm.s, gy, rec);
                ty = face_departure<P, R>(rec, 1, a.pp, R(a.eps), R(a.alpha_scaling)) > R(a.route_tau);
            }
        }
    }
    if (EVAL && a.route_mask) {
        if (gx.valid) a.route_mask[mask_index(0, gx, i0, j0, a)] = tx ? 1 : 0;
        if (gy.valid) a.route_mask[mask_index(1, gy, i0, j0, a)] = ty ? 1 : 0;
    }
    // compaction: troubled faces to the front of the list, smooth to the back
    if (gx.valid) {
        if (tx) sm.list[atomicAdd(&sm.nt, 1)] = t;
        else sm.list[T::NF - 1 - atomicAdd(&sm.ns, 1)] = t;
    }
    if (gy.valid) {
        if (ty) sm.list[atomicAdd(&sm.nt, 1)] = T::NFX + t;
        else sm.list[T::NF - 1 - atomicAdd(&sm.ns, 1)] = T::NFX + t;
    }
    __syncthreads();
    const int nt = sm.nt, n = nt + sm.ns;
    if (t == 0 && a.route_counts) {
        const int b = blockIdx.x % CFL_SPREAD;
        atomicAdd(&a.route_counts[b], (unsigned long long)nt);
        atomicAdd(&a.route_counts[CFL_SPREAD + b], (unsigned long long)n);
    }
    double ax = 0.0, ay = 0.0;
    const WenoParams<R> wp{R(a.eps), R(a.power)};
    for (int k = 0; k < 2; k++) {
        const int e = t + k * NT;
        if (e >= n) break;
        const bool full = e < nt;
        const int id = full ? sm.list[e] : sm.list[T::NF - 1 - (e - nt)];
        const int dir = id >= T::NFX ? 1 : 0;
        const int f = dir ? id - T::NFX : id;
        const FaceGeom h = face_geom<TI, TJ>(f, dir, i0, j0, a);
        R fh[M];
        R alpha;
        if (full)
            alpha = face_flux<P, R, WK, false, S::NC>(sm.s.cv, h.base, h.stride, dir, nullptr, 0, 0, 0, a.pp, wp,
                                                      R(a.alpha_scaling), fh);
        else
            alpha = face_flux_linear<P, R, S::NC>(sm.s.cv, h.base, h.stride, dir, a.pp, R(a.alpha_scaling), fh);
        if (dir) {
            for (int m = 0; m < M; m++) sm.fy[m][f] = fh[m];
            if (h.count) ay = rmax(ay, double(alpha));
            if (a.smax_faces_y) a.smax_faces_y[(long long)(i0 + h.fi) * (a.my + 1) + (j0 + h.fj)] = double(alpha);
        } else {
            for (int m = 0; m < M; m++) sm.s.face[m][f] = fh[m];
            if (h.count) ax = rmax(ax, double(alpha));
            if (a.smax_faces_x) a.smax_faces_x[(long long)(i0 + h.fi) * a.my + (j0 + h.fj)] = double(alpha);
        }
    }
    __syncthreads();
    // L = -(Fhat(i+1) - Fhat(i))/dx - (Ghat(j+1) - Ghat(j))/dy and the stage
    // update, in the operation order of the plain kernel
    {
        const int ci = t / TJ, cj = t - ci * TJ;
        const int gi = i0 + ci, gj = j0 + cj;
        if (ci < TI && gi < a.mx && gj < a.my) {
            const long long idx = (long long)gi * a.my + gj;
            const int c = (ci + 3) * S::HJ + (cj + 3);
            const bool exact_lds = sizeof(R) == sizeof(double) && !a.src_ghosts;
            const double idx_ = 1.0 / a.dx, idy_ = 1.0 / a.dy;
            for (int m = 0; m < M; m++) {
                const double Lx = -(double(sm.s.face[m][(ci + 1) * TJ + cj]) - double(sm.s.face[m][ci * TJ + cj])) * idx_;
                const double L = Lx - (double(sm.fy[m][ci * (TJ + 1) + cj + 1]) - double(sm.fy[m][ci * (TJ + 1) + cj])) * idy_;
                if (a.Lout) a.Lout[m * a.plane + idx] = L;
                if (!a.out) continue;
                const double qa = (exact_lds && a.a_is_src) ? double(sm.s.cv[m][c]) : pa[m];
                double v = a.a1 * qa;
                if (a.use_b) {
                    const double qb = (exact_lds && a.b_is_src) ? double(sm.s.cv[m][c]) : pb[m];
                    v = v + a.a2 * qb;
                }
                a.out[m * a.plane + idx] = v + a.bdt * L;
            }
        }
    }
    if (a.smax) {
        if (ax != ax) ax = __builtin_inf();
        if (ay != ay) ay = __builtin_inf();
        ax = wave_max_r(ax);
        ay = wave_max_r(ay);
        if (lane == 0) { sm.red[0][w] = ax; sm.red[1][w] = ay; }
        __syncthreads();
        if (t == 0) {
            double mx_ = sm.red[0][0], my_ = sm.red[1][0];
            for (int k = 1; k < NT / 64; k++) {
                mx_ = sm.red[0][k] > mx_ ? sm.red[0][k] : mx_;
                my_ = sm.red[1][k] > my_ ? sm.red[1][k] : my_;
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
    const dim3 grid(nti * ntj), block(256);
    if (a.route_mask_read)
        hipLaunchKernelGGL((k_stage2d_routed<Euler2D, R, WENO_JS5_P2, TI, TJ, false, false>), grid, block, 0, st, a);
    else if (a.route_exact)
        hipLaunchKernelGGL((k_stage2d_routed<Euler2D, R, WENO_JS5_P2, TI, TJ, true, true>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL((k_stage2d_routed<Euler2D, R, WENO_JS5_P2, TI, TJ, false, true>), grid, block, 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_stage2d_routed(int physics, int fp32, int wk, const StageArgs& a, hipStream_t st)
{
    if (physics != PHYS_EULER || wk != WENO_JS5_P2) return -1;
    if (a.route_mask_read && !a.route_mask) return -1;
    return fp32 ? launch_routed<float>(a, st) : launch_routed<double>(a, st);
}

} // namespace finess
