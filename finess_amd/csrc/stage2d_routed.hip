// stage2d_routed.hip -- the stage kernels of troubled-interface routing (no
// counterpart in the reference).  A face is routed to full characteristic
// WENO5 (troubled) or to the component-wise linear FD5 flux (smooth).
//
// The first stage of a step (k_stage2d_routed) decides: the MLP on the matrix
// cores for every face of the tile (or the exact weight-departure test), then
// a workgroup compaction puts the troubled faces of both directions at the
// front of one list so that at most one wave per tile runs both paths; it
// records the decisions as a compact list of troubled faces.
//
// The later stages reuse the record in two launches: k_route_full computes
// the characteristic fluxes of the listed faces only (every lane busy, no
// tile), and k_stage2d_lookup runs the tile with the linear flux at smooth
// faces and the precomputed flux at troubled ones.  Without the
// characteristic path the lookup kernel needs far fewer registers, so more
// workgroups stay resident to hide the memory latency.
#include <hip/hip_runtime.h>

#include "indicator.h"
#include "kernels.h"
#include "mlp.h"
#include "mlp_infer.h"
#include "stage2d_body.h"
#include "step_control_body.h"

namespace finess {

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

template <class P, typename R, int TI, int TJ>
__device__ __forceinline__ void stencil_records(const StageSmem<P, R, TI, TJ>& s, const FaceGeom& g,
                                                R rec[6][P::NV])
{
    for (int k = 0; k < 6; k++)
        for (int v = 0; v < P::NV; v++) rec[k][v] = s.cv[v][g.base + k * g.stride];
}

// Owner of a face shared by two tiles: the tile for which it is not the
// far edge (the domain's last face belongs to the only tile that has it).
__device__ __forceinline__ bool face_owned(int dir, const FaceGeom& g, int i0, int j0, int TI, int TJ,
                                           const StageArgs& a)
{
    return dir == 0 ? (g.fi < TI || i0 + g.fi == a.mx) : (g.fj < TJ || j0 + g.fj == a.my);
}

// Stores a troubled face in list slot `slot` of the step's record: the list
// entry is the face's cell-index pair (gi, gj) packed in 16 bits each (the
// host limits routed grids to 65535 cells per side).
__device__ __forceinline__ void record_troubled(int gi, int gj, long long slot, const StageArgs& a)
{
    a.route_list[slot] = (int)(((unsigned)gi << 16) | (unsigned)gj);
}

// The list slot of a face from a tile record: thread tt of the tile, -1 when
// the face is smooth.  The slot counts the troubled faces of lower waves and
// lanes of the same direction.
__device__ __forceinline__ int record_slot(const unsigned long long* mask, int base, int tt)
{
    const int w = tt >> 6, lane = tt & 63;
    const unsigned long long m = mask[w];
    if (!((m >> lane) & 1ull)) return -1;
    int s = base + __popcll(m & ((1ull << lane) - 1ull));
    for (int k = 0; k < w; k++) s += __popcll(mask[k]);
    return s;
}

// Loads the routing records of a tile and of its next tiles in i and j
// (rt[0], rt[1], rt[2]; zero where there is none) into LDS: threads below
// 3 x 10 read one 8-byte word each before the tile load (record_fetch) and
// store it after the load (record_store), ahead of the caller's barrier.
template <int TI>
__device__ __forceinline__ unsigned long long record_fetch(const StageArgs& a, int t, int tile, int ti, int tj)
{
    constexpr int RW = sizeof(RouteTileRec) / sizeof(unsigned long long);
    if (t >= 3 * RW) return 0ull;
    const int rk = t / RW, rwi = t - rk * RW;
    const int nti = (a.mx + TI - 1) / TI;
    const int other = rk == 0 ? tile : rk == 1 ? (ti + 1 < nti ? tile + a.ntj : -1) : (tj + 1 < a.ntj ? tile + 1 : -1);
    return other >= 0 ? reinterpret_cast<const unsigned long long*>(a.route_tile + other)[rwi] : 0ull;
}
__device__ __forceinline__ void record_store(RouteTileRec* rt, int t, unsigned long long rword)
{
    constexpr int RW = sizeof(RouteTileRec) / sizeof(unsigned long long);
    if (t < 3 * RW) reinterpret_cast<unsigned long long*>(&rt[t / RW])[t % RW] = rword;
}

// EXACT: route by the weight-departure label itself, else by the MLP; the
// decisions hold for this stage only (a record is made by k_classify).
// The learned FP32 kernel is held to 3 waves per SIMD (168 registers
// including the MFMA accumulators; its LDS allows 3 workgroups per CU).
template <class P, typename R, int WK, int TI, int TJ, bool EXACT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EXACT ? 1 : (sizeof(R) == 4 ? 3 : 2)))) void k_stage2d_routed(StageArgs a)
{
    constexpr int NT = 256, M = P::MEQN, NV = P::NV;
    using T = RouteTile<P, R, TI, TJ>;
    using S = StageSmem<P, R, TI, TJ>;
    static_assert(T::NFX <= NT && T::NFY <= NT, "one face per thread and direction");
    __shared__ T sm;
    if (stage_halted(a)) return;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int ti = tile / a.ntj, tj = tile - ti * a.ntj;
    const int i0 = ti * TI, j0 = tj * TJ;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    const FaceGeom gx = face_geom<TI, TJ>(t, 0, i0, j0, a);
    const FaceGeom gy = face_geom<TI, TJ>(t, 1, i0, j0, a);
    bool tx = false, ty = false;
    double pa[M], pb[M];
    phase_prefetch<P, R, TI, TJ>(t, i0, j0, a, pa, pb);
    if (t == 0) { sm.nt = 0; sm.ns = 0; }
    if constexpr (!EXACT) {
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
            const double idx_ = a.idx, idy_ = a.idy;
            const double bdt = stage_bdt(a);
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
                a.out[m * a.plane + idx] = v + bdt * L;
            }
        }
    }
    if (a.smax) {
        if (ax != ax) ax = __builtin_inf();
        if (ay != ay) ay = __builtin_inf();
        ax = wave_max64(ax);
        ay = wave_max64(ay);
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

// The indicator alone, for the first stage of a new routing record: every
// face of the tile is classified (by the network, or by the exact
// weight-departure label with EXACT) and the decisions are recorded; the
// stage itself then runs from the record as k_route_full + k_stage2d_lookup.
// Without the fluxes and the update the kernel keeps only the tile records
// and the network, so more workgroups stay resident than in the combined
// kernel.
template <class P, typename R, int TI, int TJ, bool EXACT>
__global__ __launch_bounds__(256) void k_classify(StageArgs a)
{
    constexpr int NT = 256, NV = P::NV;
    using S = StageSmem<P, R, TI, TJ>;
    __shared__ S ss;
    __shared__ int cnt[4][4];  // per wave: recorded troubled x, y; troubled faces, faces
    __shared__ unsigned long long bal[2][4];
    __shared__ int rbase[2];
    if (stage_halted(a)) return;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int ti = tile / a.ntj, tj = tile - ti * a.ntj;
    const int i0 = ti * TI, j0 = tj * TJ;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const FaceGeom gx = face_geom<TI, TJ>(t, 0, i0, j0, a);
    const FaceGeom gy = face_geom<TI, TJ>(t, 1, i0, j0, a);
    bool tx = false, ty = false;
    if constexpr (!EXACT) {
        __shared__ MlpTile ml;
        mlp_stage_frags(a.frags, &ml.fr, t, NT);
        phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, ss);
        __syncthreads();
        for (int dir = 0; dir < 2; dir++) {
            const FaceGeom& g = dir ? gy : gx;
            float f[MLP_IN];
            for (int k = 0; k < MLP_IN; k++) f[k] = 0.f;
            if (g.valid) {
                R rec[6][NV];
                stencil_records<P, R, TI, TJ>(ss, g, rec);
                face_features<P, R>(rec, dir, a.pp, f);
            }
            const bool tr = mlp_wave_logit(f, ml.fr, ml.feat[w], ml.lo[w], lane) > ml.fr.threshold;
            if (dir) ty = tr; else tx = tr;
        }
    } else {
        phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, ss);
        __syncthreads();
        if (gx.valid) {
            R rec[6][NV];
            stencil_records<P, R, TI, TJ>(ss, gx, rec);
            tx = face_departure<P, R>(rec, 0, a.pp, R(a.eps), R(a.alpha_scaling)) > R(a.route_tau);
        }
        if (gy.valid) {
            R rec[6][NV];
            stencil_records<P, R, TI, TJ>(ss, gy, rec);
            ty = face_departure<P, R>(rec, 1, a.pp, R(a.eps), R(a.alpha_scaling)) > R(a.route_tau);
        }
    }
    tx = tx && gx.valid;
    ty = ty && gy.valid;
    // the record: the troubled owned faces as one ballot per wave and
    // direction; their list slots from one global atomic per tile and
    // direction (x faces fill the list from slot 0, y faces from (mx+1) my)
    const bool ox = gx.valid && face_owned(0, gx, i0, j0, TI, TJ, a);
    const bool oy = gy.valid && face_owned(1, gy, i0, j0, TI, TJ, a);
    const unsigned long long bx = __ballot(ox && tx), by = __ballot(oy && ty);
    const int nt = __popcll(__ballot(tx)) + __popcll(__ballot(ty));
    const int nf = __popcll(__ballot(gx.valid)) + __popcll(__ballot(gy.valid));
    if (lane == 0) {
        cnt[0][w] = __popcll(bx);
        cnt[1][w] = __popcll(by);
        cnt[2][w] = nt;
        cnt[3][w] = nf;
        bal[0][w] = bx;
        bal[1][w] = by;
    }
    __syncthreads();
    if (t < 2) {  // thread d takes the slots of direction d
        if (a.route_n_clear && blockIdx.x == 0) a.route_n_clear[t] = 0ull;
        int c = 0;
        for (int k = 0; k < NT / 64; k++) c += cnt[t][k];
        rbase[t] = c ? (int)atomicAdd(&a.route_n[t], (unsigned long long)c) : 0;
    } else if (t < 4 && a.route_counts) {
        int c = 0;
        for (int k = 0; k < NT / 64; k++) c += cnt[t][k];
        atomicAdd(&a.route_counts[(t - 2) * CFL_SPREAD + blockIdx.x % CFL_SPREAD], (unsigned long long)c);
    }
    __syncthreads();
    RouteTileRec* rec = a.route_tile + tile;
    if (t < 8) rec->mask[t >> 2][t & 3] = bal[t >> 2][t & 3];
    else if (t < 10) rec->base[t - 8] = rbase[t - 8];
    if (bx | by) {
        long long bsx = rbase[0], bsy = (long long)(a.mx + 1) * a.my + rbase[1];
        for (int k = 0; k < w; k++) { bsx += cnt[0][k]; bsy += cnt[1][k]; }
        const unsigned long long below = (1ull << lane) - 1ull;
        if (ox && tx) record_troubled(i0 + gx.fi, j0 + gx.fj, bsx + __popcll(bx & below), a);
        if (oy && ty) record_troubled(i0 + gy.fi, j0 + gy.fj, bsy + __popcll(by & below), a);
    }
}

// Characteristic WENO5 flux of the troubled face of direction DIR in list
// slot `slot`, whose list entry is `id`.  The stencil of a face whose six
// cells all lie inside the grid is read directly, others through the
// boundary map; the records are formed as in the tile load.
template <class P, typename R, int WK, int DIR>
__device__ __forceinline__ void route_face(const StageArgs& a, unsigned id, long long slot)
{
    constexpr int M = P::MEQN, NV = P::NV;
    const int gi = (int)(id >> 16), gj = (int)(id & 0xffffu);
    R cv[NV][6];
    bool nz3 = false;  // a stencil cell with rho w != 0 (face_part's skip3)
    const bool inner = !a.src_ghosts && (DIR == 0 ? (gi >= 3 && gi + 2 < a.mx) : (gj >= 3 && gj + 2 < a.my));
    const int stride = DIR == 0 ? a.my : 1;
    const int c0 = DIR == 0 ? (gi - 3) * a.my + gj : gi * a.my + (gj - 3);
    // one record formation for both kinds of stencil (only the loads
    // differ), so that a face's flux has the same bits whichever way its
    // cells were read
    for (int s = 0; s < 6; s++) {
        double q[M], rec[NV];
        if (inner) {
            for (int m = 0; m < M; m++) q[m] = (a.src + m * a.plane)[c0 + s * stride];
        } else if (DIR == 0) {
            load_state<P>(a, gi - 3 + s, gj, q);
        } else {
            load_state<P>(a, gi, gj - 3 + s, q);
        }
        if constexpr (M == 5) nz3 = nz3 || q[3] != 0.0;
        P::template cell_record<sizeof(R) == 4>(q, a.pp, rec);
        for (int v = 0; v < NV; v++) cv[v][s] = R(rec[v]);
    }
    const WenoParams<R> wp{R(a.eps), R(a.power)};
    R fh[M];
    const R alpha = nz3 ? face_flux<P, R, WK, false, 6>(cv, 0, 1, DIR, nullptr, 0, 0, 0, a.pp, wp, R(a.alpha_scaling), fh)
                        : face_flux<P, R, WK, false, 6, true>(cv, 0, 1, DIR, nullptr, 0, 0, 0, a.pp, wp,
                                                              R(a.alpha_scaling), fh);
    R* out = static_cast<R*>(a.route_flux) + 6 * slot;
    for (int m = 0; m < M; m++) out[m] = fh[m];
    out[M] = alpha;
}

// workgroups of the troubled-face pass (a grid-stride loop over the list)
#ifndef FINESS_ROUTE_FULL_BLOCKS
#define FINESS_ROUTE_FULL_BLOCKS 1024
#endif
constexpr int ROUTE_FULL_BLOCKS = FINESS_ROUTE_FULL_BLOCKS;
static_assert(ROUTE_FULL_BLOCKS % 2 == 0, "half the workgroups per direction");

// The troubled faces of one direction: lane g0 of `step` lanes takes list
// entries g0, g0 + step, ... below n of the direction's list (slots from
// off); id0 is the lane's first entry, read ahead by the caller.  The loop
// runs whole waves (lanes past the end idle).
template <class P, typename R, int WK, int DIR>
__device__ __forceinline__ void route_dir(const StageArgs& a, long long g0, long long step, long long n, long long off,
                                          unsigned id0)
{
    const long long lane = g0 & 63;
    for (long long g = g0; g - lane < n; g += step)
        if (g < n) route_face<P, R, WK, DIR>(a, g == g0 ? id0 : (unsigned)a.route_list[off + g], off + g);
}

// Characteristic WENO5 fluxes of the listed troubled faces of this stage's
// source state, one face per lane: the x faces on the first half of the
// workgroups, the y faces on the second, so that every wave runs one
// direction (a compile-time constant in route_face) and every lane knows its
// first list slot before the list lengths arrive: the halt word, the lengths
// and that slot's entry are read together instead of one after another.
//
// With a.ctl_run the first workgroup runs the previous step's controller
// instead (the pass does not read dt; the stage launch after it does): its
// latency hides behind the pass's, and the step saves a launch.  Workgroups
// that read the halt word before the controller set it compute fluxes no
// later launch reads: the stage kernels after it return at once.
template <class P, typename R, int WK>
__global__ __launch_bounds__(256) void k_route_full(StageArgs a)
{
    int b = (int)blockIdx.x, nb = (int)gridDim.x;
    if (a.ctl_run) {
        if (b == 0) {
            if (threadIdx.x < 64) step_control_wave(a.ctl_run, a.ctl_smax, a.ctl_p, (int)threadIdx.x);
            return;
        }
        b -= 1;
        nb -= 1;
    }
    const int half = nb / 2;
    const int dir = b >= half;
    const long long g0 = (long long)(b - dir * half) * blockDim.x + threadIdx.x;
    const long long step = (long long)half * blockDim.x;
    const long long nfx = (long long)(a.mx + 1) * a.my, nfy = (long long)a.mx * (a.my + 1);
    const long long off = dir ? nfx : 0;
    // the entry is read whether or not the slot is in use (a stale entry of
    // an unused slot is never used; the list holds nfx + nfy slots)
    const unsigned id0 = g0 < (dir ? nfy : nfx) ? (unsigned)a.route_list[off + g0] : 0u;
    const long long n = (long long)a.route_n[dir];
    if (stage_halted(a) | (n <= 0)) return;
    if (dir) route_dir<P, R, WK, 1>(a, g0, step, n, off, id0);
    else route_dir<P, R, WK, 0>(a, g0, step, n, off, id0);
}

// One stage with the step's recorded routing: the linear flux at smooth
// faces, the flux k_route_full computed at troubled ones.
template <class P, typename R, int TI, int TJ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 5 : 3))) void k_stage2d_lookup(StageArgs a)
{
    constexpr int NT = 256, M = P::MEQN;
    using S = StageSmem<P, R, TI, TJ>;
    static_assert((TI + 1) * TJ <= NT && TI * (TJ + 1) <= NT, "one face per thread and direction");
    __shared__ S sm;
    __shared__ R fy[M][TI * (TJ + 1)];
    __shared__ double red[2][NT / 64];
    // the routing records of this tile (0), of the next tile in i (1, whose
    // first x faces are this tile's far x faces) and in j (2, likewise for y)
    __shared__ RouteTileRec rt[3];
    if (stage_halted(a)) return;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int ti = tile / a.ntj, tj = tile - ti * a.ntj;
    const int i0 = ti * TI, j0 = tj * TJ;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const FaceGeom gx = face_geom<TI, TJ>(t, 0, i0, j0, a);
    const FaceGeom gy = face_geom<TI, TJ>(t, 1, i0, j0, a);
    // the step's routing records and the update operands, read before the
    // tile load
    const unsigned long long rword = record_fetch<TI>(a, t, tile, ti, tj);
    double pa[M], pb[M];
    phase_prefetch<P, R, TI, TJ>(t, i0, j0, a, pa, pb);
    phase_load<P, R, TI, TJ>(t, NT, i0, j0, a, sm);
    record_store(rt, t, rword);
    __syncthreads();
    // list slots of the thread's faces (-1: smooth): an owned face from this
    // tile's record, a far-edge face from the next tile's, where it is face
    // fj (x) or fi (TJ + 1) (y) of that tile
    int sx = -1, sy = -1;
    if (gx.valid) {
        sx = face_owned(0, gx, i0, j0, TI, TJ, a) ? record_slot(rt[0].mask[0], rt[0].base[0], t)
                                                   : record_slot(rt[1].mask[0], rt[1].base[0], gx.fj);
    }
    if (gy.valid) {
        sy = face_owned(1, gy, i0, j0, TI, TJ, a) ? record_slot(rt[0].mask[1], rt[0].base[1], t)
                                                   : record_slot(rt[2].mask[1], rt[2].base[1], gy.fi * (TJ + 1));
        if (sy >= 0) sy += (a.mx + 1) * a.my;
    }
    // wave-speed maxima in the stencil precision (the maximum of values
    // that are all R is exact in R)
    R ax = R(0), ay = R(0);
    for (int dir = 0; dir < 2; dir++) {
        const FaceGeom& h = dir ? gy : gx;
        const int slot = dir ? sy : sx;
        if (!h.valid) continue;
        R fh[M];
        R alpha;
        if (slot >= 0) {
            const R* in = static_cast<const R*>(a.route_flux) + 6 * (long long)slot;
            for (int m = 0; m < M; m++) fh[m] = in[m];
            alpha = in[M];
        } else {
            alpha = face_flux_linear<P, R, S::NC>(sm.cv, h.base, h.stride, dir, a.pp, R(a.alpha_scaling), fh);
        }
        if (dir) {
            for (int m = 0; m < M; m++) fy[m][t] = fh[m];
            if (h.count) ay = rmax(ay, alpha);
        } else {
            for (int m = 0; m < M; m++) sm.face[m][t] = fh[m];
            if (h.count) ax = rmax(ax, alpha);
        }
    }
    __syncthreads();
    {
        const int ci = t / TJ, cj = t - ci * TJ;
        const int gi = i0 + ci, gj = j0 + cj;
        if (ci < TI && gi < a.mx && gj < a.my) {
            const long long idx = (long long)gi * a.my + gj;
            const int c = (ci + 3) * S::HJ + (cj + 3);
            const bool exact_lds = sizeof(R) == sizeof(double);
            const double idx_ = a.idx, idy_ = a.idy;
            const double bdt = stage_bdt(a);
            for (int m = 0; m < M; m++) {
                double L;
                if constexpr (sizeof(R) == 4) {
                    // FP32 form: the flux differences in FP32 (the fluxes are FP32
                    // values), the stage accumulation below in FP64
                    const R lx = (sm.face[m][(ci + 1) * TJ + cj] - sm.face[m][ci * TJ + cj]) * R(idx_);
                    const R ly = (fy[m][ci * (TJ + 1) + cj + 1] - fy[m][ci * (TJ + 1) + cj]) * R(idy_);
                    L = -double(lx + ly);
                } else {
                    const double Lx = -(double(sm.face[m][(ci + 1) * TJ + cj]) - double(sm.face[m][ci * TJ + cj])) * idx_;
                    L = Lx - (double(fy[m][ci * (TJ + 1) + cj + 1]) - double(fy[m][ci * (TJ + 1) + cj])) * idy_;
                }
                const double qa = (exact_lds && a.a_is_src) ? double(sm.cv[m][c]) : pa[m];
                double v = a.a1 * qa;
                if (a.use_b) {
                    const double qb = (exact_lds && a.b_is_src) ? double(sm.cv[m][c]) : pb[m];
                    v = v + a.a2 * qb;
                }
                a.out[m * a.plane + idx] = v + bdt * L;
            }
        }
    }
    if (a.smax) {
        if (ax != ax) ax = R(__builtin_inf());
        if (ay != ay) ay = R(__builtin_inf());
        ax = wave_max64(ax);
        ay = wave_max64(ay);
        if (lane == 0) { red[0][w] = double(ax); red[1][w] = double(ay); }
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

// the tile of the routed kernels (at most 256 faces of each direction)
#ifndef FINESS_ROUTE_TI
#define FINESS_ROUTE_TI 7
#define FINESS_ROUTE_TJ 32
#endif
constexpr int ROUTE_TI = FINESS_ROUTE_TI, ROUTE_TJ = FINESS_ROUTE_TJ;

template <typename R>
static int launch_routed(const StageArgs& a0, hipStream_t st)
{
    constexpr int TI = ROUTE_TI, TJ = ROUTE_TJ;
    StageArgs a = a0;
    const int nti = (a.mx + TI - 1) / TI, ntj = (a.my + TJ - 1) / TJ;
    a.ntj = ntj;
    const dim3 grid(nti * ntj), block(256);
    if (a.route_mask_read || a.route_tile) {
        // a new record: the indicator first, then the stage from the record
        if (!a.route_mask_read) {
            if (a.route_exact)
                hipLaunchKernelGGL((k_classify<Euler2D, R, TI, TJ, true>), grid, block, 0, st, a);
            else
                hipLaunchKernelGGL((k_classify<Euler2D, R, TI, TJ, false>), grid, block, 0, st, a);
        }
        hipLaunchKernelGGL((k_route_full<Euler2D, R, WENO_JS5_P2>), dim3(ROUTE_FULL_BLOCKS + (a.ctl_run ? 1 : 0)),
                           block, 0, st, a);
        hipLaunchKernelGGL((k_stage2d_lookup<Euler2D, R, TI, TJ>), grid, block, 0, st, a);
    } else if (a.route_exact)
        hipLaunchKernelGGL((k_stage2d_routed<Euler2D, R, WENO_JS5_P2, TI, TJ, true>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL((k_stage2d_routed<Euler2D, R, WENO_JS5_P2, TI, TJ, false>), grid, block, 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

long long routed_tiles(int mx, int my)
{
    return (long long)((mx + ROUTE_TI - 1) / ROUTE_TI) * ((my + ROUTE_TJ - 1) / ROUTE_TJ);
}

int launch_stage2d_routed(int physics, int fp32, int wk, const StageArgs& a, hipStream_t st)
{
    if (physics != PHYS_EULER || wk != WENO_JS5_P2) return -1;
    if ((a.route_mask_read || a.route_tile) && !(a.route_tile && a.route_list && a.route_n && a.route_flux && a.out))
        return -1;
    // the controller rides only on a stage that reuses a record (whose
    // troubled-face pass has no indicator launch before it)
    if (a.ctl_run && !(a.route_mask_read && a.route_tile && a.ctl_smax)) return -1;
    return fp32 ? launch_routed<float>(a, st) : launch_routed<double>(a, st);
}

} // namespace finess
