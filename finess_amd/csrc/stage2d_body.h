// stage2d_body.h -- one fused 2D Runge-Kutta / Lax-Wendroff stage.
//
// Replaces, per stage, the reference's sequence
//   SetBndValues(q)                    (app SetBndValues.cpp)
//   ConstructL(q, Lstar, smax)         (lib/2d/ConstructL.cpp:13-408)
//   UpdateSoln(a1, a2, b, dt, ...)     (lib/2d/UpdateSoln.cpp:14-38)
// and the wave-speed part of GetCFL (lib/2d/GetCFL.cpp:54-58).
//
// A workgroup of 256 threads owns a TI x TJ = 7 x 32 tile of cells (j, the
// contiguous index, along TJ).  Phase 1 stages the tile plus a 3-cell
// cross-shaped halo in LDS, mapping halo cells through the boundary condition
// on the fly (the halo fill) and computing each cell's record (state +
// primitives) once.  The (TI+1) x TJ = 256 x-faces are then computed one per
// thread -- characteristic projection at the mean state, local Lax-Friedrichs
// splitting, WENO5 on every characteristic field for both split parts,
// projection back -- and differenced into the x part of L; the same LDS face
// buffer then takes the TI x (TJ+1) = 231 y-faces.  The last phase forms L
// and writes  out = a1*qa + a2*qb + (b dt)*L  in FP64.  One face per thread
// and direction keeps every wave equally busy, and sharing one face buffer
// between the directions keeps the LDS footprint at three workgroups per CU.
// The per-face wave speeds feed a workgroup max that the caller reduces into
// the step's CFL slot.
//
// The phases are plain functions of the thread index so that a CPU harness
// can run the very same code one thread at a time.
#pragma once
#include "physics.h"
#include "source.h"
#include "step_control.h"
#include "weno.h"

namespace finess {

// Boundary condition kinds (values equal the FINESS_BC_* macros).
// BC_HOST sides take their ghosts from a host callback: the stage then reads
// a ghosted copy of its source, and the boundary map treats them as
// extrapolation (it only forms that copy's corners, which no stencil reads).
enum : int { BC_EXTRAP = 0, BC_PERIODIC = 1, BC_REFLECT = 2, BC_HOST = 3 };

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

// Workgroup maxima and counters are folded into CFL_SPREAD separate words
// per quantity (chosen by block index): a single word hit by every
// workgroup of a launch serialises thousands of atomics in one L2 channel.
// The host reduces the CFL_SPREAD words when it reads them.
#ifndef FINESS_CFL_SPREAD
#define FINESS_CFL_SPREAD 64
#endif
constexpr int CFL_SPREAD = FINESS_CFL_SPREAD;
static_assert(CFL_SPREAD % 64 == 0, "the readers take CFL_SPREAD / 64 words per lane");

// The maximum of one direction's spread words w[0 .. CFL_SPREAD) (bit
// patterns of non-negative doubles or +inf) over the words lane, lane + 64,
// ... of this lane; a wave maximum of the results folds them all.
__device__ __forceinline__ double spread_lane_max(const unsigned long long* w, int lane)
{
    unsigned long long m = w[lane];
    for (int k = lane + 64; k < CFL_SPREAD; k += 64) m = w[k] > m ? w[k] : m;
    return __longlong_as_double((long long)m);
}

// Maximum over the 64 lanes of a wave (the wave-speed maxima: non-negative
// doubles or +inf).
//
// A DPP ladder: the maximum over each quad (quad permutations), half row
// (row_half_mirror) and row (row_mirror), then row 0's into row 1 and row
// 2's into row 3 (row_bcast15), then rows 0-1's into rows 2-3 (row_bcast31),
// so that lane 63 holds the wave's maximum; it is returned in every lane.
// Lanes a broadcast does not write keep their own value (`old` = src).
// Register moves between lanes instead of LDS-crossbar permutes.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_move(int x)
{
    return __builtin_amdgcn_update_dpp(x, x, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_max_step(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_move<CTRL, ROWS>((int)b), hi = dpp_move<CTRL, ROWS>((int)(b >> 32));
    const double w = __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    return w > v ? w : v;
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_max_step(float v)
{
    const float w = __int_as_float(dpp_move<CTRL, ROWS>(__float_as_int(v)));
    return w > v ? w : v;
}
template <typename T>
__device__ __forceinline__ T wave_max64_dpp(T v)
{
    v = dpp_max_step<0xB1>(v);        // quad_perm [1,0,3,2]
    v = dpp_max_step<0x4E>(v);        // quad_perm [2,3,0,1]
    v = dpp_max_step<0x141>(v);       // row_half_mirror
    v = dpp_max_step<0x140>(v);       // row_mirror
    v = dpp_max_step<0x142, 0xA>(v);  // row_bcast15 into rows 1 and 3
    v = dpp_max_step<0x143, 0xC>(v);  // row_bcast31 into rows 2 and 3
    return v;
}
__device__ __forceinline__ double wave_max64(double v)
{
    v = wave_max64_dpp(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// the same for FP32 values (one move per step)
__device__ __forceinline__ float wave_max64(float v)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_max64_dpp(v)), 63));
}
// the FP32 maximum by lane permutes (the batched stage measured 1.3% faster
// with it than with the DPP ladder; the other kernels gain with the ladder)
__device__ __forceinline__ float wave_max64_xor(float v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Remapping the
// launch index gives each XCD a contiguous band of tiles, so the halo rows
// that neighbouring tiles share are found in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    constexpr int NXCD = 8;
    if (n % NXCD) return b;
    return (b % NXCD) * (n / NXCD) + b / NXCD;
}

// Routing record of one tile: the troubled faces the tile owns, as one
// 64-bit ballot per wave of 64 faces and direction (bit = lane = the face's
// thread within the wave), and the troubled-list slot of its first troubled
// face of each direction (offset of the direction's list not included).
// A face on the tile's far edge is owned by the next tile, where it is that
// tile's first row (x) or column (y) of faces.
struct RouteTileRec {
    unsigned long long mask[2][4];
    int base[2];
    int pad[2];
};

struct StageArgs {
    // state planes, field-major: q[m][i][j] at m*plane + i*ld + j (ld = my)
    const double* src;     // stencil operand
    const double* qa;      // UpdateSoln's "qstar" operand (a1)
    const double* qb;      // UpdateSoln's "qnew" operand (a2)
    double* out;
    // Lax-Wendroff: time-integrated fluxes on the extended grid
    // [-3, mx+2] x [-3, my+2], plane (mx+6)*(my+6), F at m*lplane + (i+3)*(my+6) + (j+3),
    // stored in the stencil precision R (FP32 in the FP32 form)
    const void* lxw_F;
    const void* lxw_G;
    // ghosts stored in src instead of mapped (drop-in ConstructL mode):
    // src planes are (mx+2*gh) x (my+2*gh)
    int src_ghosts;
    int mx, my;
    long long plane;       // mx*my
    int ntj;               // tiles along j
    int bc[4];             // left, right, bottom, top
    int a_is_src, b_is_src, use_b;
    double a1, a2, bdt;
    // device-controlled stepping (null: off): the stage's dt is read on the
    // device, bdt = beta * (*dtp), and every kernel of the run returns at
    // once while *halt is non-zero
    const double* dtp;
    double beta;
    const unsigned long long* halt;
    double dx, dy;
    double idx, idy;  // 1/dx, 1/dy (IEEE, formed on the host)
    double alpha_scaling;
    double eps, power;
    PhysParams pp;
    // wave-speed maxima (bit patterns of non-negative doubles),
    // [2][CFL_SPREAD] per slot: x then y
    unsigned long long* smax;
    // troubled-interface routing (routed kernel only): network fragments,
    // or the exact weight-departure test when route_exact is set
    const struct MlpFrags* frags;
    int route_exact;
    double route_tau;
    unsigned long long* route_counts;  // [2][CFL_SPREAD]: troubled faces, faces
    // per-step routing record (null: indicator only, nothing stored).
    //   route_tile[tile] the tile's troubled owned faces (RouteTileRec; tile
    //                    = ti * ntj + tj of the routed kernels' 7 x 32 tiles)
    //   route_list[slot] the face's (i, j) packed as i << 16 | j; troubled x
    //                    faces take slots 0.., y faces (mx+1)*my..
    //   route_n[2]       number of troubled x faces, y faces (zero when the
    //                    indicator starts a record)
    //   route_n_clear[2] set to zero by the indicator stage: the counter pair
    //                    of the previous record, which the next record uses
    //   route_flux[slot] 5 flux components + alpha of the troubled face, in the
    //                    stencil precision R (6 floats in the FP32 form)
    // route_mask_read: 0 = the indicator decides and records, 1 = read the record
    struct RouteTileRec* route_tile;
    int* route_list;
    unsigned long long* route_n;
    unsigned long long* route_n_clear;
    void* route_flux;
    int route_mask_read;
    // optional face-flux / L output for the drop-in ConstructL (may be null):
    double* Lout;          // L planes (mx*my)
    double* smax_faces_x;  // (mx+1) x my alphas of x-faces, row-major i*(my)+j
    double* smax_faces_y;  // mx x (my+1)
    // limited Lax-Wendroff (null: off): the interface fluxes are written
    // (x faces [m][(mx+1) my], y faces [m][mx (my+1)]) for the limiter, and
    // the faces at (0, 0) stay out of the CFL maxima, whose value there the
    // limiter sets to the global wave speeds
    double* face_x;
    double* face_y;
    // global_alpha (null: off): galpha[0..1] the global wave speeds of the
    // stage source in x and y (lib/2d/GlobalWaveSpd.cpp:20-80), formed by
    // launch_stage2d through the [2][CFL_SPREAD] words of gspread
    double* galpha;
    unsigned long long* gspread;
    // source term (src_kind SRC_OFF: none): psi of the stage source's cell
    // is added to L; aux planes (interior cells, `plane` apart per field)
    int src_kind;
    double src_coef;
    const double* aux;
    // the previous step's device controller, run by the first workgroup of
    // the troubled-face pass (null: none); its CFL words ctl_smax
    struct StepCtl* ctl_run;
    unsigned long long* ctl_smax;
    StepCtlParams ctl_p;
};

// beta * dt of the stage update (dt from the device controller when set)
__host__ __device__ __forceinline__ double stage_bdt(const StageArgs& a)
{
    return a.dtp ? a.beta * *a.dtp : a.bdt;
}

// The device controller stopped the run (a rejected step, the end, an error):
// the kernels enqueued after it do nothing.
__host__ __device__ __forceinline__ bool stage_halted(const StageArgs& a)
{
    return a.halt && *a.halt != 0ull;
}

template <class P, typename R, int TI, int TJ>
struct StageSmem {
    static constexpr int HI = TI + 6, HJ = TJ + 6, NC = HI * HJ;
    static constexpr int NFX = (TI + 1) * TJ, NFY = TI * (TJ + 1);
    static constexpr int NF = NFX > NFY ? NFX : NFY;
    R cv[P::NV][NC];
    R face[P::MEQN][NF];
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

// Returns whether a cell this thread loaded has a nonzero third momentum
// (rho w, 5-field systems; see face_part).
template <class P, typename R, int TI, int TJ>
__host__ __device__ __forceinline__ bool phase_load(int tid, int nthreads, int i0, int j0, const StageArgs& a,
                                                    StageSmem<P, R, TI, TJ>& sm)
{
    using S = StageSmem<P, R, TI, TJ>;
    // A tile whose halo lies inside the grid (most of them) reads its cells
    // directly, without the boundary map; the others map every index.
    const bool inner = !a.src_ghosts && i0 >= 3 && j0 >= 3 && i0 + TI + 3 <= a.mx && j0 + TJ + 3 <= a.my;
    // two cells per thread and round: both cells' loads are issued before
    // either record is formed, so their memory latencies overlap
    constexpr int K = 2;
    bool nz3 = false;
    for (int c0 = tid; c0 < S::NC; c0 += K * nthreads) {
        double q[K][P::MEQN];
        bool ok[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int c = c0 + k * nthreads;
            const int ci = c / S::HJ, cj = c - ci * S::HJ;
            const bool ri = ci < 3 || ci >= TI + 3, rj = cj < 3 || cj >= TJ + 3;
            const int gi = i0 + ci - 3, gj = j0 + cj - 3;
            // corners are never read; nor is anything beyond every face's stencil
            ok[k] = c < S::NC && !(ri && rj) && gi <= a.mx + 2 && gj <= a.my + 2;
            if (ok[k]) {
                if (inner) {
                    const long long idx = (long long)gi * a.my + gj;
                    for (int m = 0; m < P::MEQN; m++) q[k][m] = a.src[m * a.plane + idx];
                } else {
                    load_state<P>(a, gi, gj, q[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!ok[k]) continue;
            if constexpr (P::MEQN == 5) nz3 = nz3 || q[k][3] != 0.0;
            double rec[P::NV];
            P::template cell_record<sizeof(R) == 4>(q[k], a.pp, rec);
            for (int v = 0; v < P::NV; v++) sm.cv[v][c0 + k * nthreads] = R(rec[v]);
        }
    }
    return nz3;
}

// The time-integrated flux values of field 3 (rho w) that the tile's faces
// read: F3 along the x-face stencils (extended rows i0 .. i0+TI+5), G3
// along the y-face stencils (extended columns j0 .. j0+TJ+5), at most three
// per thread of 256 (zero where there is none).  On the Lax-Wendroff path
// face_part's SKIP3 needs them zero as well as the state's rho w; they are
// fetched before the tile load and tested after it, so that the two loads'
// latencies overlap.
template <class P, typename R, int TI, int TJ>
__device__ __forceinline__ void lxw_fetch3(int tid, int i0, int j0, const StageArgs& a, R v[3])
{
    constexpr int NX = (TI + 6) * TJ, NY = TI * (TJ + 6);
    static_assert(NX + NY <= 3 * 256, "three values per thread of 256");
    for (int k = 0; k < 3; k++) v[k] = R(0);
    if constexpr (P::MEQN == 5) {
        const long long lld = a.my + 6, lpl = (long long)(a.mx + 6) * lld;
        const R* F = static_cast<const R*>(a.lxw_F) + 3 * lpl;
        const R* G = static_cast<const R*>(a.lxw_G) + 3 * lpl;
        for (int k = 0; k < 3; k++) {
            const int c = tid + 256 * k;
            if (c < NX) {
                const int r = c / TJ, col = c - r * TJ;
                const int er = i0 + r;
                if (er <= a.mx + 5 && j0 + col < a.my) v[k] = F[er * lld + (j0 + col + 3)];
            } else if (c < NX + NY) {
                const int e = c - NX, r = e / (TJ + 6), col = e - r * (TJ + 6);
                const int ec = j0 + col;
                if (i0 + r < a.mx && ec <= a.my + 5) v[k] = G[(i0 + r + 3) * lld + ec];
            }
        }
    }
}

// The old-state operands of the thread's cell for the stage update, read at
// the start of the kernel so that their latency hides behind the tile load
// (those the FP64 tile already holds exactly are taken from LDS instead).
template <class P, typename R, int TI, int TJ>
__host__ __device__ __forceinline__ void phase_prefetch(int tid, int i0, int j0, const StageArgs& a,
                                                        double qa[P::MEQN], double qb[P::MEQN])
{
    const int ci = tid / TJ, cj = tid - ci * TJ;
    const int gi = i0 + ci, gj = j0 + cj;
    for (int m = 0; m < P::MEQN; m++) { qa[m] = 0.0; qb[m] = 0.0; }
    if (ci >= TI || gi >= a.mx || gj >= a.my || !a.out) return;
    const long long idx = (long long)gi * a.my + gj;
    const bool exact_lds = sizeof(R) == sizeof(double) && !a.src_ghosts;
    if (!(exact_lds && a.a_is_src))
        for (int m = 0; m < P::MEQN; m++) qa[m] = a.qa[m * a.plane + idx];
    if (a.use_b && !(exact_lds && a.b_is_src))
        for (int m = 0; m < P::MEQN; m++) qb[m] = a.qb[m * a.plane + idx];
}

// Mean state of a face from the two cells beside it, its eigen-system and
// its Lax-Friedrichs speed (lib/2d/ConstructL.cpp:72-77 and 149-174 for x,
// :224-229 and 300-325 for y); returns alpha and sets la = ascale * alpha.
// amin: the global speed of the direction when global_alpha is set, else 0
// (alpha = Max(alpha1, local), lib/2d/ConstructL.cpp:172,323).
template <class P, typename R, int NC>
__host__ __device__ __forceinline__ R face_setup(const R (*cv)[NC], int base, int stride, int dir,
                                                 const PhysParams& pp, R ascale,
                                                 typename P::template Eig<R>& e, R& la, R amin = R(0))
{
    constexpr int M = P::MEQN, NV = P::NV;
    R cl[NV], cr[NV], ql[M], qr[M];
    for (int v = 0; v < NV; v++) {
        cl[v] = cv[v][base + 2 * stride];
        cr[v] = cv[v][base + 3 * stride];
    }
    P::state(cl, dir, ql);
    P::state(cr, dir, qr);
    R qa[M];
    for (int k = 0; k < M; k++) qa[k] = R(0.5) * (qr[k] + ql[k]);
    P::eig(qa, pp, e);
    const R ws = P::wave_speed(ql, qr, cl, cr, pp);
    const R alpha = amin > ws ? amin : ws;  // a NaN speed stays NaN
    la = ascale * alpha;
    return alpha;
}

// One Lax-Friedrichs split part of a face flux, projected and reconstructed
// (lib/2d/ConstructL.cpp:138-192 for x, :289-342 for y).  The reference forms gp = (L f + la L q)/2
// and gm = (L f - la L q)/2; L is linear, so L(f + la q) (part 0, stencil
// cells 0..4) and L(f - la q) (part 1, cells 5..1) are projected instead and
// the 1/2 is applied after the reconstruction: every WENO5 variant satisfies
// W(2g; 4 eps) = 2 W(g; eps) exactly in binary arithmetic.  w[k] receives
// twice the reconstructed characteristic flux of field k.
// SKIP3 (5-field systems; the caller sets it only when every stencil cell
// has rho w = 0): field 3's flux, the mean state's w and so field 3's
// characteristic values are all zero, its reconstruction is zero and is
// not computed (2 of the 10 reconstructions of a face).
template <class P, typename R, int WK, bool LXW, int NC, bool SKIP3 = false>
__host__ __device__ __forceinline__ void face_part(const R (*cv)[NC], int base, int stride, int dir,
                                                   const R* lflux, long long lbase, long long lstride,
                                                   long long lplane, const typename P::template Eig<R>& e,
                                                   R la, const WenoParams<R>& wp4, int part, R w[P::MEQN])
{
    constexpr int M = P::MEQN, NV = P::NV;
    const R sg = part ? R(-1) : R(1);
    R g[M][5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
        const int s = part ? 5 - n : n;  // stencil point; g[.][n] is its position in the reconstruction
        R rec[NV], q[M], f[M];
        for (int v = 0; v < NV; v++) rec[v] = cv[v][base + s * stride];
        P::state(rec, dir, q);
        if (LXW) {
            for (int k = 0; k < M; k++) f[k] = lflux[P::rot_q(k, dir) * lplane + lbase + s * lstride];
        } else {
            P::flux(q, rec, f);
        }
        R vs[M], t[M];
        for (int k = 0; k < M; k++) vs[k] = f[k] + (sg * la) * q[k];
        P::left(e, vs, t);
        for (int k = 0; k < M; k++) g[k][n] = t[k];
    }
    for (int k = 0; k < M; k++) {
        if (SKIP3 && M == 5 && k == 3) w[k] = R(0);
        else w[k] = weno5<WK>(g[k], wp4);
    }
}

// One face: characteristic WENO-LLF flux (lib/2d/ConstructL.cpp:63-208 for
// x, :215-358 for y).  base/stride locate the 6 stencil records i-3..i+2.
template <class P, typename R, int WK, bool LXW, int NC, bool SKIP3 = false>
__host__ __device__ __forceinline__ R face_flux(const R (*cv)[NC], int base, int stride, int dir,
                                                const R* lflux, long long lbase, long long lstride,
                                                long long lplane, const PhysParams& pp,
                                                const WenoParams<R>& wp, R ascale, R fh[P::MEQN],
                                                R amin = R(0))
{
    constexpr int M = P::MEQN;
    typename P::template Eig<R> e;
    R la;
    const R alpha = face_setup<P, R, NC>(cv, base, stride, dir, pp, ascale, e, la, amin);
    const WenoParams<R> wp4{R(4) * wp.eps, wp.power};
    // The two split parts are reconstructed one after the other: only one
    // part's 25 projected values are live at a time, at the price of forming
    // the flux of cells 1..4 twice.
    R gh[M];
    auto part_flux = [&](int part) {
        R w[M];
        face_part<P, R, WK, LXW, NC, SKIP3>(cv, base, stride, dir, lflux, lbase, lstride, lplane, e, la, wp4, part, w);
        for (int k = 0; k < M; k++) gh[k] = part ? R(0.5) * (gh[k] + w[k]) : w[k];
    };
    // (5-field systems reading a tile: a rolled loop keeps the compiler from
    // interleaving the two parts, which would make both live again; a
    // register-resident 6-point stencil needs static indices instead)
    if constexpr (M >= 5 && NC > 6) {
#pragma unroll 1
        for (int part = 0; part < 2; part++) part_flux(part);
    } else {
        part_flux(0);
        part_flux(1);
    }
    R fr[M];
    P::right(e, gh, fr);
    for (int k = 0; k < M; k++) fh[P::rot_q(k, dir)] = fr[k];
    return alpha;
}

// One face per thread for direction dir (0: x-faces, thread t -> face
// (t / TJ, t % TJ), i.e. between cells fi-1 and fi; 1: y-faces, thread t ->
// (t / (TJ+1), t % (TJ+1))).  Writes the face flux to sm.face and folds its
// wave speed into amax when GetCFL counts it (faces mx+1/2 and my+1/2 are
// skipped, lib/2d/GetCFL.cpp:54-58).
// Geometry of face number f of direction dir in the tile (see phase_faces).
struct FaceGeom {
    int fi, fj, base, stride;
    long long lbase, lstride;
    bool valid, count;
};

template <int TI, int TJ>
__host__ __device__ __forceinline__ FaceGeom face_geom(int f, int dir, int i0, int j0, const StageArgs& a)
{
    constexpr int HJ = TJ + 6;
    const long long lld = a.my + 6;
    FaceGeom g;
    if (dir == 0) {
        g.fi = f / TJ; g.fj = f - g.fi * TJ;
        g.base = g.fi * HJ + (g.fj + 3); g.stride = HJ;
        const int gi = i0 + g.fi, gj = j0 + g.fj;
        g.valid = f < (TI + 1) * TJ && gi <= a.mx && gj < a.my;
        g.count = gi < a.mx;
        g.lbase = (long long)gi * lld + (gj + 3); g.lstride = lld;
    } else {
        g.fi = f / (TJ + 1); g.fj = f - g.fi * (TJ + 1);
        g.base = (g.fi + 3) * HJ + g.fj; g.stride = 1;
        const int gi = i0 + g.fi, gj = j0 + g.fj;
        g.valid = f < TI * (TJ + 1) && gi < a.mx && gj <= a.my;
        g.count = gj < a.my;
        g.lbase = (long long)(gi + 3) * lld + gj; g.lstride = 1;
    }
    return g;
}

// The face fluxes go to sm.face, or to fout (field m at fout[m * fld + tid])
// when it is given.  amax: the running maximum of the face wave speeds, in
// FP64 or in the stencil precision R (a maximum of R values is exact in R).
template <class P, typename R, int WK, bool LXW, int TI, int TJ, typename A>
__host__ __device__ __forceinline__ void phase_faces(int tid, int dir, int i0, int j0, const StageArgs& a,
                                                     StageSmem<P, R, TI, TJ>& sm, A& amax, bool skip3 = false,
                                                     R* fout = nullptr, int fld = 0)
{
    using S = StageSmem<P, R, TI, TJ>;
    constexpr int M = P::MEQN;
    const long long lpl = (long long)(a.mx + 6) * (a.my + 6);
    const FaceGeom g = face_geom<TI, TJ>(tid, dir, i0, j0, a);
    if (!g.valid) return;
    const int fi = g.fi, fj = g.fj, base = g.base, stride = g.stride;
    const long long lbase = g.lbase, lstride = g.lstride;
    bool count = g.count;
    const WenoParams<R> wp{R(a.eps), R(a.power)};
    R fh[M];
    const R amin = a.galpha ? R(a.galpha[dir]) : R(0);
    const R* lflux = static_cast<const R*>(dir ? a.lxw_G : a.lxw_F);
    // (skip3: a tile-uniform choice between two forms of the face code)
    const R alpha = skip3 ? face_flux<P, R, WK, LXW, S::NC, true>(sm.cv, base, stride, dir, lflux, lbase, lstride, lpl,
                                                                  a.pp, wp, R(a.alpha_scaling), fh, amin)
                          : face_flux<P, R, WK, LXW, S::NC>(sm.cv, base, stride, dir, lflux, lbase, lstride, lpl,
                                                            a.pp, wp, R(a.alpha_scaling), fh, amin);
    (void)sizeof(S);
    if (fout) {
        for (int m = 0; m < M; m++) fout[m * fld + tid] = fh[m];
    } else {
        for (int m = 0; m < M; m++) sm.face[m][tid] = fh[m];
    }
    if constexpr (LXW) {
        if (a.face_x) {
            const int gi = i0 + fi, gj = j0 + fj;
            if (gi == 0 && gj == 0) count = false;
            // a face shared by two tiles is written by the tile for which it is not the far edge
            const bool own = dir ? (fj < TJ || gj == a.my) : (fi < TI || gi == a.mx);
            if (own) {
                const long long nf = dir ? (long long)a.mx * (a.my + 1) : (long long)(a.mx + 1) * a.my;
                const long long f = dir ? (long long)gi * (a.my + 1) + gj : (long long)gi * a.my + gj;
                double* dst = dir ? a.face_y : a.face_x;
                for (int m = 0; m < M; m++) dst[m * nf + f] = double(fh[m]);
            }
        }
    }
    if (count) amax = rmax(amax, A(alpha));
    if (dir == 0 && a.smax_faces_x) a.smax_faces_x[(long long)(i0 + fi) * a.my + (j0 + fj)] = double(alpha);
    if (dir == 1 && a.smax_faces_y) a.smax_faces_y[(long long)(i0 + fi) * (a.my + 1) + (j0 + fj)] = double(alpha);
}

// x part of L for the thread's cell: -(Fhat(i+1) - Fhat(i))/dx
// (ConstructL.cpp:388-399).
template <class P, typename R, int TI, int TJ>
__host__ __device__ __forceinline__ void phase_lx(int tid, const StageArgs& a, const StageSmem<P, R, TI, TJ>& sm,
                                                  double Lx[P::MEQN])
{
    const int ci = tid / TJ, cj = tid - ci * TJ;
    if (ci >= TI) return;
    const double idx_ = a.idx;
    for (int m = 0; m < P::MEQN; m++)
        Lx[m] = -(double(sm.face[m][(ci + 1) * TJ + cj]) - double(sm.face[m][ci * TJ + cj])) * idx_;
}

// L = Lx - (Ghat(j+1) - Ghat(j))/dy and the stage update in FP64 (the y
// fluxes from sm.face, or from fy with leading dimension fyld).
template <class P, typename R, int TI, int TJ>
__host__ __device__ __forceinline__ void phase_update(int tid, int i0, int j0, const StageArgs& a,
                                                      const StageSmem<P, R, TI, TJ>& sm, const double Lx[P::MEQN],
                                                      const double pa[P::MEQN], const double pb[P::MEQN],
                                                      const R* fy = nullptr, int fyld = 0)
{
    using S = StageSmem<P, R, TI, TJ>;
    constexpr int M = P::MEQN;
    const int ci = tid / TJ, cj = tid - ci * TJ;
    const int gi = i0 + ci, gj = j0 + cj;
    if (ci >= TI || gi >= a.mx || gj >= a.my) return;
    const long long idx = (long long)gi * a.my + gj;
    const int c = (ci + 3) * S::HJ + (cj + 3);
    const bool exact_lds = sizeof(R) == sizeof(double) && !a.src_ghosts;
    const double idy_ = a.idy;
    const double bdt = stage_bdt(a);
    double psi[M];
    if (a.src_kind != SRC_OFF) {
        // the stage source's state of the cell: exact in an FP64 tile, else
        // the FP64 operand when the source is one (a narrowed record otherwise)
        double qs[M];
        for (int m = 0; m < M; m++) {
            qs[m] = double(sm.cv[m][c]);
            if (sizeof(R) != sizeof(double) && a.out && a.a_is_src) qs[m] = pa[m];
            else if (sizeof(R) != sizeof(double) && a.out && a.use_b && a.b_is_src) qs[m] = pb[m];
        }
        source_psi<M>(a.src_kind, a.src_coef, qs, a.aux ? a.aux + idx : nullptr, a.plane, psi);
    }
    for (int m = 0; m < M; m++) {
        const R* gm = fy ? fy + m * fyld : sm.face[m];
        double L = Lx[m] - (double(gm[ci * (TJ + 1) + cj + 1]) - double(gm[ci * (TJ + 1) + cj])) * idy_;
        if (a.src_kind != SRC_OFF) L = psi[m] + L;  // Lstar = psi + tmp (lib/2d/ConstructL.cpp:378-386)
        if (a.Lout) a.Lout[m * a.plane + idx] = L;
        if (!a.out) continue;
        const double qa = (exact_lds && a.a_is_src) ? double(sm.cv[m][c]) : pa[m];
        double v = a.a1 * qa;
        if (a.use_b) {
            const double qb = (exact_lds && a.b_is_src) ? double(sm.cv[m][c]) : pb[m];
            v = v + a.a2 * qb;
        }
        a.out[m * a.plane + idx] = v + bdt * L;
    }
}

} // namespace finess
