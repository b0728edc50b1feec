// noalloc.hip -- ConstructL on device-resident field planes without any
// device allocation (finess_construct_l_planes_noalloc).
//
// The drop-in finess_construct_l keeps its buffers in a context (staging of
// the reference layout, per-face wave speeds, CFL words), all allocated with
// hipMalloc.  Callers that cannot allocate -- code under HIP graph capture,
// tight serving loops, harnesses that forbid it -- call this variant
// instead: they pass the state and L as device field planes and receive the
// two wave-speed maxima that GetCFL takes (lib/2d/GetCFL.cpp:54-58).  The
// only memory beyond those inputs and outputs is the scratch array below, a
// file-scope device array of fixed size that every call on a device shares.
// Its device address is taken once per device with hipGetSymbolAddress; the
// array's name is never used as a device address on the host.
//
// Calls that overlap on one device race on the scratch words; callers
// serialise them (one stream, or an event between streams).
#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>

#include "../../include/finess_hip.h"
#include "kernels.h"
#include "stage2d_body.h"

namespace finess {

namespace {

// Scratch layout (bytes, with CFL_SPREAD = 64):
//   [0, 1024)      2 x CFL_SPREAD words: per-direction wave-speed maxima of
//                  the stage's workgroups (bit patterns of non-negative doubles)
//   [1024, 2048)   2 x CFL_SPREAD words: GlobalWaveSpd's partial maxima
//   [2048, 2064)   the two global wave speeds (global_alpha)
constexpr size_t SCRATCH_BYTES = 4096 * (CFL_SPREAD / 64);
constexpr size_t OFF_CFL = 0;
constexpr size_t OFF_GSPREAD = OFF_CFL + 2 * CFL_SPREAD * sizeof(unsigned long long);
constexpr size_t OFF_GALPHA = OFF_GSPREAD + 2 * CFL_SPREAD * sizeof(unsigned long long);
constexpr size_t SCRATCH_USED = OFF_GALPHA + 2 * sizeof(double);
static_assert(SCRATCH_USED <= SCRATCH_BYTES, "scratch layout exceeds the array");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_ptr[MAX_DEVICES] = {};

// The device address of this device's scratch array (taken once per device,
// with that device current).
hipError_t scratch_ptr(int dev, unsigned char** out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_ptr[dev];
    return hipSuccess;
}

// Folds the 2 x CFL_SPREAD maxima words into the two speeds (one wave).
__global__ __launch_bounds__(64) void k_fold_cfl_words(const unsigned long long* w, double* out)
{
    const int t = threadIdx.x;
    const double vx = wave_max64(spread_lane_max(w, t));
    const double vy = wave_max64(spread_lane_max(w + CFL_SPREAD, t));
    if (t == 0) {
        out[0] = vx;
        out[1] = vy;
    }
}

} // namespace

} // namespace finess

using namespace finess;

extern "C" int finess_construct_l_planes_noalloc(const finess_params* pin, const double* q, double* L,
                                                 double* smax2, void* stream)
{
    if (!pin || !q || !L || !smax2) return FINESS_ERR_ARG;
    finess_params p = *pin;
    if (finess_params_finalize(&p) != FINESS_OK) return FINESS_ERR_CONFIG;
    // the fused fifth-order stage kernel only: the other orders keep their
    // interface fluxes in grid-sized buffers, and routing keeps a record
    if (p.ndims != 2 || p.space_order != 5 || p.routing != FINESS_ROUTE_NONE) return FINESS_ERR_CONFIG;
    // the bathymetry source term reads aux planes, which this entry point does not take
    if (p.source_term && p.source_kind == FINESS_SRC_SW_BATHYMETRY) return FINESS_ERR_CONFIG;
    // the planes carry no ghosts: a host-side boundary rule has nowhere to go
    for (int k = 0; k < 4; k++)
        if (p.bc[k] == FINESS_BC_HOST) return FINESS_ERR_CONFIG;
    // face and cell indices of the stage kernel are 32-bit
    if ((long long)(p.mx + 1) * (p.my + 1) >= (1ll << 31)) return FINESS_ERR_CONFIG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return FINESS_ERR_NODEVICE;
    if (dev < 0 || dev >= MAX_DEVICES) return FINESS_ERR_NODEVICE;
    unsigned char* s = nullptr;
    if (scratch_ptr(dev, &s) != hipSuccess) return FINESS_ERR_HIP;
    hipStream_t st = static_cast<hipStream_t>(stream);

    unsigned long long* cfl = reinterpret_cast<unsigned long long*>(s + OFF_CFL);
    StageArgs a;
    memset(&a, 0, sizeof a);
    a.mx = p.mx;
    a.my = p.my;
    a.plane = (long long)p.mx * p.my;
    for (int k = 0; k < 4; k++) a.bc[k] = p.bc[k];
    a.dx = p.dx;
    a.dy = p.dy;
    a.idx = 1.0 / p.dx;
    a.idy = 1.0 / p.dy;
    a.alpha_scaling = p.alpha_scaling;
    a.eps = p.epsilon;
    a.power = p.power_param;
    a.pp.gamma = p.gamma;
    a.pp.sw_compat_hl = p.sw_compat_hl;
    if (p.global_alpha) {
        a.gspread = reinterpret_cast<unsigned long long*>(s + OFF_GSPREAD);
        a.galpha = reinterpret_cast<double*>(s + OFF_GALPHA);
    }
    a.src_kind = source_arg(p.source_term, p.source_kind);
    a.src_coef = p.source_coef;
    a.src = q;
    a.qa = q;
    a.a_is_src = 1;
    a.Lout = L;
    a.smax = cfl;
    // the words keep the previous call's maxima: clear them (a kernel, so
    // that the call can be captured in a graph)
    if (launch_fill_u64(cfl, 2 * CFL_SPREAD, 0ull, st)) return FINESS_ERR_HIP;
    const int wk = p.weno_version == FINESS_WENO_FD ? WENO_FD5
                   : p.weno_version == FINESS_WENO_Z ? WENO_Z5
                   : p.power_param == 2.0 ? WENO_JS5_P2 : WENO_JS5_POW;
    const int rc = launch_stage2d(p.physics, p.precision == FINESS_PREC_FP32, wk, 0, a, st);
    if (rc) return rc == -1 ? FINESS_ERR_CONFIG : FINESS_ERR_HIP;
    hipLaunchKernelGGL(k_fold_cfl_words, dim3(1), dim3(64), 0, st, cfl, smax2);
    return hipGetLastError() == hipSuccess ? FINESS_OK : FINESS_ERR_HIP;
}
