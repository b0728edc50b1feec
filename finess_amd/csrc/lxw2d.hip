// lxw2d.hip -- Lax-Wendroff time-integrated fluxes (ConstructIntegratedR,
// lib/2d/ConstructIntegratedR.cpp:59-296) on the extended grid
// [-3, mx+2] x [-3, my+2] that the following WENO stage reads.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "lxw_body.h"

namespace finess {

template <class P, typename R, int TI, int TJ>
__global__ __launch_bounds__(TI* TJ) void k_lxw_flux(LxwArgs a, int ntj)
{
    __shared__ LxwSmem<TI, TJ, P::MEQN, R> sm;
    const int ti = blockIdx.x / ntj, tj = blockIdx.x - ti * ntj;
    const int e0i = ti * TI, e0j = tj * TJ;
    lxw_phase_load<P, TI, TJ, R>(threadIdx.x, TI * TJ, e0i, e0j, a, sm);
    __syncthreads();
    lxw_phase_cell<P, TI, TJ, R>(threadIdx.x, e0i, e0j, a, sm);
}

// fp32: the FP32 stencil form (Euler), whose Lax-Wendroff stage reads the
// fluxes as FP32: differences, hyper-dual products and Taylor terms in FP32.
int launch_lxw_flux(int physics, int fp32, const LxwArgs& a, hipStream_t st)
{
    constexpr int TI = 8, TJ = 32;
    const int nti = (a.mx + 6 + TI - 1) / TI, ntj = (a.my + 6 + TJ - 1) / TJ;
    if (physics == PHYS_EULER && fp32)
        hipLaunchKernelGGL((k_lxw_flux<Euler2D, float, TI, TJ>), dim3(nti * ntj), dim3(TI * TJ), 0, st, a, ntj);
    else if (physics == PHYS_EULER)
        hipLaunchKernelGGL((k_lxw_flux<Euler2D, double, TI, TJ>), dim3(nti * ntj), dim3(TI * TJ), 0, st, a, ntj);
    else if (physics == PHYS_SHALLOW_WATER && !fp32)
        hipLaunchKernelGGL((k_lxw_flux<ShallowWater2D, double, TI, TJ>), dim3(nti * ntj), dim3(TI * TJ), 0, st, a,
                           ntj);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace finess
