// lxw2d.hip -- Lax-Wendroff time-integrated fluxes (ConstructIntegratedR,
// lib/2d/ConstructIntegratedR.cpp:59-296) on the extended grid
// [-3, mx+2] x [-3, my+2] that the following WENO stage reads.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "lxw_body.h"

namespace finess {

template <class P, int TI, int TJ>
__global__ __launch_bounds__(TI* TJ) void k_lxw_flux(LxwArgs a, int ntj)
{
    __shared__ LxwSmem<TI, TJ, P::MEQN> sm;
    const int ti = blockIdx.x / ntj, tj = blockIdx.x - ti * ntj;
    const int e0i = ti * TI, e0j = tj * TJ;
    lxw_phase_load<P, TI, TJ>(threadIdx.x, TI * TJ, e0i, e0j, a, sm);
    __syncthreads();
    lxw_phase_cell<P, TI, TJ>(threadIdx.x, e0i, e0j, a, sm);
}

int launch_lxw_flux(int physics, const LxwArgs& a, hipStream_t st)
{
    constexpr int TI = 8, TJ = 32;
    const int nti = (a.mx + 6 + TI - 1) / TI, ntj = (a.my + 6 + TJ - 1) / TJ;
    if (physics == PHYS_EULER)
        hipLaunchKernelGGL((k_lxw_flux<Euler2D, TI, TJ>), dim3(nti * ntj), dim3(TI * TJ), 0, st, a, ntj);
    else if (physics == PHYS_SHALLOW_WATER)
        hipLaunchKernelGGL((k_lxw_flux<ShallowWater2D, TI, TJ>), dim3(nti * ntj), dim3(TI * TJ), 0, st, a, ntj);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace finess
