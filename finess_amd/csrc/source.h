// source.h -- device forms of the app callback SourceTermFunc.
//
// With source_term set, the reference samples SourceTermFunc at every cell of
// the stage's state (SampleFunction, lib/2d/SampleFunction.cpp) and adds
// psi to the flux differences: Lstar = psi + tmp (lib/2d/ConstructL.cpp:
// 371-387, lib/1d/ConstructL.cpp:194-205).  A host callback cannot run in a
// kernel, so the kernels carry the forms the reference's apps define:
//   SRC_ZERO           psi = 0 (every app's SourceTermFunc except shallow water's)
//   SRC_SW_BATHYMETRY  psi = (0, -h b_x, -h b_y) with aux = (b, b_x, b_y)
//                      (apps/2d/shallow_water/riemann_test/SourceTermFunc.cpp)
//   SRC_DAMPING        psi = -k q, every field (a linear relaxation term)
#pragma once
#include <hip/hip_runtime.h>

namespace finess {

// Kernel argument values: SRC_OFF (no source term, the zero of a cleared
// argument block) or 1 + the public FINESS_SRC_* form.
enum : int { SRC_OFF = 0, SRC_ZERO = 1, SRC_SW_BATHYMETRY = 2, SRC_DAMPING = 3 };

// The kernel value of a context's source term (source_term, source_kind).
__host__ __device__ __forceinline__ int source_arg(int source_term, int source_kind)
{
    return source_term ? 1 + source_kind : SRC_OFF;
}

// psi of one cell: q the cell's conserved state, aux the cell's first aux
// value with successive aux fields `astride` apart (null when unused).
template <int M>
__host__ __device__ __forceinline__ void source_psi(int kind, double coef, const double q[M], const double* aux,
                                                    long long astride, double psi[M])
{
    for (int m = 0; m < M; m++) psi[m] = 0.0;
    if (kind == SRC_DAMPING) {
        for (int m = 0; m < M; m++) psi[m] = -coef * q[m];
    } else if (kind == SRC_SW_BATHYMETRY && M == 3) {
        psi[1] = -q[0] * aux[astride];
        psi[2] = -q[0] * aux[2 * astride];
    }
}

} // namespace finess
