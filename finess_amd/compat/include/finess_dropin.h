// finess_dropin.h -- the operators this library substitutes into a FINESS app
// at link time (the reference's Makefile object variables, lib/2d/Makefile.defs:
// 79-101), with the reference's C++ signatures, plus the app callbacks the
// probes consult.  Build with -DNDIMS=1 for 1D apps, -DNDIMS=2 (default) for 2D.
#ifndef FINESS_DROPIN_H
#define FINESS_DROPIN_H

#include <string>
#include "StateVars.h"
#include "tensors.h"

#if NDIMS == 1
// lib/1d/ConstructL.cpp:16
void ConstructL(const StateVars& Q, dTensorBC2& Lstar, dTensorBC1& smax);
// lib/1d/FinSolveRK.cpp:27
void FinSolveRK(StateVars& Qnew, double tend, double dtv[]);
// app callbacks (lib/1d/ConstructL.h:4-13), consulted by the probes if linked
void FluxFunc(const dTensor1& xpts, const dTensor2& Q, const dTensor2& Aux, dTensor2& flux);
void SetWaveSpd(const dTensor1& xedge, const dTensor1& Ql, const dTensor1& Qr, const dTensor1& Auxl,
                const dTensor1& Auxr, double& s1, double& s2);
// apps/1d/*/SourceTermFunc.cpp
void SourceTermFunc(const dTensor1& xpts, const dTensor2& qvals, const dTensor2& auxvals, dTensor2& psi);
#else
// lib/2d/FinSolveRK.h:18
void ConstructL(StateVars& Q, dTensorBC3& Lstar, dTensorBC3& smax);
// lib/2d/RunFinpack.h:18-21
void FinSolveRK(StateVars& Qnew, double tend, double dtv[]);
void FinSolveLxW(StateVars& Qnew, double tend, double dtv[]);
// lib/2d/FinSolveRK_PIF.cpp:31-35, 66 (RK4-PIF; the reference links it as FinSolveUser)
void FinSolveRK_PIF(StateVars& Qnew, double tend, double dtv[]);
void FinSolveUser(StateVars& Qnew, double tend, double dtv[]);
// app callbacks (lib/2d/ConstructL.h:7-34), consulted by the probes if linked
void FluxFunc(const dTensor2& xpts, const dTensor2& Q, const dTensor2& Aux, dTensor3& flux);
void ProjectLeftEig(int ixy, const dTensor1& Aux_ave, const dTensor1& Q_ave, const dTensor2& Qvals, dTensor2& Wvals);
void ProjectRightEig(int ixy, const dTensor1& Aux_ave, const dTensor1& Q_ave, const dTensor2& Wvals, dTensor2& Qvals);
void SetWaveSpd(const dTensor1& nvec, const dTensor1& xedge, const dTensor1& Ql, const dTensor1& Qr,
                const dTensor1& Auxl, const dTensor1& Auxr, double& s1, double& s2);
// apps/2d/*/SourceTermFunc.cpp (sampled by lib/2d/SampleFunction.cpp)
void SourceTermFunc(const dTensor2& xpts, const dTensor2& qvals, const dTensor2& auxvals, dTensor2& psi);
#endif
void SetBndValues(StateVars& Q);

namespace finess_dropin {
// Result of the probes run when the first operator is called.
struct ProbeReport {
    int physics;          // FINESS_PHYS_*
    int bc[4];            // FINESS_BC_* (left, right, bottom, top)
    bool physics_checked; // the app's callbacks were linked and matched
    bool bc_checked;      // the app's SetBndValues was linked and classified
    int sw_compat_hl;     // shallow water: 1 if the app's SetWaveSpd divides the right
                          // momenta by h_l, 0 if by h_r, -1 not probed
    int source_kind;      // FINESS_SRC_* matching the app's SourceTermFunc (-1: none probed)
    double source_coef;   // its coefficient (FINESS_SRC_DAMPING)
    std::string notes;
};
// Identifies the physics, boundary conditions and source term from the app's
// own callbacks (or the [gpu] section when they are not linked): FluxFunc on
// random states, ProjectLeftEig / ProjectRightEig applied to the identity and
// SetWaveSpd on unequal states, all against the device physics evaluated on
// the host (finess_physics_*), SetBndValues on coded cells, SourceTermFunc on
// random states.  Throws std::logic_error when any of them cannot be run on
// the device as the app defines it.
ProbeReport probe_app();
// Releases the device context held by the operators.
void release();
// ConSoln (lib/2d/ConSoln.cpp:13-95, lib/1d/ConSoln.cpp): the sums of the
// conserved fields of Q computed on the host in ConSoln's order, and the
// line it writes to <output_dir>/conservation.dat (the file restarts at t = 0).
// The operators append the lines of every accepted step themselves.
void con_soln_sums(const StateVars& Q, double* sums);
void con_soln(const StateVars& Q);
}

#endif
