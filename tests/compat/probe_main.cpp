// probe_main.cpp -- CPU harness for the drop-in layer (test infrastructure).
// Links the 2D drop-in library with app callbacks that forward to the CPU
// oracle's restatement of the reference callbacks, a SetBndValues chosen on
// the command line, and prints what IniParams parsed and what the probes
// detected.  No device call is made.
//   probe_main <parameters.ini> <bc: extrap|periodic|mixed|timedep>
//              <physics: euler|euler_rescaled|sw|sw_hr|bogus>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/finess_hip.h"
#include "../../oracle/finess_oracle.h"
#include "IniParams.h"
#include "finess_dropin.h"

static std::string g_bc = "extrap", g_phys = "euler";

static finess_params orc_params()
{
    finess_params p;
    finess_params_default(&p);
    p.physics = (g_phys == "sw" || g_phys == "sw_hr") ? FINESS_PHYS_SHALLOW_WATER : FINESS_PHYS_EULER;
    p.meqn = global_ini_params.get_meqn();
    p.gamma = global_ini_params.get_gamma();
    p.sw_compat_hl = 1;
    return p;
}

void FluxFunc(const dTensor2& xpts, const dTensor2& Q, const dTensor2& Aux, dTensor3& flux)
{
    finess_params p = orc_params();
    const int n = xpts.getsize(1), M = Q.getsize(2);
    for (int i = 1; i <= n; i++) {
        double q[5], f[5], g[5], a[1] = {0.0};
        for (int m = 1; m <= M; m++) q[m - 1] = Q.get(i, m);
        orc_flux(&p, q, a, f, g);
        for (int m = 1; m <= M; m++) {
            flux.set(i, m, 1, g_phys == "bogus" ? 1.5 * f[m - 1] : f[m - 1]);
            flux.set(i, m, 2, g[m - 1]);
        }
    }
}

// euler_rescaled: the first characteristic field scaled by 2 (left row) and
// 1/2 (right column) -- still a pair of inverses, but not the device's
// eigenvectors, which the probe must refuse
void ProjectLeftEig(int ixy, const dTensor1&, const dTensor1& Qa, const dTensor2& Qv, dTensor2& Wv)
{
    finess_params p = orc_params();
    const int M = Qv.getsize(1), n = Qv.getsize(2);
    double qa[5];
    for (int m = 1; m <= M; m++) qa[m - 1] = Qa.get(m);
    orc_project_left(&p, ixy, qa, n, Qv.data(), Wv.data());
    if (g_phys == "euler_rescaled")
        for (int k = 1; k <= n; k++) Wv.set(1, k, 2.0 * Wv.get(1, k));
}

void ProjectRightEig(int ixy, const dTensor1&, const dTensor1& Qa, const dTensor2& Wv, dTensor2& Qv)
{
    finess_params p = orc_params();
    const int M = Wv.getsize(1), n = Wv.getsize(2);
    double qa[5];
    for (int m = 1; m <= M; m++) qa[m - 1] = Qa.get(m);
    dTensor2 W(M, n);
    for (int m = 1; m <= M; m++)
        for (int k = 1; k <= n; k++) W.set(m, k, (g_phys == "euler_rescaled" && m == 1 ? 0.5 : 1.0) * Wv.get(m, k));
    orc_project_right(&p, ixy, qa, n, W.data(), Qv.data());
}

// sw_hr: the shallow-water wave speeds with the right momenta divided by h_r
void SetWaveSpd(const dTensor1& nvec, const dTensor1&, const dTensor1& Ql, const dTensor1& Qr, const dTensor1& Al,
                const dTensor1& Ar, double& s1, double& s2)
{
    finess_params p = orc_params();
    p.sw_compat_hl = g_phys == "sw_hr" ? 0 : 1;
    const int ixy = nvec.get(1) != 0.0 ? 1 : 2;
    orc_wave_speed(&p, ixy, Ql.data(), Qr.data(), Al.data(), Ar.data(), &s1, &s2);
}

void SetBndValues(StateVars& Q)
{
    finess_params p = orc_params();
    dTensorBC3& q = Q.ref_q();
    p.mx = q.getsize(1); p.my = q.getsize(2); p.meqn = q.getsize(3); p.mbc = q.getmbc(); p.ndims = 2; p.maux = 0;
    int bc[4] = {FINESS_BC_EXTRAP, FINESS_BC_EXTRAP, FINESS_BC_EXTRAP, FINESS_BC_EXTRAP};
    if (g_bc == "periodic") for (int k = 0; k < 4; k++) bc[k] = FINESS_BC_PERIODIC;
    if (g_bc == "mixed") { bc[1] = FINESS_BC_PERIODIC; bc[0] = FINESS_BC_PERIODIC; bc[2] = FINESS_BC_REFLECT; bc[3] = FINESS_BC_EXTRAP; }
    for (int k = 0; k < 4; k++) p.bc[k] = bc[k];
    orc_set_bnd(&p, q.data(), nullptr);
    if (g_bc == "timedep") q.set(0, 1, 1, q.get(0, 1, 1) + 1.0);  // not a copy of any cell
    // extrapolation at t = 0, an inflow value later (a clock-driven rule)
    if (g_bc == "clock" && Q.get_t() != 0.0) q.set(2, p.my + 1, 1, 1.0 + Q.get_t());
}

int main(int argc, char* argv[])
{
    if (argc < 4) return 2;
    g_bc = argv[2];
    g_phys = argv[3];
    try {
        global_ini_params.init(argv[1]);
        const IniParams& ini = global_ini_params;
        printf("ini mx=%d my=%d mbc=%d meqn=%d dx=%.17g tsm=%d weno=%d eps=%.3g alpha_scaling=%.3g gamma=%.3g\n",
               ini.get_mx(), ini.get_my(), ini.get_mbc(), ini.get_meqn(), ini.get_dx(),
               (int)ini.get_time_stepping_method(), (int)ini.get_weno_version(), ini.get_epsilon(),
               ini.get_alpha_scaling(), ini.get_gamma());
        finess_dropin::ProbeReport r = finess_dropin::probe_app();
        printf("probe physics=%d checked=%d bc=%d,%d,%d,%d bc_checked=%d sw_compat_hl=%d\n", r.physics,
               (int)r.physics_checked, r.bc[0], r.bc[1], r.bc[2], r.bc[3], (int)r.bc_checked, r.sw_compat_hl);
    } catch (const std::logic_error& e) {
        printf("logic_error: %s\n", e.what());
        return 3;
    } catch (const std::exception& e) {
        printf("error: %s\n", e.what());
        return 4;
    }
    return 0;
}
