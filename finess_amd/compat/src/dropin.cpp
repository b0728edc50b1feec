// dropin.cpp -- ConstructL / FinSolveRK / FinSolveLxW with the reference's
// C++ signatures, running on the GPU through the C-ABI of libfiness_hip.so.
// State crosses the boundary through the tensors' public numel/vget/vset, so
// this file compiles against the reference's own lib/tensors.h as well.
//
// Error behaviour follows the reference: an exceeded step count, dt > max_dt
// and CFL > 1e8 print the reference's messages and exit(1)
// (lib/2d/FinSolveRK.cpp:84-92, lib/2d/GetCFL.cpp:43-67); an unsupported
// configuration throws std::logic_error like GetWenoReconstruct
// (lib/WenoReconstruct.cpp:70); device failures throw std::runtime_error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "../../../include/finess_hip.h"
#include "IniParams.h"
#include "finess_dropin.h"

// The app callbacks are optional for the library: weak references resolve to
// null when an app (or a test) does not link them.
#if NDIMS == 1
#pragma weak FluxFunc
#pragma weak SetWaveSpd
#else
#pragma weak FluxFunc
#pragma weak ProjectLeftEig
#pragma weak ProjectRightEig
#pragma weak SetWaveSpd
#endif
#pragma weak SetBndValues
#pragma weak SourceTermFunc

namespace finess_dropin {

namespace {
finess_ctx* g_ctx = nullptr;
finess_params g_params;
ProbeReport g_report;
bool g_probed = false;

[[noreturn]] void fail_like_reference(int code, const char* what)
{
    switch (code) {
    case FINESS_ERR_NV:
        printf(" Error in FinSolveRK.cpp: ");
        printf("Exceeded allowed # of time steps \n");
        printf("Terminating program.\n");
        exit(1);
    case FINESS_ERR_DT:
        std::cout << std::endl << " Error: dt is out of bounds ... " << std::endl << std::endl;
        exit(1);
    case FINESS_ERR_CFL:
        std::cout << std::endl << " Error: CFL number is out of bounds ... " << std::endl << std::endl;
        exit(1);
    case FINESS_ERR_CONFIG:
        throw std::logic_error(std::string("finess-hip: configuration not supported on the device: ") + what);
    default:
        throw std::runtime_error(std::string("finess-hip: ") + what);
    }
}

int bc_word(const std::string& w)
{
    if (w == "periodic") return FINESS_BC_PERIODIC;
    if (w == "reflective" || w == "wall" || w == "reflect") return FINESS_BC_REFLECT;
    if (w == "extrapolation" || w == "transmissive" || w == "outflow") return FINESS_BC_EXTRAP;
    if (w == "host") return FINESS_BC_HOST;  // the app's SetBndValues on the host, before every stage
    throw std::logic_error("gpu.boundary: unknown boundary kind " + w);
}

// ---------------------------------------------------------------- physics
double rel(double a, double b) { return std::fabs(a - b) / (1.0 + std::fabs(b)); }

#if NDIMS == 2
bool euler_matches(double gamma, std::string& notes)
{
    if (!&FluxFunc) return false;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int n = 64;
    dTensor2 x(n, 2), Q(n, 5), A(n, 1);
    dTensor3 F(n, 5, 2);
    std::vector<double> qs(5 * n);
    for (int i = 1; i <= n; i++) {
        const double rho = 0.3 + U(rng), u = U(rng) - 0.5, v = U(rng) - 0.5, w = 0.0, p = 0.2 + U(rng);
        const double E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w);
        const double q[5] = {rho, rho * u, rho * v, rho * w, E};
        for (int m = 1; m <= 5; m++) Q.set(i, m, q[m - 1]);
        x.set(i, 1, U(rng)); x.set(i, 2, U(rng)); A.set(i, 1, 0.0);
    }
    FluxFunc(x, Q, A, F);
    double err = 0;
    for (int i = 1; i <= n; i++) {
        const double rho = Q.get(i, 1), u = Q.get(i, 2) / rho, v = Q.get(i, 3) / rho, E = Q.get(i, 5);
        const double p = (gamma - 1) * (E - 0.5 * rho * (u * u + v * v));
        const double f[5] = {rho * u, rho * u * u + p, rho * u * v, 0.0, u * (E + p)};
        const double g[5] = {rho * v, rho * v * u, rho * v * v + p, 0.0, v * (E + p)};
        for (int m = 1; m <= 5; m++) err = std::max(err, std::max(rel(F.get(i, m, 1), f[m - 1]), rel(F.get(i, m, 2), g[m - 1])));
    }
    if (err > 1e-12) { notes += "FluxFunc is not the Euler flux; "; return false; }
    return true;
}

// Random admissible Euler (w != 0 included) or shallow-water states.
std::vector<std::vector<double>> probe_states(int physics, double gamma, int n, unsigned seed)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<std::vector<double>> out;
    for (int k = 0; k < n; k++) {
        if (physics == FINESS_PHYS_EULER) {
            const double rho = 0.3 + U(rng), u = 2 * U(rng) - 1, v = 2 * U(rng) - 1, w = U(rng) - 0.5,
                         pr = 0.2 + U(rng);
            out.push_back({rho, rho * u, rho * v, rho * w, pr / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)});
        } else {
            const double h = 0.4 + U(rng), u = 2 * U(rng) - 1, v = 2 * U(rng) - 1;
            out.push_back({h, h * u, h * v});
        }
    }
    return out;
}

double scaled_diff(double a, double b, double scale) { return std::fabs(a - b) / scale; }

// ProjectLeftEig / ProjectRightEig applied to the identity at random mean
// states against the matrices of the device physics: the normalisation and
// the order of the characteristic fields change the WENO result, so the
// app's eigenvectors must be the device's, not merely a pair of inverses.
bool eig_conforms(const finess_params& p, std::string& notes)
{
    if (!&ProjectLeftEig || !&ProjectRightEig) return true;
    const int M = p.meqn;
    dTensor1 aux(std::max(1, p.maux)), qa(M);
    for (int k = 1; k <= aux.getsize(); k++) aux.set(k, 0.0);
    dTensor2 I(M, M), W(M, M);
    for (int a = 1; a <= M; a++)
        for (int b = 1; b <= M; b++) I.set(a, b, a == b ? 1.0 : 0.0);
    std::vector<double> Lm(M * M), Rm(M * M);
    for (const auto& st : probe_states(p.physics, p.gamma, 32, 11)) {
        for (int m = 1; m <= M; m++) qa.set(m, st[m - 1]);
        for (int ixy = 1; ixy <= 2; ixy++) {
            if (finess_physics_eig(&p, ixy, st.data(), Lm.data(), Rm.data()) != FINESS_OK) {
                notes += "no device eigenvectors for this physics; ";
                return false;
            }
            for (int pass = 0; pass < 2; pass++) {
                const std::vector<double>& D = pass ? Rm : Lm;
                if (pass) ProjectRightEig(ixy, aux, qa, I, W);
                else ProjectLeftEig(ixy, aux, qa, I, W);
                double scale = 1e-300;
                for (double v : D) scale = std::max(scale, std::fabs(v));
                for (int a = 1; a <= M; a++)
                    for (int b = 1; b <= M; b++)
                        if (!(scaled_diff(W.get(a, b), D[(a - 1) * M + (b - 1)], scale) <= 1e-10)) {
                            notes += pass ? "ProjectRightEig differs from the device's right eigenvectors; "
                                          : "ProjectLeftEig differs from the device's left eigenvectors "
                                            "(normalisation or field order); ";
                            return false;
                        }
            }
        }
    }
    return true;
}

// SetWaveSpd on unequal states against the device's max(|s1|, |s2|) (the
// only value the flux splitting and the CFL number use).
bool speed_conforms(const finess_params& p, std::string& notes)
{
    if (!&SetWaveSpd) return true;
    const int M = p.meqn;
    const auto st = probe_states(p.physics, p.gamma, 64, 13);
    dTensor1 nvec(2), xedge(2), Ql(M), Qr(M), Al(std::max(1, p.maux)), Ar(std::max(1, p.maux));
    xedge.set(1, 0.5);
    xedge.set(2, 0.5);
    for (int k = 1; k <= Al.getsize(); k++) { Al.set(k, 0.0); Ar.set(k, 0.0); }
    for (size_t k = 0; k + 1 < st.size(); k += 2) {
        for (int m = 1; m <= M; m++) { Ql.set(m, st[k][m - 1]); Qr.set(m, st[k + 1][m - 1]); }
        for (int ixy = 1; ixy <= 2; ixy++) {
            nvec.set(1, ixy == 1 ? 1.0 : 0.0);
            nvec.set(2, ixy == 2 ? 1.0 : 0.0);
            double s1 = 0, s2 = 0, dev = 0;
            SetWaveSpd(nvec, xedge, Ql, Qr, Al, Ar, s1, s2);
            if (finess_physics_wave_speed(&p, ixy, st[k].data(), st[k + 1].data(), Al.data(), Ar.data(), &dev) !=
                FINESS_OK)
                return false;
            const double app = std::max(std::fabs(s1), std::fabs(s2));
            if (!(std::fabs(app - dev) <= 1e-10 * std::max(1.0, dev))) {
                notes += "SetWaveSpd differs from the device's wave speeds; ";
                return false;
            }
        }
    }
    return true;
}

bool euler_conforms(finess_params p, std::string& notes)
{
    p.physics = FINESS_PHYS_EULER;
    return eig_conforms(p, notes) && speed_conforms(p, notes);
}

// Shallow water: the eigenvectors, then SetWaveSpd with the right momenta
// divided by h_l (the reference's own SetWaveSpd.cpp:27-28) or by h_r;
// *hl receives the form that matches.
bool sw_conforms(finess_params p, int* hl, std::string& notes)
{
    p.physics = FINESS_PHYS_SHALLOW_WATER;
    if (!eig_conforms(p, notes)) return false;
    if (!&SetWaveSpd) return true;
    for (int h : {1, 0}) {
        p.sw_compat_hl = h;
        std::string n;
        if (speed_conforms(p, n)) {
            *hl = h;
            return true;
        }
    }
    notes += "SetWaveSpd matches neither shallow-water wave-speed form; ";
    return false;
}

bool sw_matches(std::string& notes)
{
    if (!&FluxFunc) return false;
    std::mt19937_64 rng(9);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int n = 64;
    dTensor2 x(n, 2), Q(n, 3), A(n, 1);
    dTensor3 F(n, 3, 2);
    for (int i = 1; i <= n; i++) {
        const double h = 0.5 + U(rng), u = U(rng) - 0.5, v = U(rng) - 0.5;
        Q.set(i, 1, h); Q.set(i, 2, h * u); Q.set(i, 3, h * v);
        x.set(i, 1, 0.0); x.set(i, 2, 0.0); A.set(i, 1, 0.0);
    }
    FluxFunc(x, Q, A, F);
    double err = 0;
    for (int i = 1; i <= n; i++) {
        const double h = Q.get(i, 1), u = Q.get(i, 2) / h, v = Q.get(i, 3) / h;
        const double f[3] = {h * u, h * u * u + 0.5 * h * h, h * u * v};
        const double g[3] = {h * v, h * v * u, h * v * v + 0.5 * h * h};
        for (int m = 1; m <= 3; m++) err = std::max(err, std::max(rel(F.get(i, m, 1), f[m - 1]), rel(F.get(i, m, 2), g[m - 1])));
    }
    if (err > 1e-12) { notes += "FluxFunc is not the shallow-water flux (g = 1); "; return false; }
    return true;
}
#else
bool advection_matches(std::string& notes)
{
    if (!&FluxFunc) return false;
    const int n = 16;
    dTensor1 x(n);
    dTensor2 Q(n, 1), A(n, 1), F(n, 1);
    for (int i = 1; i <= n; i++) { x.set(i, 0.1 * i); Q.set(i, 1, 0.3 * i - 1.0); A.set(i, 1, 0.5 + 0.1 * i); }
    FluxFunc(x, Q, A, F);
    for (int i = 1; i <= n; i++)
        if (rel(F.get(i, 1), Q.get(i, 1) * A.get(i, 1)) > 1e-14) { notes += "FluxFunc is not aux(1)*q; "; return false; }
    return true;
}
#endif

// ---------------------------------------------------------- boundary probe
// Cells carry distinct positive codes; after the app's SetBndValues every
// ghost cell (corners excluded) names its source, which identifies the rule
// on each side.
#if NDIMS == 2
double code(int i, int j, int m, int my) { return 1.0 + m + 8.0 * (double)(j + 64 + (my + 200) * (double)(i + 64)); }

bool classify_side(StateVars& Q, int side, int mx, int my, int mbc, int meqn, int nrm, int& kind)
{
    const dTensorBC3& q = Q.const_ref_q();
    int votes[3] = {1, 1, 1};  // extrap, periodic, reflect all possible until contradicted
    auto check = [&](int gi, int gj, int ei, int ej, int pi, int pj, int ri, int rj) {
        for (int m = 1; m <= meqn; m++) {
            const double v = q.get(gi, gj, m);
            if (v != code(ei, ej, m, my)) votes[0] = 0;
            if (v != code(pi, pj, m, my)) votes[1] = 0;
            const double rv = (m == nrm ? -1.0 : 1.0) * code(ri, rj, m, my);
            if (v != rv) votes[2] = 0;
        }
    };
    for (int g = 1; g <= mbc; g++) {
        if (side == 0) for (int j = 1; j <= my; j++) check(1 - g, j, 1, j, mx + 1 - g, j, g, j);
        if (side == 1) for (int j = 1; j <= my; j++) check(mx + g, j, mx, j, g, j, mx + 1 - g, j);
        if (side == 2) for (int i = 1; i <= mx; i++) check(i, 1 - g, i, 1, i, my + 1 - g, i, g);
        if (side == 3) for (int i = 1; i <= mx; i++) check(i, my + g, i, my, i, g, i, my + 1 - g);
    }
    const int n = votes[0] + votes[1] + votes[2];
    if (n == 0) return false;
    kind = votes[1] ? FINESS_BC_PERIODIC : votes[0] ? FINESS_BC_EXTRAP : FINESS_BC_REFLECT;
    return true;
}
#else
double code(int i, int m) { return 1.0 + m + 8.0 * (double)(i + 64); }
#endif

// ------------------------------------------------------------ source term
// The app's SourceTermFunc on random states and aux values against the
// device forms (finess_physics_source): FINESS_SRC_* that matches, with *coef
// its coefficient, or -1.
int source_kind_of_app(finess_params p, double* coef)
{
    std::mt19937_64 rng(17);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int n = 32, M = p.meqn, na = std::max(1, p.maux);
#if NDIMS == 2
    dTensor2 x(n, 2);
#else
    dTensor1 x(n);
#endif
    dTensor2 Q(n, M), A(n, na), psi(n, M);
    for (int i = 1; i <= n; i++) {
#if NDIMS == 2
        x.set(i, 1, U(rng));
        x.set(i, 2, U(rng));
#else
        x.set(i, U(rng));
#endif
        for (int m = 1; m <= M; m++) Q.set(i, m, (m == 1 ? 0.5 : -0.5) + U(rng));
        for (int k = 1; k <= na; k++) A.set(i, k, 2 * U(rng) - 1);
    }
    SourceTermFunc(x, Q, A, psi);
    // the relaxation coefficient, if it is one: psi = -k q
    *coef = -psi.get(1, 1) / Q.get(1, 1);
    const int kinds[3] = {FINESS_SRC_ZERO, FINESS_SRC_SW_BATHYMETRY, FINESS_SRC_DAMPING};
    for (int kind : kinds) {
        if (kind == FINESS_SRC_SW_BATHYMETRY && (M != 3 || p.maux < 3)) continue;
        p.source_kind = kind;
        p.source_coef = kind == FINESS_SRC_DAMPING ? *coef : 0.0;
        bool ok = true;
        for (int i = 1; i <= n && ok; i++) {
            double q[8], a[16], d[8];
            for (int m = 1; m <= M; m++) q[m - 1] = Q.get(i, m);
            for (int k = 1; k <= na && k <= 16; k++) a[k - 1] = A.get(i, k);
            if (finess_physics_source(&p, q, a, d) != FINESS_OK) { ok = false; break; }
            for (int m = 1; m <= M; m++)
                if (!(std::fabs(psi.get(i, m) - d[m - 1]) <= 1e-12 * (1.0 + std::fabs(d[m - 1])))) ok = false;
        }
        if (ok) return kind;
    }
    return -1;
}

} // namespace

#if NDIMS == 2
namespace {
// aux of the running FinSolve call (reference layout), for the host
// SetBndValues of FINESS_BC_HOST sides
std::vector<double> g_bnd_aux;

// finess_bnd_fn: the app's SetBndValues on a StateVars holding q at time t
int bnd_trampoline(void*, double t, double* q)
{
    const IniParams& ini = global_ini_params;
    try {
        StateVars Q(t, ini.get_mx(), ini.get_my(), ini.get_meqn(), ini.get_maux(), ini.get_mbc());
        dTensorBC3& qq = Q.ref_q();
        for (int k = 0; k < qq.numel(); k++) qq.vset(k, q[k]);
        dTensorBC3& aux = Q.ref_aux();
        if ((long long)g_bnd_aux.size() == (long long)aux.numel())
            for (int k = 0; k < aux.numel(); k++) aux.vset(k, g_bnd_aux[(size_t)k]);
        SetBndValues(Q);
        for (int k = 0; k < qq.numel(); k++) q[k] = qq.vget(k);
        return 0;
    } catch (...) {
        return 1;
    }
}
} // namespace
#endif

ProbeReport probe_app()
{
    const IniParams& ini = global_ini_params;
    ProbeReport r;
    r.physics = -1;
    r.physics_checked = false;
    r.bc_checked = false;
    r.sw_compat_hl = -1;
    r.source_kind = -1;
    r.source_coef = 0.0;
    finess_params ip;
    ini.to_finess_params(&ip);
    for (int k = 0; k < 4; k++) r.bc[k] = FINESS_BC_EXTRAP;
    const std::string phys = ini.get_gpu("physics", "");
    const int meqn = ini.get_meqn();
#if NDIMS == 2
    if (phys == "euler" || (phys.empty() && meqn == 5)) r.physics = FINESS_PHYS_EULER;
    else if (phys == "shallow_water" || (phys.empty() && meqn == 3)) r.physics = FINESS_PHYS_SHALLOW_WATER;
    if (&FluxFunc) {
        if (r.physics == FINESS_PHYS_EULER) r.physics_checked = euler_matches(ini.get_gamma(), r.notes);
        else if (r.physics == FINESS_PHYS_SHALLOW_WATER) r.physics_checked = sw_matches(r.notes);
        if (!r.physics_checked) throw std::logic_error("finess-hip: the app's physics matches no device physics: " + r.notes);
    }
    // the eigenvectors and wave speeds the app defines must be the device's
    bool conform = true;
    if (r.physics == FINESS_PHYS_EULER) conform = euler_conforms(ip, r.notes);
    else if (r.physics == FINESS_PHYS_SHALLOW_WATER) conform = sw_conforms(ip, &r.sw_compat_hl, r.notes);
    if (!conform) throw std::logic_error("finess-hip: the app's callbacks differ from the device physics: " + r.notes);
#else
    if (phys == "advection" || (phys.empty() && meqn == 1)) r.physics = FINESS_PHYS_ADVECTION;
    if (&FluxFunc) {
        r.physics_checked = advection_matches(r.notes);
        if (!r.physics_checked) throw std::logic_error("finess-hip: the app's physics matches no device physics: " + r.notes);
    }
#endif
    if (r.physics < 0) throw std::logic_error("finess-hip: no device physics for meqn = " + std::to_string(meqn));
    if (ini.get_source_term() && &SourceTermFunc) {
        ip.physics = r.physics;
        r.source_kind = source_kind_of_app(ip, &r.source_coef);
        if (r.source_kind < 0)
            throw std::logic_error("finess-hip: the app's SourceTermFunc matches no device source term");
    }

    const std::string words = ini.get_gpu("boundary", "");
    if (!words.empty()) {
        std::istringstream in(words);
        std::string w;
        for (int k = 0; k < 4 && (in >> w); k++) r.bc[k] = bc_word(w);
    } else if (&SetBndValues) {
        const int mx = ini.get_mx(), mbc = ini.get_mbc(), maux = ini.get_maux();
#if NDIMS == 2
        const int my = ini.get_my();
        const int nx = r.physics == FINESS_PHYS_ADVECTION ? -1 : 2, ny = r.physics == FINESS_PHYS_ADVECTION ? -1 : 3;
        // at two times: a side whose rule is not one of the device's, or
        // changes with time (inflow such as apps/2d/euler/mach_reflection/
        // SetBndValues.cpp:103-139), runs the app's SetBndValues on the host
        // before every stage (FINESS_BC_HOST)
        for (int pass = 0; pass < 2; pass++) {
            StateVars Q(pass ? 1.0 : 0.0, mx, my, meqn, maux, mbc);
            dTensorBC3& q = Q.ref_q();
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                for (int j = 1 - mbc; j <= my + mbc; j++)
                    for (int m = 1; m <= meqn; m++)
                        q.set(i, j, m, (i >= 1 && i <= mx && j >= 1 && j <= my) ? code(i, j, m, my) : NAN);
            SetBndValues(Q);
            for (int side = 0; side < 4; side++) {
                int kind = 0;
                if (!classify_side(Q, side, mx, my, mbc, meqn, side < 2 ? nx : ny, kind) ||
                    (pass && kind != r.bc[side]))
                    kind = FINESS_BC_HOST;
                if (!pass || r.bc[side] != FINESS_BC_HOST) r.bc[side] = kind;
            }
        }
#else
        StateVars Q(0.0, mx, meqn, maux, mbc);
        dTensorBC2& q = Q.ref_q();
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= meqn; m++) q.set(i, m, (i >= 1 && i <= mx) ? code(i, m) : NAN);
        SetBndValues(Q);
        for (int side = 0; side < 2; side++) {
            int votes[2] = {1, 1};
            for (int g = 1; g <= mbc; g++)
                for (int m = 1; m <= meqn; m++) {
                    const int gi = side ? mx + g : 1 - g;
                    const double v = q.get(gi, m);
                    if (v != code(side ? mx : 1, m)) votes[0] = 0;
                    if (v != code(side ? g : mx + 1 - g, m)) votes[1] = 0;
                }
            if (!votes[0] && !votes[1])
                throw std::logic_error("finess-hip: 1D SetBndValues is neither periodic nor extrapolation");
            r.bc[side] = votes[1] ? FINESS_BC_PERIODIC : FINESS_BC_EXTRAP;
        }
        r.bc[2] = r.bc[3] = r.bc[0];
#endif
        r.bc_checked = true;
    }
    return r;
}

void release()
{
    if (g_ctx) finess_destroy(g_ctx);
    g_ctx = nullptr;
    g_probed = false;
}

namespace {
finess_ctx* context()
{
    if (!g_probed) {
        g_report = probe_app();
        g_probed = true;
    }
    finess_params p;
    global_ini_params.to_finess_params(&p);
    p.physics = g_report.physics;
    for (int k = 0; k < 4; k++) p.bc[k] = g_report.bc[k];
    if (g_report.sw_compat_hl >= 0) p.sw_compat_hl = g_report.sw_compat_hl;
    if (g_report.source_kind >= 0) {
        p.source_kind = g_report.source_kind;
        p.source_coef = g_report.source_coef;
    }
    int rc = finess_params_finalize(&p);
    if (rc) fail_like_reference(rc, "parameter combination");
    if (g_ctx && std::memcmp(&p, &g_params, sizeof p) == 0) return g_ctx;
    if (g_ctx) finess_destroy(g_ctx);
    g_ctx = nullptr;
    rc = finess_create(&p, &g_ctx);
    if (rc) fail_like_reference(rc, "finess_create failed");
#if NDIMS == 2
    if (p.bc[0] == FINESS_BC_HOST || p.bc[1] == FINESS_BC_HOST || p.bc[2] == FINESS_BC_HOST ||
        p.bc[3] == FINESS_BC_HOST) {
        if (!&SetBndValues) throw std::logic_error("finess-hip: gpu.boundary = host needs the app's SetBndValues");
        rc = finess_set_bnd_callback(g_ctx, bnd_trampoline, nullptr);
        if (rc) fail_like_reference(rc, finess_last_error(g_ctx));
    }
#endif
    // ConSoln after every accepted step, reduced on the device (without a
    // capacity function; with one, con_soln_frame writes a line per call)
    if (global_ini_params.get_mcapa() < 1) {
        rc = finess_set_step_log(g_ctx, 1);
        if (rc) fail_like_reference(rc, finess_last_error(g_ctx));
    }
    if (p.routing == FINESS_ROUTE_LEARNED) {
        // [gpu] indicator: the trained network, FINESS_MLP_NPARAM float32
        // parameters followed by the float32 logit threshold
        const std::string path = global_ini_params.get_gpu("indicator", "");
        std::vector<float> w(FINESS_MLP_NPARAM + 1);
        FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "rb");
        const size_t got = f ? std::fread(w.data(), sizeof(float), w.size(), f) : 0;
        if (f) std::fclose(f);
        if (got != w.size())
            throw std::logic_error("finess-hip: gpu.routing = learned needs gpu.indicator = <file of " +
                                   std::to_string(FINESS_MLP_NPARAM + 1) + " float32 values>; could not read '" +
                                   path + "'");
        rc = finess_set_indicator(g_ctx, w.data(), w[FINESS_MLP_NPARAM]);
        if (rc) fail_like_reference(rc, finess_last_error(g_ctx));
    }
    g_params = p;
    return g_ctx;
}

template <class T>
std::vector<double> gather(const T& t)
{
    std::vector<double> v((size_t)t.numel());
    for (int k = 0; k < t.numel(); k++) v[k] = t.vget(k);
    return v;
}
template <class T>
void scatter(const std::vector<double>& v, T& t)
{
    for (int k = 0; k < t.numel(); k++) t.vset(k, v[k]);
}

} // namespace
// ConSoln's sums on the host in its order (lib/2d/ConSoln.cpp:44-73,
// lib/1d/ConSoln.cpp:36-60): field by field, i then j, dx*dy*q (times the
// capacity aux(mcapa) when mcapa >= 1).
void con_soln_sums(const StateVars& Q, double* s)
{
    const IniParams& ini = global_ini_params;
    const int meqn = ini.get_meqn(), mx = ini.get_mx(), mcapa = ini.get_mcapa();
    const double dx = ini.get_dx();
#if NDIMS == 1
    for (int m = 1; m <= meqn; m++) {
        double acc = 0.0;
        for (int i = 1; i <= mx; i++) {
            const double qtmp = Q.const_ref_q().get(i, m);
            acc = mcapa < 1 ? acc + dx * qtmp : acc + Q.const_ref_aux().get(i, mcapa) * dx * qtmp;
        }
        s[m - 1] = acc;
    }
#else
    const int my = ini.get_my();
    const double dy = ini.get_dy();
    for (int m = 1; m <= meqn; m++) {
        double acc = 0.0;
        for (int i = 1; i <= mx; i++)
            for (int j = 1; j <= my; j++) {
                const double qtmp = Q.const_ref_q().get(i, j, m);
                acc = mcapa < 1 ? acc + dx * dy * qtmp : acc + Q.const_ref_aux().get(i, j, mcapa) * dx * dy * qtmp;
            }
        s[m - 1] = acc;
    }
#endif
}
namespace {

// One line of conservation.dat in ConSoln's format (lib/2d/ConSoln.cpp:
// 82-94): t and the sums, %24.16e, tiny sums printed as zero.
void con_soln_line(std::ofstream& f, double t, const double* s, int meqn)
{
    f << std::setprecision(16);
    f << std::setw(24) << std::scientific << t << " ";
    for (int m = 0; m < meqn; m++) f << std::setw(24) << std::scientific << (std::fabs(s[m]) < 1.0e-99 ? 0.0 : s[m]) << " ";
    f << std::endl;
}

// Appends the steps' conservation records of the last FinSolve call; with a
// capacity function, one line for the state at its end, summed on the host
// as ConSoln does.
void con_soln_steps(finess_ctx* c, const StateVars& Q)
{
    const IniParams& ini = global_ini_params;
    const int meqn = ini.get_meqn();
    std::ofstream f((ini.get_output_dir() + "/conservation.dat").c_str(), std::ofstream::app);
    if (ini.get_mcapa() < 1) {
        std::vector<double> rec((size_t)(meqn + 1) * 256);
        int n;
        while ((n = finess_read_step_log(c, rec.data(), 256)) > 0)
            for (int k = 0; k < n; k++) con_soln_line(f, rec[(size_t)k * (meqn + 1)], &rec[(size_t)k * (meqn + 1) + 1], meqn);
        if (n < 0) fail_like_reference(n, finess_last_error(c));
        return;
    }
    std::vector<double> s((size_t)meqn);
    con_soln_sums(Q, s.data());
    con_soln_line(f, Q.get_t(), s.data(), meqn);
}

void fin_solve(StateVars& Qnew, double tend, double dtv[], int ts)
{
    finess_ctx* c = context();
    std::vector<double> q = gather(Qnew.const_ref_q()), aux = gather(Qnew.const_ref_aux());
    double t = Qnew.get_t();
#if NDIMS == 2
    g_bnd_aux = aux;
#endif
    const int rc = ts == FINESS_TS_LXW   ? finess_fin_solve_lxw(c, q.data(), aux.data(), &t, tend, dtv)
                   : ts == FINESS_TS_PIF ? finess_fin_solve_rk_pif(c, q.data(), aux.data(), &t, tend, dtv)
                                         : finess_fin_solve_rk(c, q.data(), aux.data(), &t, tend, dtv);
    if (rc) fail_like_reference(rc, finess_last_error(c));
    scatter(q, Qnew.ref_q());
    Qnew.set_t(t);
    con_soln_steps(c, Qnew);
}
} // namespace

void con_soln(const StateVars& Q)
{
    const IniParams& ini = global_ini_params;
    const int meqn = ini.get_meqn();
    std::vector<double> s((size_t)meqn);
    con_soln_sums(Q, s.data());
    std::ofstream f((ini.get_output_dir() + "/conservation.dat").c_str(),
                    Q.get_t() == 0 ? std::ofstream::out : std::ofstream::app);
    con_soln_line(f, Q.get_t(), s.data(), meqn);
}
} // namespace finess_dropin

using namespace finess_dropin;

#if NDIMS == 2
void ConstructL(StateVars& Q, dTensorBC3& Lstar, dTensorBC3& smax)
{
    finess_ctx* c = context();
    std::vector<double> q = gather(Q.const_ref_q()), aux = gather(Q.const_ref_aux());
    std::vector<double> L = gather(Lstar), sm = gather(smax);
    const int rc = finess_construct_l(c, q.data(), aux.data(), L.data(), sm.data());
    if (rc) fail_like_reference(rc, finess_last_error(c));
    scatter(L, Lstar);
    scatter(sm, smax);
}

void FinSolveRK(StateVars& Qnew, double tend, double dtv[]) { fin_solve(Qnew, tend, dtv, FINESS_TS_RK); }
void FinSolveLxW(StateVars& Qnew, double tend, double dtv[]) { fin_solve(Qnew, tend, dtv, FINESS_TS_LXW); }
// lib/2d/FinSolveRK_PIF.cpp:66, and FinSolveUser wrapping it as that file does (:31-35)
void FinSolveRK_PIF(StateVars& Qnew, double tend, double dtv[]) { fin_solve(Qnew, tend, dtv, FINESS_TS_PIF); }
void FinSolveUser(StateVars& Qnew, double tend, double dtv[]) { FinSolveRK_PIF(Qnew, tend, dtv); }
#else
void ConstructL(const StateVars& Q, dTensorBC2& Lstar, dTensorBC1& smax)
{
    finess_ctx* c = context();
    std::vector<double> q = gather(Q.const_ref_q()), aux = gather(Q.const_ref_aux());
    std::vector<double> L = gather(Lstar), sm = gather(smax);
    const int rc = finess_construct_l(c, q.data(), aux.data(), L.data(), sm.data());
    if (rc) fail_like_reference(rc, finess_last_error(c));
    scatter(L, Lstar);
    scatter(sm, smax);
}

void FinSolveRK(StateVars& Qnew, double tend, double dtv[]) { fin_solve(Qnew, tend, dtv, FINESS_TS_RK); }
#endif
