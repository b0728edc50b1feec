// refapp_main.cpp -- CPU harness (test infrastructure) that links an app's own,
// unchanged callback sources, compiled from their files by the test, with the
// drop-in layer and runs the drop-in probes on them.  No device call is made.
// Prints what the probes found, the app's initial state at two points and,
// with -DAPP_OPT, the app's own parameter accessor get_OPT().
//   refapp_main <parameters.ini>
#include <cstdio>
#include <stdexcept>

#include "IniParams.h"
#include "finess_dropin.h"

#if NDIMS == 2
void QinitFunc(const dTensor2& xpts, dTensor2& qvals);
#else
void QinitFunc(const dTensor1& xpts, dTensor2& qvals);
#endif

int main(int argc, char* argv[])
{
    if (argc < 2) return 2;
    try {
        global_ini_params.init(argv[1]);
        const finess_dropin::ProbeReport r = finess_dropin::probe_app();
        printf("probe physics=%d checked=%d bc=%d,%d,%d,%d bc_checked=%d sw_compat_hl=%d source_kind=%d\n",
               r.physics, (int)r.physics_checked, r.bc[0], r.bc[1], r.bc[2], r.bc[3], (int)r.bc_checked,
               r.sw_compat_hl, r.source_kind);
        const int meqn = global_ini_params.get_meqn();
#if NDIMS == 2
        dTensor2 x(2, 2), q(2, meqn);
        x.set(1, 1, 0.25); x.set(1, 2, 0.25);
        x.set(2, 1, 0.75); x.set(2, 2, 0.75);
#else
        dTensor1 x(2);
        dTensor2 q(2, meqn);
        x.set(1, 0.25);
        x.set(2, 0.75);
#endif
        QinitFunc(x, q);
        printf("qinit %.6g %.6g\n", q.get(1, 1), q.get(2, 1));
#ifdef APP_OPT
        printf("OPT=%d\n", global_ini_params.get_OPT());
#endif
    } catch (const std::logic_error& e) {
        printf("logic_error: %s\n", e.what());
        return 3;
    } catch (const std::exception& e) {
        printf("error: %s\n", e.what());
        return 4;
    }
    return 0;
}
