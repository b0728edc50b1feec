// Four-quadrant Riemann data for 2D Euler (gamma from [euler]), split at
// (0.5, 0.5); (rho, u, v, p) per quadrant.
#include "IniParams.h"
#include "dogdefs.h"

void QinitFunc(const dTensor2& xpts, dTensor2& qvals)
{
    const double gamma = global_ini_params.get_gamma();
    for (int i = 1; i <= xpts.getsize(1); i++) {
        const double x = xpts.get(i, 1), y = xpts.get(i, 2);
        double rho, u, v, p;
        if (y >= 0.5) {
            if (x >= 0.5) { rho = 1.5; u = 0.0; v = 0.0; p = 1.5; }
            else { rho = 0.5323; u = 1.206; v = 0.0; p = 0.3; }
        } else {
            if (x >= 0.5) { rho = 0.5323; u = 0.0; v = 1.206; p = 0.3; }
            else { rho = 0.138; u = 1.206; v = 1.206; p = 0.029; }
        }
        qvals.set(i, 1, rho);
        qvals.set(i, 2, rho * u);
        qvals.set(i, 3, rho * v);
        qvals.set(i, 4, 0.0);
        qvals.set(i, 5, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v));
    }
}
