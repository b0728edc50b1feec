// Kelvin-Helmholtz shear layer: rho = 2, u = 0.5 in |y - 1/2| < 1/4, else
// rho = 1, u = -0.5; p = 2.5; v = 0.01 sum_{k=1..8} sin(2 pi k x + phi_k)
// exp(-((|y-1/2| - 1/4)/0.05)^2) with phases from splitmix64(seed).
#include <cmath>
#include <cstdint>
#include "IniParams.h"
#include "dogdefs.h"

static double splitmix_uniform(uint64_t seed, int k)
{
    uint64_t z = seed + (uint64_t)k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void QinitFunc(const dTensor2& xpts, dTensor2& qvals)
{
    const double gamma = global_ini_params.get_gamma();
    const uint64_t seed = (uint64_t)std::stoull(global_ini_params.get_gpu("seed", "20240611"));
    double phi[8];
    for (int k = 1; k <= 8; k++) phi[k - 1] = 2.0 * pi * splitmix_uniform(seed, k);
    for (int i = 1; i <= xpts.getsize(1); i++) {
        const double x = xpts.get(i, 1), y = xpts.get(i, 2);
        const bool band = std::fabs(y - 0.5) < 0.25;
        const double rho = band ? 2.0 : 1.0, u = band ? 0.5 : -0.5, p = 2.5;
        double s = 0.0;
        for (int k = 1; k <= 8; k++) s += std::sin(2.0 * pi * k * x + phi[k - 1]);
        const double d = (std::fabs(y - 0.5) - 0.25) / 0.05;
        const double v = 0.01 * s * std::exp(-d * d);
        qvals.set(i, 1, rho);
        qvals.set(i, 2, rho * u);
        qvals.set(i, 3, rho * v);
        qvals.set(i, 4, 0.0);
        qvals.set(i, 5, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v));
    }
}
