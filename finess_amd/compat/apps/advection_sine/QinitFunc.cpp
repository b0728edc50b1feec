// q(x, 0) = sin(2 pi x), advection speed aux(1) = 1.
#include <cmath>
#include "dogdefs.h"

void QinitFunc(const dTensor1& xpts, dTensor2& qvals)
{
    for (int i = 1; i <= xpts.getsize(); i++) qvals.set(i, 1, std::sin(2.0 * pi * xpts.get(i)));
}

void AuxFunc(const dTensor1& xpts, dTensor2& auxvals)
{
    for (int i = 1; i <= xpts.getsize(); i++) auxvals.set(i, 1, 1.0);
}

void FluxFunc(const dTensor1& xpts, const dTensor2& Q, const dTensor2& Aux, dTensor2& flux)
{
    for (int i = 1; i <= xpts.getsize(); i++) flux.set(i, 1, Q.get(i, 1) * Aux.get(i, 1));
}
