// Periodic boundary for a 1D grid.
#include "StateVars.h"

void SetBndValues(StateVars& Q)
{
    dTensorBC2& q = Q.ref_q();
    dTensorBC2& a = Q.ref_aux();
    const int mx = q.getsize(1), meqn = q.getsize(2), mbc = q.getmbc(), na = a.getsize(2);
    for (int g = 1; g <= mbc; g++) {
        for (int m = 1; m <= meqn; m++) {
            q.set(1 - g, m, q.get(mx + 1 - g, m));
            q.set(mx + g, m, q.get(g, m));
        }
        for (int m = 1; m <= na; m++) {
            a.set(1 - g, m, a.get(mx + 1 - g, m));
            a.set(mx + g, m, a.get(g, m));
        }
    }
}
