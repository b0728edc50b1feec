// Zeroth-order extrapolation on all four sides (corners left untouched).
#include "IniParams.h"
#include "StateVars.h"

void SetBndValues(StateVars& Q)
{
    dTensorBC3& q = Q.ref_q();
    const int mx = q.getsize(1), my = q.getsize(2), meqn = q.getsize(3), mbc = q.getmbc();
    for (int g = 1; g <= mbc; g++)
        for (int m = 1; m <= meqn; m++) {
            for (int j = 1; j <= my; j++) {
                q.set(1 - g, j, m, q.get(1, j, m));
                q.set(mx + g, j, m, q.get(mx, j, m));
            }
            for (int i = 1; i <= mx; i++) {
                q.set(i, 1 - g, m, q.get(i, 1, m));
                q.set(i, my + g, m, q.get(i, my, m));
            }
        }
}
