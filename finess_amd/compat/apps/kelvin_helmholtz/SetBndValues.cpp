// Doubly periodic boundary including the corners (needed by Lax-Wendroff).
#include "StateVars.h"

void SetBndValues(StateVars& Q)
{
    dTensorBC3& q = Q.ref_q();
    const int mx = q.getsize(1), my = q.getsize(2), meqn = q.getsize(3), mbc = q.getmbc();
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            if (i >= 1 && i <= mx && j >= 1 && j <= my) continue;
            const int si = i < 1 ? i + mx : (i > mx ? i - mx : i);
            const int sj = j < 1 ? j + my : (j > my ? j - my : j);
            for (int m = 1; m <= meqn; m++) q.set(i, j, m, q.get(si, sj, m));
        }
}
