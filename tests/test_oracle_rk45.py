"""The oracle's 10-stage fourth-order (SSP(10,4)) and 8-stage fifth-order
Runge-Kutta schemes (lib/2d/FinSolveRK.cpp:213-305, coefficients
lib/2d/SetRKinfo.cpp:66-129), pinned by their observed temporal order: on a
fixed grid (the spatial operator is then the same in every run) the
difference to a run with a 64 times smaller step falls as dt^4, resp. dt^5.
A wrong coefficient or a stage taken from the wrong register drops the order
to one or two."""
import math

import numpy as np
import pytest

import orc
from finess_amd import problems


def _run(order, dt, T, mx=24):
    p, q, _ = problems.density_wave(mx=mx, amp=0.3, time_order=order)
    orc.fixed_steps(p, q, dt, int(round(T / dt)))
    g = p.mbc
    return q[g:g + mx, g:g + mx]


@pytest.mark.parametrize("order,least", [(4, 3.8), (5, 4.6)])
def test_temporal_order(order, least):
    T = 0.1
    ref = _run(5, T / 256, T)
    errs = [np.abs(_run(order, T / k, T) - ref).max() for k in (4, 8, 16)]
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print("RK%d errors" % order, errs, "orders", orders)
    assert min(orders) > least
