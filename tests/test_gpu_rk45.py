"""The 10-stage fourth-order and 8-stage fifth-order Runge-Kutta schemes on
the GPU against the oracle: fixed steps and the controlled run, 1e-10
relative per field (FP64)."""
import numpy as np
import pytest

import orc
from finess_amd import Solver, problems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [4, 5])
def test_fixed_steps(order):
    p, q, _ = problems.riemann4(mx=64, time_order=order)
    q0 = q.copy()
    orc.fixed_steps(p, q0, 2e-3, 6)
    with Solver(p) as s:
        s.set_state(q)
        s.run_fixed_steps(2e-3, 6)
        q1, _ = s.get_state()
    err = orc.rel_linf(orc.interior(p, q1), orc.interior(p, q0))
    print("RK%d fixed steps rel err" % order, err)
    assert np.all(err <= 1e-10), err


@pytest.mark.parametrize("order", [4, 5])
def test_controlled_run(order):
    p, q, _ = problems.riemann4(mx=96, time_order=order)
    q0 = q.copy()
    dtv0 = np.array([0.0, p.initial_dt, p.max_dt])
    rc, t0, st0 = orc.fin_solve(p, q0, 0.0, 0.02, dtv0)
    assert rc == 0
    q1 = q.copy()
    dtv1 = np.array([0.0, p.initial_dt, p.max_dt])
    with Solver(p) as s:
        t1 = s.fin_solve_rk(q1, 0.0, 0.02, dtv1)
    assert abs(t1 - t0) <= 1e-14
    assert abs(dtv1[1] - dtv0[1]) <= 1e-12 * dtv0[1]
    err = orc.rel_linf(orc.interior(p, q1), orc.interior(p, q0))
    print("RK%d controlled run: steps" % order, st0.steps_accepted, "rel err", err)
    assert np.all(err <= 1e-10), err
