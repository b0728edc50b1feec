"""RK4-PIF (FinSolveRK_PIF, lib/2d/FinSolveRK_PIF.cpp) on the GPU against the
oracle: fixed steps on the four-quadrant problem and with mixed boundary
rules, and the controlled run; 1e-10 relative per field (FP64)."""
import numpy as np
import pytest

import orc
from finess_amd import BC_EXTRAP, BC_PERIODIC, BC_REFLECT, TS_PIF, Solver, problems

pytestmark = pytest.mark.gpu


def _fixed(p, q, dt, n):
    q0 = q.copy()
    orc.fixed_steps(p, q0, dt, n)
    with Solver(p) as s:
        s.set_state(q)
        s.run_fixed_steps(dt, n)
        q1, _ = s.get_state()
    err = orc.rel_linf(orc.interior(p, q1), orc.interior(p, q0))
    print("PIF %d steps rel err" % n, err)
    assert np.all(err <= 1e-10), err


def test_pif_fixed_steps_riemann4():
    p, q, _ = problems.riemann4(mx=64, time_order=4, time_stepping_method=TS_PIF)
    _fixed(p, q, 2e-3, 10)


def test_pif_fixed_steps_mixed_boundaries():
    p, q, _ = problems.riemann4(mx=48, my=40, time_order=4, time_stepping_method=TS_PIF,
                                bc=[BC_EXTRAP, BC_PERIODIC, BC_REFLECT, BC_REFLECT])
    _fixed(p, q, 2e-3, 10)


def test_pif_controlled_run():
    p, q, _ = problems.riemann4(mx=96, time_order=4, time_stepping_method=TS_PIF)
    q0 = q.copy()
    dtv0 = np.array([0.0, p.initial_dt, p.max_dt])
    rc, t0, st0 = orc.fin_solve(p, q0, 0.0, 0.02, dtv0, pif=True)
    assert rc == 0
    q1 = q.copy()
    dtv1 = np.array([0.0, p.initial_dt, p.max_dt])
    with Solver(p) as s:
        t1 = s.fin_solve_rk_pif(q1, 0.0, 0.02, dtv1)
    assert abs(t1 - t0) <= 1e-14
    assert abs(dtv1[1] - dtv0[1]) <= 1e-12 * dtv0[1]
    err = orc.rel_linf(orc.interior(p, q1), orc.interior(p, q0))
    print("PIF controlled run: steps", st0.steps_accepted, "rel err", err)
    assert np.all(err <= 1e-10), err


def test_pif_shallow_water():
    """Shallow water (g = 1): fixed steps on the periodic bump with random
    momenta, and a controlled run."""
    p, q, _ = problems.sw_bump(nx=40, ny=36, time_order=4, time_stepping_method=TS_PIF)
    rng = np.random.default_rng(17)
    q[..., 1] = 0.1 * rng.random(q.shape[:-1])
    q[..., 2] = 0.1 * rng.random(q.shape[:-1])
    _fixed(p, q.copy(), 4e-3, 10)
    q0 = q.copy()
    dtv0 = np.array([0.0, p.initial_dt, p.max_dt])
    rc, t0, st0 = orc.fin_solve(p, q0, 0.0, 0.05, dtv0, pif=True)
    assert rc == 0
    q1 = q.copy()
    dtv1 = np.array([0.0, p.initial_dt, p.max_dt])
    with Solver(p) as s:
        t1 = s.fin_solve_rk_pif(q1, 0.0, 0.05, dtv1)
    assert abs(t1 - t0) <= 1e-14
    err = orc.rel_linf(orc.interior(p, q1), orc.interior(p, q0))
    print("SW PIF controlled run: steps", st0.steps_accepted, "rel err", err)
    assert np.all(err <= 1e-10), err
