"""The oracle's RK4-PIF stepper (FinSolveRK_PIF, lib/2d/FinSolveRK_PIF.cpp:
three component-wise WENO stages without characteristic projection or
wave-speed splitting (ConstructL_NOC), the time-averaged fluxes
(R1 + 2 R2 + 2 R3 + R4)/6, characteristic WENO on them and one Euler step),
pinned by what the mathematics fixes:

- on the smooth 2D Euler density wave, whose exact solution is a
  translation, the error falls at the scheme's order (fourth in time,
  fifth in space) as the grid and the step are refined together; a wrong
  stage weight, a stage taken from the wrong state or a lost flux sample
  drops the order to one or two;
- on a periodic grid the totals of all fields are conserved (flux form);
- a uniform state stays uniform.
"""
import math

import numpy as np

import orc
from finess_amd import TS_PIF, problems


def _pif_wave(mx, **kw):
    return problems.density_wave(mx=mx, time_order=4, time_stepping_method=TS_PIF, **kw)


def _run(p, q, tend, dt):
    dtv = np.array([0.0, dt, dt])
    rc, t, st = orc.fin_solve(p, q, 0.0, tend, dtv, pif=True)
    assert rc == 0 and abs(t - tend) < 1e-14
    return st


def test_order_of_accuracy_density_wave():
    errs = []
    T = 0.1
    for mx in (16, 32, 64):
        p, q, _ = _pif_wave(mx, cfl=10.0)
        dt = 0.5 / mx / 3.0  # CFL about 0.45 with |u| + c below 1.9
        _run(p, q, T, dt)
        g = p.mbc
        X = p.xlow + (np.arange(mx) + 0.5) * p.dx
        Y = p.ylow + (np.arange(mx) + 0.5) * p.dy
        XX, YY = np.meshgrid(X, Y, indexing="ij")
        exact = problems.density_wave_state(XX, YY, T)[..., 0]
        errs.append(np.mean(np.abs(q[g:g + mx, g:g + mx, 0] - exact)))
    orders = [math.log2(errs[k] / errs[k + 1]) for k in range(2)]
    print("PIF density wave L1 errors", errs, "orders", orders)
    assert orders[-1] > 3.7


def test_conservation_periodic():
    p, q, _ = _pif_wave(32, amp=0.5)
    g = p.mbc
    before = q[g:g + 32, g:g + 32].sum(axis=(0, 1))
    _run(p, q, 0.05, 0.004)
    after = q[g:g + 32, g:g + 32].sum(axis=(0, 1))
    assert np.all(np.abs(after - before) <= 1e-12 * np.abs(before).max())


def test_uniform_state_stays_uniform():
    p, q, _ = _pif_wave(16, amp=0.0)
    q0 = q.copy()
    _run(p, q, 0.02, 0.005)
    g = p.mbc
    assert np.allclose(q[g:g + 16, g:g + 16], q0[g:g + 16, g:g + 16], rtol=1e-13, atol=1e-13)


def test_shallow_water_against_rk4():
    """RK4-PIF on the smooth shallow-water bump (periodic, g = 1) against the
    10-stage fourth-order Runge-Kutta solution on the same grid, with grid
    and step refined together (dt = 0.16 / mx): the difference of the two
    fourth-order schemes falls at fourth order (3.1, then 4.4 from 16^2 to
    64^2); a wrong stage weight or a stage taken from the wrong state leaves
    an O(h) difference.  (With the grid fixed, RK4-PIF's step-to-step
    differences fall only at first order, for Euler as well: the scheme mixes
    two spatial operators, so its time error is not separable from h.)"""
    T = 0.16
    diffs = []
    for mx in (16, 32, 64):
        out = []
        for kw in (dict(time_stepping_method=TS_PIF), {}):
            p, q, _ = problems.sw_bump(nx=mx, ny=mx, cfl=10.0, time_order=4, **kw)
            dtv = np.array([0.0, T / mx, T / mx])
            rc, t, _ = orc.fin_solve(p, q, 0.0, T, dtv, pif=bool(kw))
            assert rc == 0 and abs(t - T) < 1e-13
            g = p.mbc
            out.append(q[g:g + mx, g:g + mx].copy())
        diffs.append(np.mean(np.abs(out[0] - out[1])))
    orders = [math.log2(diffs[k] / diffs[k + 1]) for k in range(2)]
    print("shallow-water PIF vs RK4 differences", diffs, "orders", orders)
    assert orders[-1] > 3.7
    # flux form on the periodic grid: the totals are conserved
    p, q, _ = problems.sw_bump(nx=32, ny=32, time_order=4, time_stepping_method=TS_PIF)
    g = p.mbc
    before = q[g:g + 32, g:g + 32].sum(axis=(0, 1))
    _run(p, q, 0.05, 0.004)
    after = q[g:g + 32, g:g + 32].sum(axis=(0, 1))
    assert np.all(np.abs(after - before) <= 1e-12 * np.abs(before).max())
