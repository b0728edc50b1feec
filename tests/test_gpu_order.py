"""Design order of accuracy on the GPU: workload 1 (sin(2 pi x) advected once
around [0, 1], WENO5-JS with LLF splitting, SSP-RK3 at CFL 0.4) at 50, 100,
200 and 400 cells against the exact shifted solution: with the CFL number
fixed the time error O(dt^3) dominates, so the observed order tends to 3
(3.6, 3.2, 3.0 between successive grids); with CFL 0.05 the fifth order in
space shows (5.0 between 25, 50 and 100 cells).  The oracle gives the same
orders on the CPU."""
import numpy as np
import pytest

from finess_amd import Solver, problems

pytestmark = pytest.mark.gpu


def _l1(mx, cfl):
    p, q, aux = problems.advection_1d(mx=mx, cfl=cfl, initial_dt=cfl / mx)
    with Solver(p) as s:
        t = s.fin_solve_rk(q, 0.0, 1.0, np.array([0.0, p.initial_dt, p.max_dt]), aux=aux)
    x = p.xlow + (np.arange(mx) + 0.5) * p.dx
    return float(np.mean(np.abs(q[p.mbc:p.mbc + mx, 0] - np.sin(2 * np.pi * (x - t)))))


def test_observed_order_cfl04():
    errs = [_l1(mx, 0.4) for mx in (50, 100, 200, 400)]
    obs = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("L1 errors", errs, "observed orders", obs)
    assert np.all(obs > 2.8), obs


def test_observed_order_small_cfl():
    errs = [_l1(mx, 0.05) for mx in (25, 50, 100)]
    obs = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("L1 errors", errs, "observed orders", obs)
    assert np.all(obs > 4.7), obs
