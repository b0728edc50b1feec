"""Pins of the oracle's 1D Lax-Wendroff step (lib/1d/FinSolveLxW.cpp,
ConstructIntegratedF.cpp, ConstructLxWL.cpp) and of its 1D Runge-Kutta
schemes of orders 4 and 5 (lib/1d/FinSolveRK.cpp:204-297):

* for linear advection with constant speed a the time-integrated flux is
  F = a (q - (a dt / 2) q_x + (a dt)^2 / 6 q_xx) (the Taylor series in time
  with q_t = -a q_x; the reference's symbolic_tools/weno_diff.py:24-38 has the
  same expression); on smooth periodic data the oracle's F matches it to the
  O(dt dx^4) error of the 5-point differences, and drops the dt^2 term at
  time order 2;
* the whole 1D Lax-Wendroff run converges to the exact translation at
  third order (time order 3) and second order (time order 2);
* with zero transport and the relaxation source psi = -k q, one step of the
  1D fourth- and fifth-order schemes equals exp(-k dt) to O(dt^5), O(dt^6)."""
import math

import numpy as np
import pytest

import orc
from finess_amd import SRC_DAMPING, TS_LXW, problems


def _lxw1d(mx, time_order=3, **kw):
    return problems.advection_1d(mx=mx, mbc=5, time_stepping_method=TS_LXW, time_order=time_order, **kw)


@pytest.mark.parametrize("time_order", [2, 3])
def test_integrated_flux_is_the_taylor_series(time_order):
    a, dt, mx = 0.7, 0.004, 64
    p, q, aux = _lxw1d(mx, time_order)
    aux[:] = a
    orc.set_bnd(p, q, aux)
    F, _ = orc.integrated_r(p, dt, q, aux=aux)
    x = p.xlow + (np.arange(p.mx) + 0.5) * p.dx
    k = 2 * np.pi
    qe, qx, qxx = np.sin(k * x), k * np.cos(k * x), -k * k * np.sin(k * x)
    expect = a * (qe - (a * dt / 2) * qx + (time_order > 2) * (a * dt) ** 2 / 6 * qxx)
    err = np.abs(orc.interior(p, F)[:, 0] - expect).max()
    print("integrated flux error", err)
    assert err <= 1e-6


@pytest.mark.parametrize("time_order,order", [(2, 1.8), (3, 2.7)])
def test_lxw1d_converges(time_order, order):
    tend, errs = 0.5, []
    for mx in (40, 80, 160):
        p, q, aux = _lxw1d(mx, time_order)
        dtv = np.array([0.0, p.initial_dt, p.max_dt])
        rc, t, _ = orc.fin_solve(p, q, 0.0, tend, dtv, aux=aux, lxw=True)
        assert rc == 0 and abs(t - tend) < 1e-14
        x = p.xlow + (np.arange(p.mx) + 0.5) * p.dx
        errs.append(np.mean(np.abs(orc.interior(p, q)[:, 0] - np.sin(2 * np.pi * (x - tend)))))
    obs = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("1D LxW time order", time_order, "errors", errs, "observed", obs)
    assert np.all(obs > order), obs


@pytest.mark.parametrize("time_order,expect", [(4, 5), (5, 6)])
def test_rk45_1d_relaxation_order(time_order, expect):
    k = 1.0
    errs = []
    for dt in (0.2, 0.1):
        p, q, aux = problems.advection_1d(mx=16, time_order=time_order, source_term=1, source_kind=SRC_DAMPING,
                                          source_coef=k)
        aux[:] = 0.0
        q0 = orc.interior(p, q).copy()
        orc.fixed_steps(p, q, dt, 1, aux=aux)
        errs.append(np.abs(orc.interior(p, q) - q0 * math.exp(-k * dt)).max() / np.abs(q0).max())
    obs = math.log2(errs[0] / errs[1])
    print("1D order", time_order, "one-step errors", errs, "observed", obs)
    assert obs > expect - 0.3
