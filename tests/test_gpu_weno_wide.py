"""GPU parity of the wider WENO orders (face-list kernels, stage2d_wide.hip)
against the CPU oracle: JS 7/9/11, Z 7/9, linear 7/9 and first order
(lib/WenoReconstruct.cpp:39-73).  Tolerance, relative L-inf per field: 1e-10
(the FP64 bar) or ten times the oracle's own spread under 1e-15 relative
input noise, whichever is larger, for ConstructL (face speeds to 1e-13) and
after the stated Runge-Kutta steps (with the same accepted-step count).  The
operator bar is not the 1e-12 of the fifth-order kernel: the wide smoothness
indicators are integer quadratic forms whose terms cancel by up to eight
digits, and the device sums them in another order (with fused multiply-adds)
than the oracle, so the weights differ in their last digits by more."""
TOL = 1e-10
import numpy as np
import pytest

import orc
from finess_amd import (BC_EXTRAP, BC_PERIODIC, BC_REFLECT, PREC_FP32, ROUTE_ORACLE, TS_LXW, WENO_FD, WENO_JS,
                        WENO_Z, FinessError, Solver, problems)

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2718)
CASES = [(WENO_JS, 7), (WENO_JS, 9), (WENO_JS, 11), (WENO_Z, 7), (WENO_Z, 9), (WENO_FD, 7), (WENO_FD, 9),
         (WENO_JS, 1)]


def _mbc(order):
    return max(3, (order + 1) // 2)


def _l_tol(p, q, L0):
    """max(TOL, ten times the oracle's own ConstructL spread under 1e-15
    relative input noise)."""
    noise = np.random.default_rng(98)
    spread = 0.0
    for _ in range(2):
        qn = q * (1.0 + 1e-15 * (noise.random(q.shape) - 0.5))
        Ln, _ = orc.construct_l(p, qn)
        spread = max(spread, np.max(orc.rel_linf(orc.interior(p, Ln), orc.interior(p, L0))))
    return max(TOL, 10.0 * spread), spread


def _perturbed_riemann(mx, my, version, order, **kw):
    p, q, _ = problems.riemann4(mx=mx, my=my, space_order=order, weno_version=version, mbc=_mbc(order), **kw)
    q = q.copy()
    q[..., 0] *= 1.0 + 0.02 * RNG.random(q.shape[:-1])
    return p, q


@pytest.mark.parametrize("version,order", CASES)
def test_construct_l_wide_euler(version, order):
    print("case", version, order, flush=True)
    p, q = _perturbed_riemann(37, 30, version, order)
    orc.set_bnd(p, q)
    L0, s0 = orc.construct_l(p, q)
    with Solver(p) as s:
        L1, s1 = s.construct_l(q)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    tol, spread = _l_tol(p, q, L0)
    print("construct_l rel err", err, "oracle spread", spread)
    assert np.all(err <= tol), err
    assert np.max(np.abs(s1 - s0)) <= 1e-13 * max(1.0, np.max(s0))


@pytest.mark.parametrize("order", [7, 11])
def test_construct_l_wide_euler_nonzero_rho_w(order):
    """rho w != 0 in the first half of the rows (36 of 70): the face tiles
    that see it reconstruct every characteristic field, the others skip the
    zero one; both against the oracle."""
    p, q = _perturbed_riemann(70, 48, WENO_JS, order)
    n = q.shape[0] // 2
    w = 0.3 * np.sin(np.linspace(0.0, 3.0, q.shape[1]))[None, :] * np.ones((n, 1))
    q[:n, :, 3] = q[:n, :, 0] * w
    q[:n, :, 4] += 0.5 * q[:n, :, 0] * w * w
    orc.set_bnd(p, q)
    L0, s0 = orc.construct_l(p, q)
    with Solver(p) as s:
        L1, s1 = s.construct_l(q)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    tol, spread = _l_tol(p, q, L0)
    print("order", order, "construct_l rel err", err, "oracle spread", spread)
    assert np.all(err <= tol), err


@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_REFLECT, BC_EXTRAP])
def test_construct_l_order11_boundaries(bc):
    """Eleventh order reaches six cells into the ghosts on every side."""
    p, q = _perturbed_riemann(24, 26, WENO_JS, 11, bc=[bc] * 4)
    # momenta in both directions (the reflecting walls flip them), small
    # enough to keep the pressure positive everywhere
    q[..., 1] += 0.01 * q[..., 0] * (RNG.random(q.shape[:-1]) - 0.5)
    q[..., 2] -= 0.01 * q[..., 0] * (RNG.random(q.shape[:-1]) - 0.5)
    orc.set_bnd(p, q)
    L0, _ = orc.construct_l(p, q)
    with Solver(p) as s:
        L1, _ = s.construct_l(q)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    tol, spread = _l_tol(p, q, L0)
    print("bc", bc, "rel err", err, "oracle spread", spread)
    assert np.all(err <= tol), err


@pytest.mark.parametrize("order,bc", [(11, BC_PERIODIC), (11, BC_REFLECT), (9, BC_EXTRAP), (7, BC_PERIODIC)])
def test_construct_l_smallest_grid(order, bc):
    """The smallest grid an order accepts (mx = max(r, 5), my = mx + 1; 2D
    grids need 5 cells a side): at r = 5, 6 the stencils reach the boundary
    map's whole range [-n, 2n-1]."""
    r = (order + 1) // 2
    n = max(r, 5)
    p, q = _perturbed_riemann(n, n + 1, WENO_JS, order, bc=[bc] * 4)
    orc.set_bnd(p, q)
    L0, s0 = orc.construct_l(p, q)
    with Solver(p) as s:
        L1, s1 = s.construct_l(q)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    tol, spread = _l_tol(p, q, L0)
    print("order", order, "bc", bc, "rel err", err, "oracle spread", spread)
    assert np.all(err <= tol), err
    assert np.max(np.abs(s1 - s0)) <= 1e-13 * max(1.0, np.max(s0))


@pytest.mark.parametrize("version,order", [(WENO_JS, 7), (WENO_Z, 9)])
def test_construct_l_wide_shallow_water(version, order):
    p, q, _ = problems.sw_bump(nx=33, ny=28, space_order=order, weno_version=version, mbc=_mbc(order))
    q[..., 1] = 0.1 * RNG.random(q.shape[:-1])
    q[..., 2] = 0.1 * RNG.random(q.shape[:-1])
    orc.set_bnd(p, q)
    L0, _ = orc.construct_l(p, q)
    with Solver(p) as s:
        L1, _ = s.construct_l(q)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    tol, spread = _l_tol(p, q, L0)
    print("sw rel err", err, "oracle spread", spread)
    assert np.all(err <= tol), err


@pytest.mark.parametrize("version,order,time_order", [(WENO_JS, 7, 3), (WENO_Z, 9, 3), (WENO_JS, 11, 2),
                                                      (WENO_FD, 7, 4), (WENO_FD, 9, 3), (WENO_JS, 1, 1)])
def test_fin_solve_rk_wide(version, order, time_order):
    """FinSolveRK from t = 0 to 0.01 at 64^2 (the device controller for time
    orders up to 3, the host controller above): the four-quadrant problem for
    the nonlinear weights, the smooth density wave for the linear ones (which
    have no shock capturing)."""
    print("case", version, order, time_order, flush=True)
    make = problems.density_wave if version == WENO_FD else problems.riemann4
    p, q, _ = make(mx=64, space_order=order, weno_version=version, mbc=_mbc(order), time_order=time_order)
    q0 = q.copy()
    dtv0 = np.array([0.0, p.initial_dt, p.max_dt])
    rc, t0, st0 = orc.fin_solve(p, q0, 0.0, 0.01, dtv0)
    assert rc == 0
    # the oracle's own spread under input noise at the rounding level: the
    # wide indicators' cancellation amplifies it over the steps (JS11 on the
    # four-quadrant problem: 1e-15 -> ~1e-9 in four steps)
    spread = 0.0
    noise = np.random.default_rng(99)
    for _ in range(2):
        qn = q * (1.0 + 1e-15 * (noise.random(q.shape) - 0.5))
        dtn = np.array([0.0, p.initial_dt, p.max_dt])
        assert orc.fin_solve(p, qn, 0.0, 0.01, dtn)[0] == 0
        spread = max(spread, np.max(orc.rel_linf(orc.interior(p, qn), orc.interior(p, q0))))
    tol = max(TOL, 10.0 * spread)
    q1 = q.copy()
    dtv1 = np.array([0.0, p.initial_dt, p.max_dt])
    with Solver(p) as s:
        t1 = s.fin_solve_rk(q1, 0.0, 0.01, dtv1)
    assert abs(t1 - t0) <= 1e-14
    assert abs(dtv1[1] - dtv0[1]) <= tol * dtv0[1]
    err = orc.rel_linf(orc.interior(p, q1), orc.interior(p, q0))
    print("steps", st0.steps_accepted, "rel err", err, "oracle spread", spread, "tol", tol)
    assert np.all(err <= tol), err


@pytest.mark.parametrize("version,order", CASES)
def test_construct_l_wide_1d(version, order):
    """1D advection (the face kernel of stage1d.hip with a 2r-cell stencil):
    ConstructL on a rough periodic state."""
    r = (order + 1) // 2
    p, q, aux = problems.advection_1d(mx=53, space_order=order, weno_version=version, mbc=max(3, r))
    q = q + 0.3 * RNG.random(q.shape)
    orc.set_bnd(p, q, aux)
    L0, _ = orc.construct_l(p, q, aux)
    with Solver(p) as s:
        L1, _ = s.construct_l(q, aux=aux)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    print("1d construct_l rel err", err)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("version,order", [(WENO_JS, 11), (WENO_Z, 9), (WENO_FD, 7)])
def test_fin_solve_rk_wide_1d(version, order):
    """The 1D advection run to t = 1 (500 steps at CFL 0.4) against the oracle
    and against the exact shifted sine."""
    r = (order + 1) // 2
    p, q, aux = problems.advection_1d(mx=200, space_order=order, weno_version=version, mbc=max(3, r),
                                      initial_dt=0.002)
    q0 = q.copy()
    dtv0 = np.array([0.0, p.initial_dt, p.max_dt])
    rc, t0, st0 = orc.fin_solve(p, q0, 0.0, 1.0, dtv0, aux=aux)
    assert rc == 0
    q1 = q.copy()
    dtv1 = np.array([0.0, p.initial_dt, p.max_dt])
    with Solver(p) as s:
        t1 = s.fin_solve_rk(q1, 0.0, 1.0, dtv1, aux=aux)
    assert abs(t1 - t0) <= 1e-14
    err = orc.rel_linf(orc.interior(p, q1), orc.interior(p, q0))
    x = problems.cell_centres(p)
    l1 = np.mean(np.abs(orc.interior(p, q1)[:, 0] - orc.interior(p, np.sin(2 * np.pi * (x - 1.0)))))
    print("steps", st0.steps_accepted, "rel err", err, "L1 vs exact", l1)
    assert np.all(err <= TOL), err
    assert l1 < 2e-6


def test_wide_orders_rejected_where_the_device_has_no_path():
    """Orders other than 5 run the 2D Runge-Kutta / ConstructL path in FP64
    only; the rest is refused at configuration time, as are the pairs
    GetWenoReconstruct lacks."""
    bad = [dict(space_order=7, precision=PREC_FP32), dict(space_order=7, time_stepping_method=TS_LXW, mbc=5),
           dict(space_order=7, routing=ROUTE_ORACLE), dict(space_order=9, mbc=4),
           dict(space_order=11, weno_version=WENO_Z, mbc=6), dict(space_order=3)]
    for kw in bad:
        with pytest.raises(FinessError):
            problems.riemann4(mx=32, **kw)


@pytest.mark.parametrize("version,order", [(WENO_JS, 5), (WENO_JS, 11), (WENO_Z, 9)])
def test_construct_l_1d_large_grid(version, order):
    """Above the single-workgroup limit (1536 cells for WENO5, 512 for
    WENO11) the 1D stage runs as two launches (one thread per face, then one
    per cell) instead of one workgroup."""
    r = (order + 1) // 2
    p, q, aux = problems.advection_1d(mx=9001, space_order=order, weno_version=version, mbc=max(3, r))
    q = q + 0.3 * RNG.random(q.shape)
    orc.set_bnd(p, q, aux)
    L0, s0 = orc.construct_l(p, q, aux)
    with Solver(p) as s:
        L1, s1 = s.construct_l(q, aux=aux)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    print("1d large grid rel err", err)
    assert np.all(err <= TOL), err
    assert np.max(np.abs(s1 - s0)) <= 1e-13 * max(1.0, np.max(s0))
