"""GPU parity of the host SetBndValues path (FINESS_BC_HOST: the app's
boundary rule runs on the host before every stage, as lib/2d/FinSolveRK.cpp:
171-206 calls it, and the stage reads a ghosted copy of its source):

* a host callback that restates the device's reflecting and extrapolation
  rules gives bit for bit the run of those device rules (controlled SSP-RK3
  and fixed steps at orders 1-3; the FP32 stencil form, global alpha and
  seventh-order WENO with mbc = 4 at order 3, and the routed stages with the
  exact-departure and the learned FP32 indicator): the callback's ghosts reach
  the stencils unchanged, on every side, layer and field;
* a time-dependent inflow rule of the kind of apps/2d/euler/mach_reflection/
  SetBndValues.cpp:14-146 (fixed inflow on the left, a reflecting wall right
  of x0 on the bottom, a shock-tracking top), stepped against the oracle's
  ConstructL with the same rule filling the ghosts at the stage times
  UpdateSoln forms (lib/2d/UpdateSoln.cpp:20-27), to 1e-10;
* the host checks: a rule that reads deeper than mbc cells, or a callback that
  fails, stops the run with FINESS_ERR_CONFIG, as does a missing callback."""
import numpy as np
import pytest

import orc
from finess_amd import (BC_EXTRAP, BC_HOST, BC_REFLECT, ERR_CONFIG, PREC_FP32, ROUTE_LEARNED, ROUTE_ORACLE, FinessError,
                        Mlp, Solver, problems)

pytestmark = pytest.mark.gpu


def _cons(rho, u, v, pr, gamma):
    return np.array([rho, rho * u, rho * v, 0.0, pr / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


def _state(p, seed):
    X, Y = problems.cell_centres(p)
    rng = np.random.default_rng(seed)
    ph = rng.random(4) * 6.0
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y + ph[1])
    u = 0.4 * np.sin(2 * np.pi * Y + ph[2])
    v = -0.3 * np.cos(2 * np.pi * X + ph[3])
    pr = 1.0 + 0.2 * np.cos(2 * np.pi * (X + Y))
    return np.ascontiguousarray(problems.euler_conserved(rho, u, v, pr))


def _device_rules(q, g, rules):
    """Edge ghosts (corners untouched) by the device's rules, restated on the
    host: rules[side] is BC_EXTRAP or BC_REFLECT (left, right, bottom, top)."""
    mx, my = q.shape[0] - 2 * g, q.shape[1] - 2 * g
    I, J = slice(g, g + mx), slice(g, g + my)
    for l in range(g):
        for side, (dst, src, nrm) in enumerate([((l, J), (2 * g - 1 - l, J), 1),
                                                  ((g + mx + l, J), (g + mx - 1 - l, J), 1),
                                                  ((I, l), (I, 2 * g - 1 - l), 2),
                                                  ((I, g + my + l), (I, g + my - 1 - l), 2)]):
            if rules[side] == BC_REFLECT:
                q[dst] = q[src]
                q[dst + (nrm,)] = -q[src + (nrm,)]
            else:
                edge = [(g, J), (g + mx - 1, J), (I, g), (I, g + my - 1)][side]
                q[dst] = q[edge]


RULES = [BC_REFLECT, BC_EXTRAP, BC_EXTRAP, BC_REFLECT]


@pytest.mark.parametrize("order,kw", [(1, {}), (2, {}), (3, {}), (3, dict(precision=PREC_FP32)),
                                      (3, dict(global_alpha=1)), (3, dict(space_order=7, mbc=4))])
def test_host_rule_equals_device_rule_fixed_steps(order, kw):
    p, _, _ = problems.riemann4(mx=40, my=36, bc=RULES, time_order=order, **kw)
    q = _state(p, order)
    ph, _, _ = problems.riemann4(mx=40, my=36, bc=[BC_HOST] * 4, time_order=order, **kw)
    calls = []

    def rule(t, qq):
        calls.append(t)
        _device_rules(qq, p.mbc, RULES)

    dt = 0.4 * p.dx / 2.5
    with Solver(p) as s:
        s.set_state(q, t=0.0)
        s.run_fixed_steps(dt, 5)
        q0, t0 = s.get_state()
    with Solver(ph) as s:
        s.set_bnd_callback(rule)
        s.set_state(q, t=0.0)
        s.run_fixed_steps(dt, 5)
        q1, t1 = s.get_state()
    assert len(calls) == 5 * order
    assert np.array_equal(orc.interior(p, q1), orc.interior(p, q0))
    assert t1 == t0


@pytest.mark.parametrize("kw", [dict(routing=ROUTE_ORACLE, route_tau=0.05, route_per_step=2),
                                dict(routing=ROUTE_LEARNED, precision=PREC_FP32, route_per_step=4)])
def test_host_rule_equals_device_rule_routed(kw):
    """The routed stages (indicator record, troubled-face pass, lookup) read
    the ghosted source like the plain stage: the same bits as the device rules."""
    p, q, _ = problems.riemann4(mx=64, my=72, bc=RULES, **kw)
    ph, _, _ = problems.riemann4(mx=64, my=72, bc=[BC_HOST] * 4, **kw)
    net = None
    if kw["routing"] == ROUTE_LEARNED:
        # an untrained network flags faces at random: on smooth data both
        # fluxes it chooses between stay stable
        m = Mlp(seed=5, capacity=1024)
        net = m.params
        m.close()
        q = _state(p, 7)
    dt = 0.3 / 72 / 2.5
    out = []
    runs = [(p, None), (ph, lambda t, qq: _device_rules(qq, p.mbc, RULES))]
    if kw["routing"] == ROUTE_LEARNED:
        runs.append((p, None))
    for pp, cb in runs:
        with Solver(pp) as s:
            if net is not None:
                s.set_indicator(net, 0.0)
            if cb:
                s.set_bnd_callback(cb)
            s.set_state(q, t=0.0)
            print("run with", "host" if cb else "device", "rules", flush=True)
            s.run_fixed_steps(dt, 6)
            out.append((s.get_state()[0], s.route_stats()))
    print("routed", kw["routing"], "troubled / faces", out[0][1], out[1][1])
    if kw["routing"] == ROUTE_ORACLE:
        assert 0 < out[0][1][0] < out[0][1][1]
    assert out[0][1] == out[1][1]
    if kw["routing"] == ROUTE_LEARNED:
        assert np.array_equal(orc.interior(p, out[2][0]), orc.interior(p, out[0][0]))  # repeatable
    assert np.array_equal(orc.interior(p, out[1][0]), orc.interior(p, out[0][0]))


def test_host_rule_equals_device_rule_controlled():
    p, _, _ = problems.riemann4(mx=48, my=40, bc=RULES, host_controller=1)
    ph, _, _ = problems.riemann4(mx=48, my=40, bc=[BC_HOST, BC_EXTRAP, BC_HOST, BC_HOST])
    q = _state(p, 11)
    res = []
    for pp, cb in ((p, None), (ph, lambda t, qq: _device_rules(qq, p.mbc, RULES))):
        qq = q.copy()
        dtv = np.array([0.0, pp.initial_dt, pp.max_dt])
        with Solver(pp) as s:
            if cb:
                s.set_bnd_callback(cb)
            t = s.fin_solve_rk(qq, 0.0, 0.05, dtv)
        res.append((orc.interior(p, qq), t, dtv[1]))
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1] and res[0][2] == res[1][2]


X0 = 1.0 / 6.0
OSQ3 = 1.0 / np.sqrt(3.0)
POST = (8.0, 8.25 * np.cos(np.pi / 6.0), -8.25 * np.sin(np.pi / 6.0), 116.5)
PRE = (1.4, 0.0, 0.0, 1.0)


def _inflow_rule(p):
    """Ghosts of a double-Mach-type rule: left the post-shock state, right
    extrapolation, bottom the post-shock state left of x0 and a reflecting
    wall right of it, top post-shock behind the moving front x0 + (20 t + 1)/sqrt(3)
    and pre-shock ahead of it."""
    g, mx, my = p.mbc, p.mx, p.my
    xs = p.xlow + (np.arange(-g, mx + g) + 0.5) * p.dx
    post = _cons(*POST, gamma=p.gamma)
    pre = _cons(*PRE, gamma=p.gamma)
    J, I = slice(g, g + my), slice(g, g + mx)

    def rule(t, q):
        q[:g, J] = post
        q[g + mx:, J] = q[g + mx - 1, J][None]
        wall = xs[I] >= X0 - 1e-12
        for l in range(g):
            row = q[I, l]
            row[~wall] = post
            mirror = q[I, 2 * g - 1 - l].copy()
            mirror[:, 2] = -mirror[:, 2]
            row[wall] = mirror[wall]
            q[I, l] = row
            top = np.where((xs[I] < X0 + (20.0 * t + 1.0) * OSQ3)[:, None], post[None], pre[None])
            q[I, g + my + l] = top
    return rule


def test_time_dependent_inflow_against_oracle():
    # the wedge problem's aspect, [0, 3.2] x [0, 1]
    p, _, _ = problems.riemann4(mx=40, my=32, bc=[BC_HOST, BC_EXTRAP, BC_HOST, BC_HOST], xhigh=3.2)
    X, Y = problems.cell_centres(p)
    behind = X < X0 + Y * OSQ3
    q0 = np.where(behind[..., None], _cons(*POST, gamma=p.gamma)[None, None],
                  _cons(*PRE, gamma=p.gamma)[None, None])
    q0 = np.ascontiguousarray(q0)
    rule = _inflow_rule(p)
    dt, nsteps = 0.2 * p.dy / 13.0, 12
    seen = []
    with Solver(p) as s:
        s.set_bnd_callback(lambda t, q: (seen.append(t), rule(t, q)))
        s.set_state(q0, t=0.0)
        s.run_fixed_steps(dt, nsteps)
        qg, tg = s.get_state()
    # the oracle: the same rule fills the ghosts of each stage's source at its
    # UpdateSoln time, then ConstructL and the SSP-RK3 combinations
    po = p.copy()
    po.bc[0] = po.bc[2] = po.bc[3] = BC_EXTRAP
    q = q0.copy()
    t = 0.0
    times = []

    def L_of(qs, ts):
        qs = qs.copy()
        rule(ts, qs)
        times.append(ts)
        L, _ = orc.construct_l(po, qs)
        return L

    for _ in range(nsteps):
        t0 = t
        q1 = 1.0 * q + (1.0 * dt) * L_of(q, t0)
        t1 = 1.0 * t0 + 1.0 * dt
        q2 = 0.75 * q + 0.25 * q1 + (0.25 * dt) * L_of(q1, t1)
        t2 = 0.75 * t0 + 0.25 * t1 + 0.25 * dt
        q = (2.0 / 3.0) * q2 + (1.0 / 3.0) * q + ((2.0 / 3.0) * dt) * L_of(q2, t2)
        t = t0 + dt
    assert seen == times
    err = orc.rel_linf(orc.interior(p, qg), orc.interior(p, q))
    print("time-dependent inflow: rel err", err)
    assert np.all(err <= 1e-10), err
    # the front crossed a cell centre: the top ghosts changed during the run
    xs = p.xlow + (np.arange(p.mx) + 0.5) * p.dx
    behind = [int(np.sum(xs < X0 + (20.0 * tt + 1.0) * OSQ3)) for tt in (times[0], times[-1])]
    assert behind[1] > behind[0], behind


def test_host_boundary_errors():
    p, _, _ = problems.riemann4(mx=24, my=20, bc=[BC_EXTRAP, BC_EXTRAP, BC_EXTRAP, BC_HOST])
    q = _state(p, 3)
    dt = 0.4 * p.dx / 2.5
    g = p.mbc

    def deep(t, qq):  # copies a cell 2 mbc rows in: NaN in the frame it is given
        qq[:, g + p.my:] = qq[:, g + p.my - 2 * g - 1][:, None]

    def broken(t, qq):
        raise RuntimeError("callback failure")

    with Solver(p) as s:
        s.set_state(q, t=0.0)
        with pytest.raises(FinessError) as e:
            s.run_fixed_steps(dt, 1)
        assert e.value.code == ERR_CONFIG
        for cb in (deep, broken):
            s.set_bnd_callback(cb)
            s.set_state(q, t=0.0)
            with pytest.raises(FinessError) as e:
                s.run_fixed_steps(dt, 1)
            assert e.value.code == ERR_CONFIG, cb
        s.set_bnd_callback(lambda t, qq: _device_rules(qq, g, [BC_EXTRAP] * 4))
        s.set_state(q, t=0.0)
        s.run_fixed_steps(dt, 2)
