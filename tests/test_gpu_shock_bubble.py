"""The Mach-3 shock-bubble problem (workload 5: [0,4]x[0,1], shock at x = 0.5,
light bubble, +-1% seeded density noise; extrapolation in x, reflecting walls
in y) in its routed forms, at 512 x 128 cells (dx = dy = 1/128):

* exact-departure routing (the label itself decides, FP64) against the
  routed oracle over 5 steps: 1e-10 relative L-inf per field;
* the learned form (trained network, FP32 stencil, FP64 accumulation,
  indicator every 4th step) against the FP64 oracle's plain run to t = 0.05
  (about 75 steps): the L1 density difference is stated relative to the
  oracle's L1 density norm and bounded by 5e-3, with the plain FP64 GPU
  run matching the oracle to 1e-10 on the same steps."""
import numpy as np
import pytest

import orc
from finess_amd import PREC_FP32, ROUTE_LEARNED, ROUTE_ORACLE, Solver, problems
from finess_amd.indicator import train_indicator

pytestmark = pytest.mark.gpu

MX, MY = 512, 128
T_END = 0.05


def test_routed_exact_shock_bubble_matches_oracle():
    p, q, _ = problems.shock_bubble(mx=MX, my=MY, routing=ROUTE_ORACLE, route_tau=0.05, route_per_step=1)
    dt = 0.4 / MY / 4.6
    q_ref = q.copy()
    orc.fixed_steps_routed(p, q_ref, dt, 5, 0.05, 1)
    with Solver(p) as s:
        s.set_state(q)
        s.run_fixed_steps(dt, 5)
        q_gpu, _ = s.get_state()
        nt, nf = s.route_stats()
    err = orc.rel_linf(orc.interior(p, q_gpu), orc.interior(p, q_ref))
    print("shock-bubble exact routing: troubled", nt, "of", nf, "rel err", err)
    assert 0 < nt < nf
    assert np.all(err <= 1e-10), err


def _run(p, q, tend):
    with Solver(p) as s:
        s.set_state(q)
        st = s.advance(tend, np.array([0.0, p.initial_dt, p.max_dt]))
        qn, _ = s.get_state()
        share = None
        if p.routing != 0:
            nt, nf = s.route_stats()
            share = nt / max(nf, 1)
    return qn, st, share


def test_learned_fp32_shock_bubble_against_the_oracle():
    ind = train_indicator(mx=512, holdout=(384, 0.12))
    p0, q, _ = problems.shock_bubble(mx=MX, my=MY)
    q_orc = q.copy()
    rc, t0, st0 = orc.fin_solve(p0, q_orc, 0.0, T_END, np.array([0.0, p0.initial_dt, p0.max_dt]))
    assert rc == 0
    q_plain, st_plain, _ = _run(p0, q.copy(), T_END)
    e_plain = orc.rel_linf(orc.interior(p0, q_plain), orc.interior(p0, q_orc))
    pl, _, _ = problems.shock_bubble(mx=MX, my=MY, precision=PREC_FP32, routing=ROUTE_LEARNED, route_per_step=4)
    with Solver(pl) as s:
        s.set_state(q.copy())
        ind.install(s)
        st_l = s.advance(T_END, np.array([0.0, pl.initial_dt, pl.max_dt]))
        q_l, t_l = s.get_state()
        nt, nf = s.route_stats()
    g = p0.mbc
    rho_o = q_orc[g:g + MX, g:g + MY, 0]
    rho_l = q_l[g:g + MX, g:g + MY, 0]
    assert np.all(np.isfinite(q_l))
    rel_l1 = np.abs(rho_l - rho_o).sum() / np.abs(rho_o).sum()
    print("shock-bubble %dx%d to t=%g: oracle steps %d, plain GPU steps %d rel err %s; learned steps %d, "
          "characteristic share %.3f, L1 density difference from the oracle %.3e (relative)" %
          (MX, MY, T_END, st0.steps_accepted, st_plain.steps_accepted, e_plain, st_l.steps_accepted,
           nt / max(nf, 1), rel_l1))
    assert st_plain.steps_accepted == st0.steps_accepted and np.all(e_plain <= 1e-10)
    assert abs(t_l - T_END) <= 1e-15
    assert rel_l1 <= 5e-3
