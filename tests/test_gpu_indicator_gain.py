"""What the trained troubled-interface network gains over the method without
it.  Four-quadrant Riemann problem at 256^2 to t = 0.1 (SSP-RK3, the learned
form's FP32 stencil), L1 density error against a 512^2 plain-FP64 run
averaged 2 x 2:

* plain FP64 (characteristic WENO at every face): the accuracy to match;
* learned routing with the trained network: within 5% of the plain error
  while sending only a small share of the faces to the characteristic flux;
* the same network untrained (its initial weights), with the threshold set so
  that it flags the same share of faces: the faces it picks miss the waves,
  and the error grows or the run breaks down (negative pressure behind the
  linear flux at a shock: a non-finite wave speed stops it);
* no network (every face on the linear flux): likewise.
"""
import numpy as np
import pytest

from finess_amd import PREC_FP32, ROUTE_LEARNED, ROUTE_ORACLE, FinessError, Mlp, Solver, problems
from finess_amd.indicator import train_indicator

pytestmark = pytest.mark.gpu

T_END = 0.1


def _density(mx):
    """Interior density of the plain FP64 run at t = T_END."""
    p, q, _ = problems.riemann4(mx=mx)
    with Solver(p) as s:
        s.set_state(q)
        s.advance(T_END, np.array([0.0, p.initial_dt, p.max_dt]))
        qn, _ = s.get_state()
    g = p.mbc
    return qn[g:g + mx, g:g + mx, 0].copy()


def _routed(mx, params, threshold):
    """Interior density at T_END (None if the run breaks down: a non-finite
    wave speed stops it) and the share of faces sent to the characteristic
    flux."""
    p, q, _ = problems.riemann4(mx=mx, precision=PREC_FP32, routing=ROUTE_LEARNED, route_per_step=4)
    with Solver(p) as s:
        s.set_state(q)
        s.set_indicator(params, threshold)
        dtv = np.array([0.0, p.initial_dt, p.max_dt])
        try:
            s.advance(T_END, dtv)
        except FinessError:
            return None, float("nan")
        qn, _ = s.get_state()
        nt, nf = s.route_stats()
    g = p.mbc
    return qn[g:g + mx, g:g + mx, 0].copy(), nt / max(nf, 1)


def test_trained_network_gain():
    mx = 256
    ref = _density(2 * mx)
    ref = ref.reshape(mx, 2, mx, 2).mean(axis=(1, 3))

    def l1(rho):
        # a run that broke down counts as an unbounded error
        return float(np.mean(np.abs(rho - ref))) if rho is not None and np.all(np.isfinite(rho)) else float("inf")

    plain = _density(mx)
    e_plain = l1(plain)
    ind = train_indicator(mx=mx, epochs=4, stride=1, holdout=(192, 0.06))
    rho_l, frac = _routed(mx, ind.params, ind.threshold)
    e_learned = l1(rho_l)
    # the untrained network at the threshold that flags the same share of
    # faces on a harvested snapshot of this problem
    net = Mlp(seed=3, capacity=1 << 21)
    try:
        p, q, _ = problems.riemann4(mx=mx)
        with Solver(p) as s:
            s.set_state(q)
            s.advance(0.05, np.array([0.0, p.initial_dt, p.max_dt]))
            net.harvest(s, stride=1, tau=0.2)
        X, _, _ = net.get_data()
        z = net.logits(X, path=1)
        raw = net.params
    finally:
        net.close()
    thr_u = float(np.quantile(z, 1.0 - frac))
    rho_u, frac_u = _routed(mx, raw, thr_u)
    e_untrained = l1(rho_u)
    # no network: the exact label with tau > 1 flags nothing
    p, q, _ = problems.riemann4(mx=mx, precision=PREC_FP32, routing=ROUTE_ORACLE, route_tau=2.0)
    rho_n = None
    with Solver(p) as s:
        s.set_state(q)
        try:
            s.advance(T_END, np.array([0.0, p.initial_dt, p.max_dt]))
            qn, _ = s.get_state()
            rho_n = qn[p.mbc:p.mbc + mx, p.mbc:p.mbc + mx, 0]
        except FinessError:
            pass
    e_linear = l1(rho_n)
    print("L1 density error: plain %.4e, trained %.4e (%.1f%% faces full WENO), untrained %.4e (%.1f%%), "
          "no network %.4e" % (e_plain, e_learned, 100 * frac, e_untrained, 100 * frac_u, e_linear))
    assert frac < 0.25
    assert e_learned <= 1.05 * e_plain
    assert e_untrained > 1.2 * e_learned
    assert e_linear > 1.2 * e_learned
