"""The device-side time-step controller (k_step_control, one launch per step
attempt; in the routed form with records, the first workgroup of the next
step's troubled-face pass) against the host controller: the same accepted and rejected
attempts, the same final dt and bit-identical states.  The controller's
arithmetic is the host's operation for operation and the stage kernels are
the same, so nothing may differ at all.  Cases: frames with tend clipping,
a rejected first attempt, fixed step counts across the 64-step chunks, the
routed form with a record reused over two steps (the routing state is
restored after a rejection), and the error paths (nv limit, dt > max_dt);
in 1D (CFL from the x speeds only, bound 1e10; lib/1d/GetCFL.cpp) frames
with a rejection at each time order, steps across chunks with
global_alpha, and the nv limit."""
import numpy as np
import pytest

from finess_amd import PREC_FP32, ROUTE_ORACLE, FinessError, Solver, problems

pytestmark = pytest.mark.gpu


def _run(p, q, host, action, dtv0=None, aux=None):
    p = p.copy()
    p.host_controller = 1 if host else 0
    with Solver(p) as s:
        s.set_state(q, aux)
        dtv = np.array([0.0, p.initial_dt, p.max_dt] if dtv0 is None else dtv0, dtype=np.float64)
        try:
            st = action(s, dtv)
            err = None
        except FinessError as e:
            st, err = None, e.code
        qo, t = s.get_state()
    return qo, t, dtv, st, err


def _stats(st):
    if st is None:
        return None
    if isinstance(st, list):
        return [_stats(x) for x in st]
    return (st.steps_accepted, st.steps_rejected, st.dt_min, st.dt_max, st.last_cfl, st.t)


def _same(p, q, action, dtv0=None, aux=None):
    qh, th, dh, sh, eh = _run(p, q, True, action, dtv0, aux)
    qd, td, dd, sd, ed = _run(p, q, False, action, dtv0, aux)
    print("host", _stats(sh), eh, "device", _stats(sd), ed)
    assert eh == ed
    assert th == td
    assert np.array_equal(dh, dd)
    assert _stats(sh) == _stats(sd)
    assert np.array_equal(qh, qd)
    return sd, ed


def test_frames_with_clipping():
    p, q, _ = problems.riemann4(mx=96)
    st, _ = _same(p, q, lambda s, dtv: [s.advance(t, dtv) for t in (0.01, 0.02, 0.035)])
    assert st[-1].t == 0.035


def test_rejected_first_attempt():
    p, q, _ = problems.riemann4(mx=64, initial_dt=0.01)
    st, _ = _same(p, q, lambda s, dtv: s.advance(0.01, dtv))
    assert st.steps_rejected >= 1


def test_fixed_step_count_across_chunks():
    p, q, _ = problems.riemann4(mx=64)
    st, _ = _same(p, q, lambda s, dtv: s.advance_steps(70, dtv))
    assert st.steps_accepted == 70


def test_routed_record_reuse_with_rejection():
    p, q, _ = problems.riemann4(mx=64, initial_dt=0.01, routing=ROUTE_ORACLE, route_per_step=2,
                                precision=PREC_FP32)
    st, _ = _same(p, q, lambda s, dtv: s.advance(0.02, dtv))
    assert st.steps_rejected >= 1 and st.steps_accepted > 4


def test_routed_fixed_steps_across_chunks():
    """The routed form with a record every third step over 70 steps (a chunk
    boundary after 64): a step's controller runs in the next step's
    troubled-face pass when that step reuses the record, on its own before an
    indicator launch, and on its own before each chunk's read-back."""
    p, q, _ = problems.riemann4(mx=64, routing=ROUTE_ORACLE, route_per_step=3, precision=PREC_FP32)
    st, _ = _same(p, q, lambda s, dtv: s.advance_steps(70, dtv))
    assert st.steps_accepted == 70


def test_nothing_to_do():
    p, q, _ = problems.riemann4(mx=32)
    st, _ = _same(p, q, lambda s, dtv: s.advance(0.0, dtv))
    assert st.steps_accepted == 0


def test_nv_limit():
    p, q, _ = problems.riemann4(mx=32, nv=3)
    _, err = _same(p, q, lambda s, dtv: s.advance(1.0, dtv))
    assert err == -3


def test_dt_above_max_dt():
    p, q, _ = problems.riemann4(mx=32)
    _, err = _same(p, q, lambda s, dtv: s.advance(1.0, dtv), dtv0=[0.0, 0.5, 0.1])
    assert err == -5


def _advection_varying_speed(mx, **kw):
    """1D advection with aux(1) = 1 + 0.5 sin(2 pi x): the step's CFL moves
    with the state's speed field rather than staying fixed."""
    p, q, _ = problems.advection_1d(mx=mx, **kw)
    x = problems.cell_centres(p)
    aux = np.ascontiguousarray((1.0 + 0.5 * np.sin(2 * np.pi * x))[:, None])
    return p, q, aux


@pytest.mark.parametrize("time_order", [1, 2, 3])
def test_1d_frames_with_rejection(time_order):
    # initial dt at CFL 6: the first attempt is rejected
    p, q, aux = _advection_varying_speed(100, time_order=time_order, initial_dt=0.04)
    st, _ = _same(p, q, lambda s, dtv: [s.advance(t, dtv) for t in (0.05, 0.1, 0.23)], aux=aux)
    assert st[0].steps_rejected >= 1 and st[-1].t == 0.23


def test_1d_fixed_steps_global_alpha():
    p, q, aux = _advection_varying_speed(150, global_alpha=1)
    st, _ = _same(p, q, lambda s, dtv: s.advance_steps(70, dtv), aux=aux)
    assert st.steps_accepted == 70


def test_1d_nv_limit():
    p, q, aux = _advection_varying_speed(64, nv=3)
    _, err = _same(p, q, lambda s, dtv: s.advance(1.0, dtv), aux=aux)
    assert err == -3


@pytest.mark.parametrize("order", [5, 9])
def test_1d_large_grid_two_launch_path(order):
    """Above the single-workgroup limit (1536 cells for WENO5, 768 for
    WENO9) the 1D stage is a face launch and a cell launch, each reading dt
    and the halt word from the control block: a rejected first attempt
    (CFL 6), global_alpha, three chunks of steps."""
    r = (order + 1) // 2
    p, q, aux = _advection_varying_speed(10000, global_alpha=1, space_order=order, mbc=max(3, r), initial_dt=4e-4)
    st, _ = _same(p, q, lambda s, dtv: [s.advance(t, dtv) for t in (0.002, 0.005)], aux=aux)
    assert st[0].steps_rejected >= 1 and st[-1].steps_accepted > 64
