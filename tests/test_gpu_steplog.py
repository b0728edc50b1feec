"""ConSoln after every accepted step (finess_set_step_log /
finess_read_step_log): one record (t, sum of dx dy q per field) per accepted
step under the device controller (Runge-Kutta, 2D and 1D) and the host
controller (Lax-Wendroff), none for rejected attempts; the sums are those of the
states (1e-13 relative to the host sums in ConSoln's order) and stay constant
on periodic problems."""
import numpy as np
import pytest

from finess_amd import Solver, problems

pytestmark = pytest.mark.gpu


def _host_sums(p, q):
    g = p.mbc
    if p.ndims == 1:
        return (q[g:g + p.mx] * p.dx).sum(axis=0)
    return (q[g:g + p.mx, g:g + p.my] * (p.dx * p.dy)).sum(axis=(0, 1))


def _run(p, q, tend, dtv=None):
    with Solver(p) as s:
        s.set_step_log(True)
        s.set_state(q)
        dtv = np.array([0.0, p.initial_dt, p.max_dt]) if dtv is None else dtv
        st = s.advance(tend, dtv)
        rec = s.read_step_log()
        q1, t1 = s.get_state()
        assert len(s.read_step_log()) == 0  # read once
    return st, rec, q1, t1


@pytest.mark.parametrize("case", ["rk_device", "lxw_host", "advection_1d"])
def test_records_per_accepted_step(case):
    if case == "rk_device":
        p, q, _ = problems.density_wave(mx=64, amp=0.3)
    elif case == "lxw_host":
        p, q, _ = problems.kelvin_helmholtz(mx=64, my=48)
    else:
        p, q, _ = problems.advection_1d(mx=200)
    s0 = _host_sums(p, q)
    st, rec, q1, t1 = _run(p, q, 0.05)
    print(case, "steps", st.steps_accepted, "records", len(rec))
    assert len(rec) == st.steps_accepted > 10
    assert np.all(np.diff(rec[:, 0]) > 0) and rec[-1, 0] == t1 == 0.05
    scale = np.abs(_host_sums(p, np.abs(q))).max()  # sum of |q|: the sine's total is 0
    assert np.all(np.abs(rec[-1, 1:] - _host_sums(p, q1)) <= 1e-13 * scale)
    # periodic: every record carries the initial totals
    assert np.all(np.abs(rec[:, 1:] - s0) <= 1e-12 * scale)


def test_rejected_attempts_leave_no_record():
    p, q, _ = problems.riemann4(mx=64, initial_dt=0.01)
    st, rec, q1, _ = _run(p, q, 0.01, np.array([0.0, 0.01, 1.0]))
    assert st.steps_rejected >= 1
    assert len(rec) == st.steps_accepted
    assert np.all(np.abs(rec[-1, 1:] - _host_sums(p, q1)) <= 1e-13 * np.abs(_host_sums(p, np.abs(q1))).max())
    # every record is the state at its time: the first one after one accepted step
    assert rec[0, 0] > 0
