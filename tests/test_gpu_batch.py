"""GPU tests of the batched mode: every problem of a batch follows the path
the single-problem solver takes (the same stage arithmetic up to the
compiler's fused multiply-adds, the same CFL controller decisions), whatever
the tile shape the autotuner picks; the single-problem solver is itself
checked against the oracle elsewhere.  Tolerances: 1e-12 relative L-inf per
field in FP64, 5e-5 in FP32 (the two kernels may fuse different
multiply-adds; FP32 rounding differences grow over the steps)."""
import math

import numpy as np
import pytest

import orc
from finess_amd import PREC_FP32, PREC_FP64, Batch, Solver, problems

pytestmark = pytest.mark.gpu

SIZES = [(64, 96), (96, 64), (128, 128), (64, 64), (192, 96)]


def _members(sizes, prec):
    out = []
    for k, (nx, ny) in enumerate(sizes):
        p, q, _ = problems.sw_bump(nx=nx, ny=ny, A=0.1 + 0.08 * k, cx=0.4 + 0.05 * k, cy=0.6 - 0.04 * k,
                                   s=0.06 + 0.015 * k, precision=prec)
        out.append((p, q))
    return out


def _single(p, q, steps, dt0):
    with Solver(p) as s:
        s.set_state(q)
        st = s.advance_steps(steps, np.array([0.0, dt0, p.max_dt]))
        qs, t = s.get_state()
    return qs, t, st.steps_accepted, st.steps_rejected


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_batch_matches_single_solver(prec):
    mem = _members(SIZES, prec)
    steps = 12
    with Batch(mem[0][0], SIZES) as b:
        dts = []
        for k, (p, q) in enumerate(mem):
            dts.append(p.initial_dt * (1.0 + 0.1 * k))
            b.set_state(k, q, 0.0, dts[-1])
        b.run(steps)
        for k, (p, q) in enumerate(mem):
            qb, tb, _, acc, rej = b.get_state(k)
            qs, ts, acc_s, rej_s = _single(p, q, steps, dts[k])
            err = orc.rel_linf(orc.interior(p, qb), orc.interior(p, qs))
            print(k, SIZES[k], "accepted", acc, acc_s, "err", err)
            assert acc == steps and acc_s == steps and rej == rej_s
            assert abs(tb - ts) <= (1e-15 if prec == PREC_FP64 else 1e-6) * ts  # dt follows FP32 wave speeds
            assert np.all(err <= (1e-12 if prec == PREC_FP64 else 5e-5)), err


def test_batch_rejected_steps_follow_the_controller():
    """A first dt far above the CFL limit is rejected and retried with the
    controller's dt, as in the single-problem solver."""
    mem = _members(SIZES[:3], PREC_FP64)
    with Batch(mem[0][0], SIZES[:3]) as b:
        for k, (p, q) in enumerate(mem):
            b.set_state(k, q, 0.0, 6.0 * p.initial_dt)
        b.run(5)
        for k, (p, q) in enumerate(mem):
            qb, tb, _, acc, rej = b.get_state(k)
            qs, ts, acc_s, rej_s = _single(p, q, 5, 6.0 * p.initial_dt)
            assert rej >= 1 and rej == rej_s and acc == 5
            assert np.all(orc.rel_linf(orc.interior(p, qb), orc.interior(p, qs)) <= 1e-12)


def test_batch_autotune_and_conservation():
    """The autotuner leaves the state alone and any tile shape gives the same
    result; the periodic shallow-water mass of every member is conserved."""
    draws = problems.sw_batch_members(n=48, seed=11)
    sizes = [(d[0], d[1]) for d in draws]
    mem = []
    for nx, ny, A, cx, cy, s in draws:
        p, q, _ = problems.sw_bump(nx=nx, ny=ny, A=A, cx=cx, cy=cy, s=s, precision=PREC_FP32)
        mem.append((p, q))
    results = []
    for tune in (False, True):
        with Batch(mem[0][0], sizes) as b:
            for k, (p, q) in enumerate(mem):
                b.set_state(k, q, 0.0, p.initial_dt)
            if tune:
                ti, tj = b.autotune(3)
                print("autotuned tile", ti, tj)
                assert (ti + 1) * tj <= 256 and ti * (tj + 1) <= 256
            b.run(20)
            results.append([b.get_state(k)[0] for k in range(len(mem))])
    for k, (p, q) in enumerate(mem):
        a, c = orc.interior(p, results[0][k]), orc.interior(p, results[1][k])
        assert np.all(orc.rel_linf(a, c) <= 1e-13)
        m0 = math.fsum(orc.interior(p, q)[..., 0].ravel())
        m1 = math.fsum(a[..., 0].ravel())
        assert abs(m1 - m0) <= 1e-12 * m0, (k, m0, m1)


def _workload_batch(prec, steps):
    """Workload 4: the 512 seeded draws (grids of 64..384 cells a side,
    periodic shallow-water bumps) advanced `steps` accepted steps each in one
    batch; returns the members and their final states, times and counts."""
    draws = problems.sw_batch_members(n=512, seed=7)
    sizes = [(d[0], d[1]) for d in draws]
    mem = []
    for nx, ny, A, cx, cy, s in draws:
        p, q, _ = problems.sw_bump(nx=nx, ny=ny, A=A, cx=cx, cy=cy, s=s, precision=prec)
        mem.append((p, q))
    with Batch(mem[0][0], sizes) as b:
        for k, (p, q) in enumerate(mem):
            b.set_state(k, q, 0.0, p.initial_dt)
        b.autotune(3)
        b.run(steps)
        out = [b.get_state(k) for k in range(len(mem))]
    return mem, out


def _oracle_members(mem):
    """Eight members of different shapes, the largest 256 x 192 cells (the
    oracle's share of the test time)."""
    picked, shapes = [], set()
    for k, (p, _) in enumerate(mem):
        if p.mx * p.my <= 256 * 192 and (p.mx, p.my) not in shapes:
            picked.append(k)
            shapes.add((p.mx, p.my))
        if len(picked) == 8:
            break
    return picked


def _against_oracle(mem, out, k, steps):
    p, q = mem[k]
    qb, tb, _, acc, rej = out[k]
    p64 = p.copy()
    p64.precision = PREC_FP64
    q0 = q.copy()
    dtv = np.array([0.0, p.initial_dt, p.max_dt])
    rc, t0, st = orc.fin_solve(p64, q0, 0.0, tb, dtv)
    assert rc == 0
    # the oracle ends its run at the member's time: the same steps, unless its
    # dt sequence, equal to rounding, leaves a last step of ~1e-17
    assert st.steps_accepted in (steps, steps + 1) and abs(t0 - tb) <= 1e-15
    return orc.rel_linf(orc.interior(p, qb), orc.interior(p, q0)), st


def test_workload_batch_fp64_members_match_the_oracle():
    """512 members x 200 steps in FP64: eight members against the oracle's
    FinSolveRK run to each member's final time, 1e-10 relative per field."""
    steps = 200
    mem, out = _workload_batch(PREC_FP64, steps)
    assert all(o[3] == steps for o in out)
    for k in _oracle_members(mem):
        err, st = _against_oracle(mem, out, k, steps)
        print("member", k, (mem[k][0].mx, mem[k][0].my), "rejected", out[k][4], st.steps_rejected, "rel err", err)
        assert out[k][4] == st.steps_rejected
        assert np.all(err <= 1e-10), err


def test_workload_batch_fp32_properties_and_error():
    """The workload as specified: 512 members x 200 steps with the FP32
    stencil and FP64 accumulation.  Every member makes its 200 accepted
    steps and conserves its mass and both momenta (periodic grids: the flux
    differences telescope; relative defect <= 1e-12 of the field's absolute
    sum); eight members stay within 1e-4 relative L-inf per field of the FP64
    oracle run to the same time (FP32 rounding of the fluxes, 6e-8 per
    evaluation, over 600 stages)."""
    steps = 200
    mem, out = _workload_batch(PREC_FP32, steps)
    worst = 0.0
    for k, (p, q) in enumerate(mem):
        qb, tb, _, acc, rej = out[k]
        assert acc == steps, (k, acc)
        a0, a1 = orc.interior(p, q), orc.interior(p, qb)
        assert np.all(np.isfinite(a1))
        for m in range(3):
            s0, s1 = math.fsum(a0[..., m].ravel()), math.fsum(a1[..., m].ravel())
            scale = math.fsum(np.abs(a0[..., m]).ravel()) + math.fsum(np.abs(a1[..., m]).ravel())
            worst = max(worst, abs(s1 - s0) / scale)
    print("worst relative conservation defect over 512 members x 3 fields", worst)
    assert worst <= 1e-12
    errs = []
    for k in _oracle_members(mem):
        err, _ = _against_oracle(mem, out, k, steps)
        errs.append(err.max())
        print("member", k, (mem[k][0].mx, mem[k][0].my), "FP32 vs FP64 oracle rel err", err)
    assert max(errs) <= 1e-4, errs
