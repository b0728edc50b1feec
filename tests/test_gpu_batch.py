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
