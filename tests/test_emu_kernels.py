"""The device kernels' phase functions (finess_amd/csrc/*_body.h), run on the
CPU one workgroup and one thread at a time, against the oracle.  Covers the
tiling (partial tiles, non-square grids), the halo built through the
boundary map for every boundary kind, both physics, all WENO variants and
the Lax-Wendroff fluxes -- without a GPU."""
import numpy as np
import pytest

import orc
from finess_amd import BC_EXTRAP, BC_PERIODIC, BC_REFLECT, PREC_FP32, WENO_FD, WENO_Z, problems

RNG = np.random.default_rng(5)


def _cl(p, q, tol=1e-13):
    orc.set_bnd(p, q)
    L0, s0 = orc.construct_l(p, q)
    L1, s1 = orc.emu_construct_l(p, q)
    err = orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0))
    assert np.all(err <= tol), err
    assert np.abs(s1 - s0).max() <= 1e-14 * max(1.0, s0.max())


@pytest.mark.parametrize("shape", [(8, 32), (37, 70), (41, 33)])
def test_construct_l_euler(shape):
    p, q, _ = problems.riemann4(mx=shape[0], my=shape[1])
    q[..., 0] *= 1 + 0.05 * RNG.random(q.shape[:-1])
    _cl(p, q)


@pytest.mark.parametrize("kw", [dict(weno_version=WENO_Z), dict(weno_version=WENO_FD), dict(power_param=1.5)])
def test_construct_l_euler_variants(kw):
    p, q, _ = problems.riemann4(mx=24, my=40, **kw)
    _cl(p, q)


def test_construct_l_sw_and_1d():
    p, q, _ = problems.sw_bump(nx=45, ny=29)
    q[..., 1] = 0.1 * RNG.random(q.shape[:-1])
    q[..., 2] = 0.1 * RNG.random(q.shape[:-1])
    _cl(p, q, 1e-13)
    p, q, _ = problems.advection_1d(mx=50)
    _cl(p, q)


def test_construct_l_fp32():
    p, q, _ = problems.riemann4(mx=40, precision=PREC_FP32)
    orc.set_bnd(p, q)
    L0, _ = orc.construct_l(p, q)
    L1, _ = orc.emu_construct_l(p, q)
    assert np.all(orc.rel_linf(orc.interior(p, L1), orc.interior(p, L0)) <= 1e-5)


@pytest.mark.parametrize("bc", [[BC_EXTRAP] * 4, [BC_PERIODIC] * 4, [BC_EXTRAP, BC_PERIODIC, BC_REFLECT, BC_REFLECT],
                                [BC_REFLECT, BC_REFLECT, BC_PERIODIC, BC_EXTRAP]])
def test_steps_through_boundary_map(bc):
    p, q, _ = problems.riemann4(mx=37, my=45, bc=bc)
    q[..., 0] *= 1 + 0.02 * RNG.random(q.shape[:-1])
    q0 = q.copy()
    orc.fixed_steps(p, q0, 1e-3, 3)
    orc.emu_fixed_steps(p, q, 1e-3, 3)
    err = orc.rel_linf(orc.interior(p, q), orc.interior(p, q0))
    assert np.all(err <= 1e-12), err


def test_steps_shallow_water_and_lxw():
    p, q, _ = problems.sw_bump(nx=40, ny=24)
    q0 = q.copy()
    orc.fixed_steps(p, q0, 2e-3, 3)
    orc.emu_fixed_steps(p, q, 2e-3, 3)
    assert np.all(orc.rel_linf(orc.interior(p, q), orc.interior(p, q0)) <= 1e-12)
    p, q, _ = problems.kelvin_helmholtz(mx=24, my=40)
    q0 = q.copy()
    orc.fixed_steps(p, q0, 2e-3, 3)
    orc.emu_fixed_steps(p, q, 2e-3, 3)
    assert np.all(orc.rel_linf(orc.interior(p, q), orc.interior(p, q0)) <= 1e-12)


def test_integrated_r():
    p, q, _ = problems.kelvin_helmholtz(mx=24, my=40)
    q[..., 0] *= 1 + 0.05 * RNG.random(q.shape[:-1])
    orc.set_bnd(p, q)
    F0, G0 = orc.integrated_r(p, 0.004, q)
    F1, G1 = orc.emu_integrated_r(p, 0.004, q)
    g = p.mbc
    sl = (slice(g - 3, g + p.mx + 3), slice(g - 3, g + p.my + 3))
    assert np.all(orc.rel_linf(F1[sl], F0[sl]) <= 1e-13)
    assert np.all(orc.rel_linf(G1[sl], G0[sl]) <= 1e-13)
