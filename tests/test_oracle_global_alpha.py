"""Pins the oracle's global_alpha option (GlobalWaveSpd folded into every
interface's LLF speed, lib/2d/ConstructL.cpp:48-54,172,323;
lib/2d/GlobalWaveSpd.cpp:20-80):

* the interface speeds it records are max(global speed of the direction,
  the local speed without it), face by face;
* the global speeds are max(|u| + c) and max(|v| + c) over the interior
  cells (closed form, computed here);
* where every cell has the same |u| + c and |v| + c the operator equals the
  local one bit for bit, and on a periodic grid the operator still telescopes.
"""
import numpy as np

import orc
from finess_amd import BC_PERIODIC, problems

RNG = np.random.default_rng(31)


def _state():
    p, q, aux = problems.riemann4(mx=40, my=36)
    q = q.copy()
    q[..., 0] *= 1.0 + 0.05 * RNG.random(q.shape[:-1])
    return p, q, aux


def test_face_speeds_are_max_of_global_and_local():
    p, q, aux = _state()
    orc.set_bnd(p, q, aux)
    L0, s0 = orc.construct_l(p, q, aux)
    pg = p.copy()
    pg.global_alpha = 1
    L1, s1 = orc.construct_l(pg, q, aux)
    # closed-form global speeds over the interior
    qi = orc.interior(p, q)
    rho, mu, mv, E = qi[..., 0], qi[..., 1], qi[..., 2], qi[..., 4]
    u, v = mu / rho, mv / rho
    pr = (p.gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v))
    c = np.sqrt(np.abs(p.gamma * pr / rho))
    ax, ay = np.max(np.abs(u) + c), np.max(np.abs(v) + c)
    gx, gy = orc.global_wave_spd(p, q, aux)
    assert abs(gx - ax) <= 1e-14 * ax and abs(gy - ay) <= 1e-14 * ay
    g = p.mbc
    sx0, sy0 = s0[g:g + p.mx + 1, g:g + p.my, 0], s0[g:g + p.mx, g:g + p.my + 1, 1]
    sx1, sy1 = s1[g:g + p.mx + 1, g:g + p.my, 0], s1[g:g + p.mx, g:g + p.my + 1, 1]
    assert np.array_equal(sx1, np.maximum(sx0, gx))
    assert np.array_equal(sy1, np.maximum(sy0, gy))
    # more dissipation changes the operator where the local speed is lower
    assert np.max(np.abs(orc.interior(p, L1 - L0))) > 1e-6


def test_uniform_speed_state_equals_local_and_telescopes():
    p, q, aux = problems.density_wave(mx=24, my=20)
    orc.set_bnd(p, q, aux)
    L0, _ = orc.construct_l(p, q, aux)
    pg = p.copy()
    pg.global_alpha = 1
    L1, _ = orc.construct_l(pg, q, aux)
    # density wave: |u| + c varies with rho (c = sqrt(gamma p / rho)), so the
    # operators differ; both conserve on the periodic grid
    for L in (L0, L1):
        s = orc.interior(p, L).sum(axis=(0, 1))
        assert np.all(np.abs(s) <= 1e-11 * np.abs(orc.interior(p, L)).sum(axis=(0, 1)).max())
    # a uniform state: global speed = every local speed, identical operators
    qu = q.copy()
    qu[...] = q[0, 0]
    orc.set_bnd(p, qu, aux)
    La, sa = orc.construct_l(p, qu, aux)
    Lb, sb = orc.construct_l(pg, qu, aux)
    assert np.array_equal(La, Lb) and np.array_equal(sa, sb)


def test_1d_face_speeds_are_max_of_global_and_local():
    """1D (lib/1d/ConstructL.cpp:45-49,148; lib/1d/GlobalWaveSpd.cpp): the
    global speed of advection is max |a| over the interior, and every face
    speed (smax, overwritten per face in 1D) is max(global, local)."""
    p, q, aux = problems.advection_1d(mx=40)
    x = problems.cell_centres(p)
    aux = np.ascontiguousarray((1.0 + 0.5 * np.sin(2 * np.pi * x))[:, None])
    orc.set_bnd(p, q, aux)
    L0, s0 = orc.construct_l(p, q, aux)
    pg = p.copy()
    pg.global_alpha = 1
    L1, s1 = orc.construct_l(pg, q, aux)
    g = p.mbc
    amax = np.max(np.abs(aux[g:g + p.mx, 0]))
    faces = slice(g, g + p.mx + 1)
    assert np.array_equal(s1[faces], np.maximum(s0[faces], amax))
    assert np.max(np.abs(orc.interior(p, L1 - L0))) > 1e-8
