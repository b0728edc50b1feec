"""The device physics evaluated on the host (finess_physics_eig,
finess_physics_wave_speed, finess_physics_source: the code the kernels are
compiled from, built into libfiness_hip.so) against the pinned oracle's
callback restatements.  These are the reference values of the drop-in
layer's callback probe, so they must be the reference's callbacks: the
eigenvector matrices of ProjectLeftEig / ProjectRightEig applied to the
identity (normalisation included), SetWaveSpd's max(|s1|, |s2|) for unequal
states (the shallow-water h_l quirk on and off), and SourceTermFunc's forms.
CPU only: no device call."""
import ctypes

import numpy as np
import pytest

import orc
from finess_amd import (PHYS_ADVECTION, PHYS_EULER, PHYS_SHALLOW_WATER, SRC_DAMPING, SRC_SW_BATHYMETRY, SRC_ZERO,
                        default_params, load_library)

RNG = np.random.default_rng(8)
PD = orc.ptr


def _proj(p, fn, ixy, qa, v):
    v = np.ascontiguousarray(v.reshape(p.meqn, -1))
    out = np.zeros_like(v)
    fn(ctypes.byref(p), ixy, PD(np.ascontiguousarray(qa)), v.shape[1], PD(v), PD(out))
    return out


def _euler(rho, u, v, pr, g=1.4):
    return np.array([rho, rho * u, rho * v, 0.0, pr / (g - 1) + 0.5 * rho * (u * u + v * v)])


def _sw(h, u, v):
    return np.array([h, h * u, h * v])


def _states(phys, n):
    if phys == PHYS_EULER:
        return [_euler(0.3 + RNG.random(), RNG.normal(), RNG.normal(), 0.2 + RNG.random()) for _ in range(n)]
    return [_sw(0.4 + RNG.random(), RNG.normal(), RNG.normal()) for _ in range(n)]


@pytest.mark.parametrize("phys,meqn", [(PHYS_EULER, 5), (PHYS_SHALLOW_WATER, 3)])
def test_eigenvector_matrices_are_the_callbacks(phys, meqn):
    lib, olib = load_library(), orc.oracle()
    p = default_params(physics=phys, meqn=meqn)
    for qa in _states(phys, 20):
        for ixy in (1, 2):
            Lm, Rm = np.zeros((meqn, meqn)), np.zeros((meqn, meqn))
            assert lib.finess_physics_eig(ctypes.byref(p), ixy, PD(qa), PD(Lm), PD(Rm)) == 0
            L0 = _proj(p, olib.orc_project_left, ixy, qa, np.eye(meqn))
            R0 = _proj(p, olib.orc_project_right, ixy, qa, np.eye(meqn))
            np.testing.assert_allclose(Lm, L0, rtol=1e-12, atol=1e-12 * np.abs(L0).max())
            np.testing.assert_allclose(Rm, R0, rtol=1e-12, atol=1e-12 * np.abs(R0).max())


@pytest.mark.parametrize("phys,meqn,hl", [(PHYS_EULER, 5, 1), (PHYS_SHALLOW_WATER, 3, 1),
                                          (PHYS_SHALLOW_WATER, 3, 0)])
def test_wave_speed_of_unequal_states(phys, meqn, hl):
    lib, olib = load_library(), orc.oracle()
    p = default_params(physics=phys, meqn=meqn, sw_compat_hl=hl)
    aux = np.zeros(1)
    st = _states(phys, 40)
    for ql, qr in zip(st[::2], st[1::2]):
        for ixy in (1, 2):
            s = ctypes.c_double()
            assert lib.finess_physics_wave_speed(ctypes.byref(p), ixy, PD(ql), PD(qr), PD(aux), PD(aux),
                                                 ctypes.byref(s)) == 0
            s1, s2 = ctypes.c_double(), ctypes.c_double()
            olib.orc_wave_speed(ctypes.byref(p), ixy, PD(ql), PD(qr), PD(aux), PD(aux), ctypes.byref(s1),
                                ctypes.byref(s2))
            assert abs(s.value - max(abs(s1.value), abs(s2.value))) <= 1e-13 * s.value


def test_wave_speed_advection_1d():
    lib = load_library()
    p = default_params(ndims=1, physics=PHYS_ADVECTION, meqn=1, maux=1, mx=10, my=1)
    q = np.zeros(1)
    for al, ar in [(1.0, 1.0), (-2.0, 0.5), (0.3, -0.9)]:
        s = ctypes.c_double()
        assert lib.finess_physics_wave_speed(ctypes.byref(p), 1, PD(q), PD(q), PD(np.array([al])),
                                             PD(np.array([ar])), ctypes.byref(s)) == 0
        m = 0.5 * (al + ar)
        assert s.value == max(abs(min(al, m)), abs(max(ar, m)))


def test_source_forms():
    lib = load_library()
    q, aux = np.array([1.3, 0.2, -0.4]), np.array([0.5, -0.7, 0.25])
    for kind, coef, expect in [(SRC_ZERO, 0.0, [0.0, 0.0, 0.0]), (SRC_DAMPING, 2.0, -2.0 * q),
                               (SRC_SW_BATHYMETRY, 0.0, [0.0, -1.3 * -0.7, -1.3 * 0.25])]:
        p = default_params(physics=PHYS_SHALLOW_WATER, meqn=3, maux=3)
        p.source_kind, p.source_coef = kind, coef
        psi = np.zeros(3)
        assert lib.finess_physics_source(ctypes.byref(p), PD(q), PD(aux), PD(psi)) == 0
        np.testing.assert_allclose(psi, expect, rtol=1e-15)
