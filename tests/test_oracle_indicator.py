"""Pins the oracle of the troubled-interface indicator (new functionality,
no reference counterpart): BF16 rounding, the MLP gradient against finite
differences, the invariances the features are defined to have, the limits
of the weight-departure label, and the identities of the routed ConstructL."""
import numpy as np
import pytest

import orc
from finess_amd import problems

RNG = np.random.default_rng(21)


def bf16_reference(x):
    """Round-to-nearest-even to 8 significant bits, from the float's integer image."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def test_bf16_rounding():
    x = np.concatenate([RNG.normal(size=1000).astype(np.float32), np.float32([1.0, 1.00390625, 1.01171875, -2.5])])
    got = np.array([orc.oracle().orc_bf16(float(v)) for v in x], np.float32)
    assert np.array_equal(got, bf16_reference(x))
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert orc.oracle().orc_bf16(1.00390625) == 1.0


def bce(z, y):
    z = np.asarray(z, np.float64)
    return np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))


def np_forward(P, X):
    """Float64 restatement of the network for finite differences."""
    P = P.astype(np.float64)
    W1 = P[:768].reshape(24, 32); b1 = P[768:800]
    W2 = P[800:1824].reshape(32, 32); b2 = P[1824:1856]
    w3 = P[1856:1888]; b3 = P[1888]
    h1 = np.maximum(X.astype(np.float64) @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    return h2 @ w3 + b3


def test_mlp_gradient_matches_finite_differences():
    orc.oracle().orc_mlp_set_rounding(0)
    try:
        P = (RNG.normal(size=1889) * 0.3).astype(np.float32)
        X = RNG.normal(size=(40, 24)).astype(np.float32)
        y = (RNG.random(40) > 0.5).astype(np.float32)
        g, loss = orc.mlp_gradient(P, X, y)
        assert abs(loss - bce(np_forward(P, X), y)) < 1e-6
        assert np.allclose(orc.mlp_forward(P, X), np_forward(P, X), atol=1e-5)
        idx = RNG.choice(1889, 80, replace=False)
        for k in list(idx) + [1888, 768 + 5, 1824 + 3]:
            h = 1e-4
            Pp = P.astype(np.float64); Pp[k] += h
            Pm = P.astype(np.float64); Pm[k] -= h
            fd = (bce(np_forward(Pp, X), y) - bce(np_forward(Pm, X), y)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])
    finally:
        orc.oracle().orc_mlp_set_rounding(1)


def _euler_stencil():
    rho = 1.0 + 0.5 * RNG.random(6)
    u = RNG.normal(size=6) * 0.3
    v = RNG.normal(size=6) * 0.3
    pr = 1.0 + RNG.random(6)
    return problems.euler_conserved(rho, u, v, pr), (rho, u, v, pr)


def test_features_are_dimensionless():
    p, _, _ = problems.riemann4(mx=16)
    qs, (rho, u, v, pr) = _euler_stencil()
    f0 = orc.features(p, 1, qs)
    for lam in (1e-3, 7.0, 1e4):                      # same velocities, scaled rho and p
        f1 = orc.features(p, 1, problems.euler_conserved(lam * rho, u, v, lam * pr))
        assert np.allclose(f0, f1, atol=1e-5)
    f2 = orc.features(p, 1, problems.euler_conserved(rho, u, v + 0.7, pr))  # tangential shift
    assert np.allclose(f0, f2, atol=1e-5)
    # the y-face of the transposed stencil sees the same rotated primitives
    f3 = orc.features(p, 2, problems.euler_conserved(rho, v, u, pr))
    assert np.allclose(f0, f3, atol=1e-6)
    assert np.abs(f0).max() <= 1.0 + 1e-6


def test_departure_limits():
    p, _, _ = problems.riemann4(mx=16)
    flat = problems.euler_conserved(np.full(6, 1.2), np.full(6, 0.3), np.full(6, -0.1), np.full(6, 2.0))
    assert orc.departure(p, 1, flat) < 1e-12
    step = problems.euler_conserved(np.array([1, 1, 1, 0.125, 0.125, 0.125]), np.zeros(6), np.zeros(6),
                                    np.array([1, 1, 1, 0.1, 0.1, 0.1]))
    assert orc.departure(p, 1, step) > 0.4


def test_routed_construct_l_identities():
    p, q, _ = problems.riemann4(mx=24, my=20)
    q[..., 0] *= 1 + 0.05 * RNG.random(q.shape[:-1])
    orc.set_bnd(p, q)
    L0, s0 = orc.construct_l(p, q)
    La, sa, na = orc.construct_l_routed(p, q, 0)
    Lb, sb, nb = orc.construct_l_routed(p, q, 2, tau=-1.0)      # every face troubled
    assert np.array_equal(orc.interior(p, La), orc.interior(p, L0)) and np.array_equal(sa, s0)
    assert np.array_equal(orc.interior(p, Lb), orc.interior(p, L0))
    Lc, _, nc = orc.construct_l_routed(p, q, 1)
    Ld, _, nd = orc.construct_l_routed(p, q, 2, tau=2.0)        # no face troubled
    assert nc == 0 and nd == 0 and np.array_equal(orc.interior(p, Lc), orc.interior(p, Ld))
    p2, q2, _ = problems.riemann4(mx=24, my=20)       # piecewise constant: only the jumps are troubled
    orc.set_bnd(p2, q2)
    Le, _, ne = orc.construct_l_routed(p2, q2, 2, tau=0.05)
    assert 0 < ne < na // 3


def test_linear_flux_is_fifth_order():
    """The component-wise linear flux is the FD5 scheme: on a smooth density
    wave its L converges to the exact -div f at fifth order."""
    import math
    errs = []
    for mx in (16, 32, 64):
        p, q, _ = problems.density_wave(mx=mx, u0=1.0, v0=0.5)
        L, _, _ = orc.construct_l_routed(p, q, 1)
        X, Y = problems.cell_centres(p)
        # rho_t = -(u rho_x + v rho_y) for the point values
        ex = -0.2 * 2 * math.pi * np.cos(2 * math.pi * (X + Y)) * 1.5
        errs.append(np.abs(orc.interior(p, L)[..., 0] - orc.interior(p, ex[..., None])[..., 0]).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > 4.7, (errs, rates)


def test_linear_flux_equals_weno_on_linear_data():
    """With u, v, p constant and rho linear in x, both split fluxes are linear
    along the stencil, every smoothness indicator is equal, the JS weights are
    exactly the linear weights, and characteristic WENO5 reduces to the linear
    flux: interior L of the two routes agree and the departure vanishes."""
    p, _, _ = problems.riemann4(mx=24, my=12)
    X, _ = problems.cell_centres(p)
    rho = 1.0 + 0.5 * X
    q = problems.euler_conserved(rho, np.full_like(X, 0.4), np.full_like(X, -0.2), np.full_like(X, 1.3))
    orc.set_bnd(p, q)
    Lw, sw, _ = orc.construct_l_routed(p, q, 0)
    Ll, sl, _ = orc.construct_l_routed(p, q, 1)
    g = p.mbc
    sl_ = (slice(g + 4, g + p.mx - 4), slice(g, g + p.my))   # cells whose faces see no boundary ghost
    assert np.abs(Lw[sl_] - Ll[sl_]).max() <= 1e-12 * np.abs(Lw[sl_]).max()
    assert np.array_equal(sw, sl)
    qs = orc.face_stencil(p, q, 1, 10, 5)
    assert orc.departure(p, 1, qs) < 1e-12


@pytest.mark.parametrize("dir_", [1, 2])
def test_device_indicator_functions_match_oracle(dir_):
    """The stage kernel's feature, label and linear-flux routines, run on the
    host, against the oracle on random stencils."""
    p, _, _ = problems.riemann4(mx=16)
    for _ in range(50):
        qs, _ = _euler_stencil()
        f_o = orc.features(p, dir_, qs)
        assert np.allclose(orc.emu_face_features(p, dir_, qs), f_o, atol=1e-6)
        assert np.allclose(orc.emu_face_features(p, dir_, qs, fp32=True), f_o, atol=2e-5)
        assert abs(orc.emu_face_departure(p, dir_, qs) - orc.departure(p, dir_, qs)) < 1e-12
        fh_o, a_o = orc.face_flux_linear(p, dir_, qs)
        fh_e, a_e = orc.emu_face_flux_linear(p, dir_, qs)
        assert abs(a_e - a_o) <= 1e-14 * a_o
        assert np.allclose(fh_e, fh_o, rtol=1e-13, atol=1e-13 * np.abs(fh_o).max())


def test_routed_stepping_identities():
    """Routed SSP-RK3 stepping: with every face troubled it is the plain
    stepper bit for bit, whether the routing is decided per stage or per step;
    mixed routing decided once per step stays close to deciding per stage."""
    p, q0, _ = problems.riemann4(mx=24, my=20)
    dt = 0.3 / 24 / 2.5
    qa = q0.copy()
    orc.fixed_steps(p, qa, dt, 3)
    for per_step in (0, 1, 2):
        qb = q0.copy()
        orc.fixed_steps_routed(p, qb, dt, 3, -1.0, per_step)
        assert np.array_equal(orc.interior(p, qb), orc.interior(p, qa))
    qc, qd = q0.copy(), q0.copy()
    orc.fixed_steps_routed(p, qc, dt, 3, 0.05, 0)
    orc.fixed_steps_routed(p, qd, dt, 3, 0.05, 1)
    dif = orc.rel_linf(orc.interior(p, qd), orc.interior(p, qc))
    assert 0 < dif.max() < 0.05, dif
