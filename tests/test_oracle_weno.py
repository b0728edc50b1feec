"""Pins the oracle's WENO reconstructions (oracle/finess_oracle.c, restating
lib/WenoReconstruct.cpp:90-141 JS5, :385-438 Z5, :580-616 FD5) against an
independent construction from their mathematical definitions:

* the candidate values u_k are the values at x_{i+1/2} of the quadratics p_k
  whose cell averages match the 3-cell sub-stencils (Jiang & Shu 1996);
* the smoothness indicators are beta_k = sum_{l=1,2} int_{I_i} h^{2l-1}
  (p_k^{(l)})^2 dx (Jiang & Shu 1996, eq. 3.3), evaluated here by exact
  polynomial integration -- the 13/12 and 1/4 of the closed form are not
  typed anywhere in this file;
* the linear weights d = (1/10, 6/10, 3/10) are the unique combination that
  is exact for degree-4 polynomials (solved for here, not typed).
"""
import numpy as np
import numpy.polynomial.polynomial as P
import pytest

import orc
from finess_amd import WENO_FD, WENO_JS, WENO_Z

RNG = np.random.default_rng(7)


def cell_avgs_of_poly(c, centres, h=1.0):
    """Cell averages of polynomial c (coefficients, low first) over cells of width h."""
    ci = P.polyint(c)
    return np.array([(P.polyval(x + h / 2, ci) - P.polyval(x - h / 2, ci)) / h for x in centres])


def quad_with_avgs(avgs, centres):
    """Quadratic whose averages over unit cells at `centres` are `avgs`."""
    A = np.array([cell_avgs_of_poly(np.eye(3)[k], centres) for k in range(3)]).T
    return np.linalg.solve(A, avgs)


def reference_parts(g):
    """Candidates u_k and indicators beta_k of stencil g = u_{i-2..i+2}
    (cells centred at -2..2, unit width, x_{i+1/2} = 0.5)."""
    u, beta = [], []
    for k, cells in enumerate(([-2, -1, 0], [-1, 0, 1], [0, 1, 2])):
        c = quad_with_avgs(np.array([g[j + 2] for j in cells]), np.array(cells, float))
        u.append(P.polyval(0.5, c))
        b = 0.0
        for l in (1, 2):
            d = P.polyder(c, l)
            sq = P.polyint(P.polymul(d, d))
            b += P.polyval(0.5, sq) - P.polyval(-0.5, sq)   # h = 1
        beta.append(b)
    return np.array(u), np.array(beta)


def linear_weights():
    """The d_k making sum d_k u_k exact for all degree-4 polynomials."""
    rows, rhs = [], []
    for deg in range(5):
        c = np.zeros(deg + 1)
        c[deg] = 1.0
        g = cell_avgs_of_poly(c, np.arange(-2, 3, dtype=float))
        u, _ = reference_parts(g)
        rows.append(u)
        rhs.append(P.polyval(0.5, c))
    d, *_ = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)
    return d


def weno(version, g, eps=1e-6, pw=2.0):
    g = np.ascontiguousarray(g, dtype=np.float64)
    return orc.oracle().orc_weno5(version, eps, pw, orc.ptr(g))


def test_linear_weights_are_the_textbook_ones():
    d = linear_weights()
    assert np.allclose(d, [0.1, 0.6, 0.3], atol=1e-12)


@pytest.mark.parametrize("version", [WENO_JS, WENO_FD, WENO_Z])
def test_constant_is_reproduced(version):
    for c in (0.0, 1.0, -3.25, 1e6):
        assert abs(weno(version, [c] * 5) - c) <= 1e-15 * max(1.0, abs(c))


def test_fd5_exact_for_quartics():
    for _ in range(20):
        c = RNG.normal(size=5)
        g = cell_avgs_of_poly(c, np.arange(-2, 3, dtype=float))
        assert abs(weno(WENO_FD, g) - P.polyval(0.5, c)) <= 1e-12 * (1 + np.abs(c).sum())


def test_js5_matches_definition():
    d = linear_weights()
    for eps in (1e-6, 1e-2, 1e-12):
        for _ in range(50):
            g = RNG.normal(size=5) * RNG.choice([1e-3, 1.0, 30.0])
            u, beta = reference_parts(g)
            a = d / (eps + beta) ** 2
            ref = np.dot(a, u) / a.sum()
            w = np.zeros(3)
            orc.oracle().orc_weno5_js_weights(eps, 2.0, orc.ptr(np.ascontiguousarray(g)), orc.ptr(w))
            assert np.allclose(w, a / a.sum(), rtol=1e-10, atol=1e-14)
            assert abs(weno(WENO_JS, g, eps) - ref) <= 1e-11 * (1 + np.abs(g).max())


def test_js5_general_power():
    d = linear_weights()
    for pw in (1.0, 1.5, 3.0):
        g = RNG.normal(size=5)
        u, beta = reference_parts(g)
        a = d / (1e-6 + beta) ** pw
        assert abs(weno(WENO_JS, g, 1e-6, pw) - np.dot(a, u) / a.sum()) <= 1e-11


def test_z5_matches_definition():
    d = linear_weights()
    for _ in range(50):
        g = RNG.normal(size=5)
        u, beta = reference_parts(g)
        tau = abs(beta[0] - beta[2])
        a = d * (1.0 + (tau / (1e-6 + beta)) ** 2)
        assert abs(weno(WENO_Z, g) - np.dot(a, u) / a.sum()) <= 1e-11 * (1 + np.abs(g).max())


def test_eno_property_picks_smooth_substencil():
    # jump between cells i+1 and i+2: only sub-stencil 0 is smooth
    assert abs(weno(WENO_JS, [0, 0, 0, 0, 1])) < 1e-6
    # jump between cells i-2 and i-1: sub-stencil 2 (cells i..i+2) is smooth
    assert abs(weno(WENO_JS, [0, 1, 1, 1, 1]) - 1.0) < 1e-6


def test_smooth_data_weights_approach_linear():
    """For smooth data omega_k - d_k = O(h^2) away from critical points."""
    errs = []
    for h in (0.04, 0.02, 0.01, 0.005):
        x = 0.3 + h * np.arange(-2, 3)
        g = np.array([(np.sin(xx + h / 2) - np.sin(xx - h / 2)) / h for xx in x])  # averages of cos
        w = np.zeros(3)
        orc.oracle().orc_weno5_js_weights(1e-40, 2.0, orc.ptr(g), orc.ptr(w))
        errs.append(np.abs(w - [0.1, 0.6, 0.3]).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > 1.9 and np.all(rates > 1.7), (errs, rates)
