"""The example apps, linked against the C++ drop-in operators, run on the GPU
end to end (parameters.ini -> QinitFunc -> frames of FinSolveRK/FinSolveLxW ->
q%04d.dat) and their frames match the oracle run with the same parameters."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import orc
from finess_amd import TS_PIF, WENO_Z, problems

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APPS = os.path.join(ROOT, "finess_amd", "compat", "apps")


def run_app(name, tmp_path):
    shutil.copy(os.path.join(APPS, name, "parameters.ini"), tmp_path / "parameters.ini")
    print("running app", name, flush=True)
    r = subprocess.run([os.path.join(APPS, name, "finess.exe")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0


def read_frame(path, shape2d=None, meqn=1):
    v = np.loadtxt(path)
    t, q = v[0], v[1:]
    if shape2d is None:
        return t, q.reshape(meqn, -1).T
    mx, my = shape2d
    return t, np.transpose(q.reshape(meqn, my, mx), (2, 1, 0))


def oracle_frames(p, q, aux, nout, tfinal, lxw=False, pif=False):
    dtv = np.array([0.0, p.initial_dt, p.max_dt])
    t = 0.0
    for n in range(1, nout + 1):
        rc, t, _ = orc.fin_solve(p, q, t, tfinal * n / nout, dtv, aux=aux, lxw=lxw, pif=pif)
        assert rc == 0
    return t, q


def test_euler_quadrants_app(tmp_path):
    run_app("euler_quadrants", tmp_path)
    p, q, aux = problems.riemann4(mx=64, initial_dt=1e-3, max_dt=1.0, nv=500000)
    dtv = np.array([0.0, p.initial_dt, p.max_dt])
    q0 = q.copy()
    steps = 0
    for n in (1, 2):
        rc, t0, st = orc.fin_solve(p, q0, 0.01 * (n - 1), 0.01 * n, dtv, aux=aux)
        assert rc == 0
        steps += st.steps_accepted
    t1, q1 = read_frame(tmp_path / "output" / "q0002.dat", (64, 64), 5)
    assert abs(t1 - t0) < 1e-15
    err = orc.rel_linf(q1, orc.interior(p, q0))
    print("app vs oracle", err)
    assert np.all(err <= 1e-10)
    # the other outputs: qhelp.dat, aux frames, conservation.dat with the t = 0
    # line and one line per accepted step (ConSoln's format)
    out = tmp_path / "output"
    qhelp = (out / "qhelp.dat").read_text().split("\n")
    assert qhelp[0].split()[0] == "2" and qhelp[4].split()[0] == "64" and qhelp[5].split()[0] == "64"
    assert (out / "a0000.dat").exists() and (out / "a0002.dat").exists()
    cons = np.loadtxt(out / "conservation.dat")
    assert cons.shape == (1 + steps, 6)
    g = p.mbc
    sums0 = (q[g:g + 64, g:g + 64] * (p.dx * p.dy)).sum(axis=(0, 1))
    assert cons[0, 0] == 0.0 and np.allclose(cons[0, 1:], sums0, rtol=1e-15, atol=1e-300)
    assert abs(cons[-1, 0] - 0.02) < 1e-15
    sums1 = (q1 * (p.dx * p.dy)).sum(axis=(0, 1))
    print("conservation.dat last line", cons[-1], "frame sums", sums1)
    assert np.allclose(cons[-1, 1:], sums1, rtol=1e-12, atol=1e-14)


def test_advection_app(tmp_path):
    run_app("advection_sine", tmp_path)
    p, q, aux = problems.advection_1d(mx=200, initial_dt=0.002, max_dt=1.0, nv=500000)
    t0, q0 = oracle_frames(p, q, aux, 1, 1.0)
    t1, q1 = read_frame(tmp_path / "output" / "q0001.dat", None, 1)
    assert abs(t1 - t0) < 1e-15
    err = orc.rel_linf(q1, orc.interior(p, q0))
    assert np.all(err <= 1e-10), err


def test_euler_quadrants_app_pif(tmp_path):
    """The app with `time_stepping_method = User-Defined`, `time_order = 4`:
    RunFinpack calls FinSolveUser, which this library links as RK4-PIF, as
    lib/2d/FinSolveRK_PIF.cpp does for the reference's apps."""
    ini = open(os.path.join(APPS, "euler_quadrants", "parameters.ini")).read()
    ini = ini.replace("time_stepping_method = Runge-Kutta", "time_stepping_method = User-Defined")
    ini = ini.replace("time_order  = 3", "time_order  = 4")
    (tmp_path / "parameters.ini").write_text(ini)
    r = subprocess.run([os.path.join(APPS, "euler_quadrants", "finess.exe")], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode == 0
    p, q, aux = problems.riemann4(mx=64, initial_dt=1e-3, max_dt=1.0, nv=500000, time_order=4,
                                  time_stepping_method=TS_PIF)
    t0, q0 = oracle_frames(p, q, aux, 2, 0.02, pif=True)
    t1, q1 = read_frame(tmp_path / "output" / "q0002.dat", (64, 64), 5)
    assert abs(t1 - t0) < 1e-15
    err = orc.rel_linf(q1, orc.interior(p, q0))
    print("PIF app vs oracle", err)
    assert np.all(err <= 1e-10), err


def test_euler_quadrants_app_wide_weno_global_alpha(tmp_path):
    """The app with `space_order = 9`, `weno_version = Z`, `mbc = 5` and
    `global_alpha = true` read from parameters.ini: the face-tile kernels with
    the global wave speeds, frames against the oracle (1e-10 or ten times the
    oracle's own spread under 1e-15 input noise: the ninth-order indicators
    cancel by several digits)."""
    ini = open(os.path.join(APPS, "euler_quadrants", "parameters.ini")).read()
    ini = ini.replace("space_order = 5", "space_order = 9").replace("mbc   = 3", "mbc   = 5")
    ini = ini.replace("weno_version  = JS", "weno_version  = Z")
    ini = ini.replace("source_term = false", "source_term = false\nglobal_alpha = true")
    (tmp_path / "parameters.ini").write_text(ini)
    r = subprocess.run([os.path.join(APPS, "euler_quadrants", "finess.exe")], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode == 0
    kw = dict(mx=64, initial_dt=1e-3, max_dt=1.0, nv=500000, space_order=9, mbc=5, weno_version=WENO_Z,
              global_alpha=1)
    p, q, aux = problems.riemann4(**kw)
    t0, q0 = oracle_frames(p, q.copy(), aux, 2, 0.02)
    noise = np.random.default_rng(5)
    spread = 0.0
    for _ in range(2):
        _, qn = oracle_frames(p, q * (1.0 + 1e-15 * (noise.random(q.shape) - 0.5)), aux, 2, 0.02)
        spread = max(spread, np.max(orc.rel_linf(orc.interior(p, qn), orc.interior(p, q0))))
    t1, q1 = read_frame(tmp_path / "output" / "q0002.dat", (64, 64), 5)
    assert abs(t1 - t0) < 1e-15
    err = orc.rel_linf(q1, orc.interior(p, q0))
    print("wide app vs oracle", err, "oracle spread", spread)
    assert np.all(err <= max(1e-10, 10.0 * spread)), err


def test_kelvin_helmholtz_lxw_app(tmp_path):
    run_app("kelvin_helmholtz", tmp_path)
    p, q, aux = problems.kelvin_helmholtz(mx=64, seed=20240611, cfl=0.3, max_cfl=0.35, initial_dt=1e-3,
                                          max_dt=1.0, nv=500000)
    t0, q0 = oracle_frames(p, q, aux, 2, 0.02, lxw=True)
    t1, q1 = read_frame(tmp_path / "output" / "q0002.dat", (64, 64), 5)
    assert abs(t1 - t0) < 1e-15
    err = orc.rel_linf(q1, orc.interior(p, q0))
    print("LxW app vs oracle", err)
    assert np.all(err <= 1e-10), err


def test_euler_quadrants_app_learned(tmp_path):
    """The same app with `[gpu] routing = learned, precision = fp32` and a
    network trained here and written with Indicator.save: it runs through the
    drop-in layer and stays close to the plain result (the routing is decided
    by the network, so the comparison is an error bound, not parity)."""
    from finess_amd.indicator import train_indicator
    ind = train_indicator(mx=128, epochs=2, stride=1, times=(0.0, 0.02, 0.05), holdout=(96, 0.03))
    ind.save(str(tmp_path / "net.bin"))
    ini = open(os.path.join(APPS, "euler_quadrants", "parameters.ini")).read()
    ini = ini.replace("precision = fp64", "precision = fp32\nrouting   = learned\nindicator = net.bin")
    (tmp_path / "parameters.ini").write_text(ini)
    r = subprocess.run([os.path.join(APPS, "euler_quadrants", "finess.exe")], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode == 0
    p, q, aux = problems.riemann4(mx=64, initial_dt=1e-3, max_dt=1.0, nv=500000)
    t0, q0 = oracle_frames(p, q, aux, 2, 0.02)
    t1, q1 = read_frame(tmp_path / "output" / "q0002.dat", (64, 64), 5)
    assert abs(t1 - t0) < 1e-6
    rho0 = orc.interior(p, q0)[..., 0]
    d = np.abs(q1[..., 0] - rho0)
    print("learned app vs plain oracle: L1", d.mean(), "Linf", d.max())
    assert d.mean() < 2e-3 * rho0.mean() and d.max() < 0.1 * rho0.max()
