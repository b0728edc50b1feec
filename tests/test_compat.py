"""The C++ drop-in layer (finess_amd/compat): it builds, parameters.ini is
read with the reference's grammar and accessor semantics, and the probes
identify the app's physics and boundary conditions from the app's own
callbacks (here forwarding to the oracle's restatement of the reference
callbacks).  CPU only: the harness makes no device call."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT = os.path.join(ROOT, "finess_amd", "compat")
HARNESS = os.path.join(ROOT, "tests", "compat", "probe_main")
INI = os.path.join(COMPAT, "apps", "euler_quadrants", "parameters.ini")


@pytest.fixture(scope="module")
def harness():
    subprocess.run(["make", "-s", "-C", COMPAT], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "compat")], check=True)
    return HARNESS


def run(h, ini, bc="extrap", phys="euler"):
    r = subprocess.run([h, ini, bc, phys], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


def test_apps_link(harness):
    for app in ("euler_quadrants", "advection_sine", "kelvin_helmholtz"):
        assert os.path.exists(os.path.join(COMPAT, "apps", app, "finess.exe"))


def test_ini_values(harness):
    rc, out = run(harness, INI)
    assert rc == 0, out
    assert "mx=64 my=64 mbc=3 meqn=5 dx=0.015625 tsm=0" in out
    assert "eps=1e-06 alpha_scaling=1.1 gamma=1.4" in out


@pytest.mark.parametrize("bc,expect", [("extrap", "0,0,0,0"), ("periodic", "1,1,1,1"), ("mixed", "1,1,2,0")])
def test_probe_identifies_boundaries(harness, bc, expect):
    rc, out = run(harness, INI, bc)
    assert rc == 0, out
    assert "probe physics=1 checked=1 bc=%s bc_checked=1" % expect in out


@pytest.mark.parametrize("bc,expect", [("timedep", "3,0,0,0"), ("clock", "0,0,0,3")])
def test_probe_sends_other_boundaries_to_the_host(harness, bc, expect):
    """A side whose ghosts are no copy of a cell, or change between t = 0 and
    t = 1, runs the app's SetBndValues on the host (FINESS_BC_HOST)."""
    rc, out = run(harness, INI, bc)
    assert rc == 0, out
    assert "probe physics=1 checked=1 bc=%s bc_checked=1" % expect in out


def test_probe_rejects_foreign_physics(harness):
    rc, out = run(harness, INI, "extrap", "bogus")
    assert rc == 3 and "matches no device physics" in out


def test_ini_grammar(harness, tmp_path):
    txt = open(INI).read()
    # comments, trailing spaces and repeated keys (the last assignment wins)
    txt = txt.replace("mx    = 64", "mx    = 32   ; first\nmx = 48    \n; a comment line\n   mx=64  ; last wins")
    f = tmp_path / "p.ini"
    f.write_text(txt)
    rc, out = run(harness, str(f))
    assert rc == 0 and "mx=64" in out, out
    # a check of the schema (max_cfl >= desired_cfl) fails like terminate()
    f.write_text(open(INI).read().replace("max_cfl     = 0.45", "max_cfl     = 0.30"))
    rc, out = run(harness, str(f))
    assert rc == 4 and "max_cfl" in out, out
    # a missing required key
    f.write_text(open(INI).read().replace("nv          = 500000\n", ""))
    rc, out = run(harness, str(f))
    assert rc == 4 and "finess.nv is required" in out, out


def test_probe_rejects_rescaled_eigenvectors(harness):
    """Eigenvector projections that are inverses of each other but scaled
    differently from the device's change the WENO result: refused."""
    rc, out = run(harness, INI, "extrap", "euler_rescaled")
    assert rc == 3 and "ProjectLeftEig differs from the device's left eigenvectors" in out, out


@pytest.mark.parametrize("phys,hl", [("sw", 1), ("sw_hr", 0)])
def test_probe_detects_the_shallow_water_wave_speed_form(harness, tmp_path, phys, hl):
    """SetWaveSpd dividing the right momenta by h_l (the reference's own,
    apps/2d/shallow_water/lib/SetWaveSpd.cpp:27-28) or by h_r: the probe
    finds which one the app has."""
    f = tmp_path / "sw.ini"
    f.write_text(open(INI).read().replace("meqn        = 5", "meqn        = 3").replace("physics   = euler",
                                                                                        "physics   = shallow_water"))
    rc, out = run(harness, str(f), "periodic", phys)
    assert rc == 0, out
    assert "probe physics=2 checked=1 bc=1,1,1,1 bc_checked=1 sw_compat_hl=%d" % hl in out, out
