"""The reference's own app callback sources, unchanged, against the drop-in
layer: compiled from their files (where the reference tree is present; the
test is skipped elsewhere) with the compat headers, the per-app parameter
accessors written by gen_app_params.py from the app's this_app_params.py, and
linked with the drop-in library; the drop-in probes then run on them (CPU
only, no device call).

* Euler riemann_test (apps/2d/euler/lib + riemann_test/{QinitFunc,
  SetBndValues,SourceTermFunc}.cpp): the Euler physics with the device's
  eigenvectors and wave speeds, transmissive boundaries on all four sides,
  the app's own get_OPT() accessor (QinitFunc.cpp:31) and its initial state;
* shallow-water riemann_test (apps/2d/shallow_water/lib + riemann_test): the
  h_l form of SetWaveSpd (apps/2d/shallow_water/lib/SetWaveSpd.cpp:27-28) and,
  with source_term on, the bathymetry SourceTermFunc;
* an app whose eigenvectors are rescaled is refused (test_compat.py)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
COMPAT = os.path.join(ROOT, "finess_amd", "compat")
LIB = os.path.join(ROOT, "finess_amd")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "apps", "2d", "euler", "riemann_test")),
                                reason="the reference tree is not present")


def _build(tmp_path, name, sources, app_params=None, defines=()):
    subprocess.run(["make", "-s", "-C", COMPAT], check=True)
    inc = tmp_path / "app_params.inc"
    if app_params:
        subprocess.run(["python3", os.path.join(COMPAT, "gen_app_params.py"), app_params, str(inc)], check=True)
    else:
        inc.write_text("")
    exe = tmp_path / name
    cmd = ["g++", "-std=c++17", "-O1", "-DNDIMS=2", '-DFINESS_APP_PARAMS="app_params.inc"', "-I", str(tmp_path),
           "-I", os.path.join(COMPAT, "include")] + list(defines) + \
          ["-o", str(exe), os.path.join(ROOT, "tests", "compat", "refapp_main.cpp")] + sources + \
          [os.path.join(LIB, "libfiness_dropin2d.a"), "-L" + LIB, "-lfiness_hip", "-Wl,-rpath," + LIB,
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return str(exe)


def _run(exe, ini):
    r = subprocess.run([exe, ini], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


def _lib(app, names):
    return [os.path.join(REF, "apps", "2d", app, "lib", n + ".cpp") for n in names]


LIBFILES = ["FluxFunc", "ProjectLeftEig", "ProjectRightEig", "SetWaveSpd", "DFluxFunc", "D2FluxFunc", "AuxFunc"]


def test_euler_riemann_test_relinks_unchanged(tmp_path):
    app = os.path.join(REF, "apps", "2d", "euler", "riemann_test")
    exe = _build(tmp_path, "euler_rt",
                 _lib("euler", LIBFILES) + [os.path.join(app, f + ".cpp") for f in
                                            ("QinitFunc", "SetBndValues", "SourceTermFunc")],
                 app_params=os.path.join(app, "this_app_params.py"), defines=["-DAPP_OPT"])
    rc, out = _run(exe, os.path.join(app, "parameters.ini"))
    print(out)
    assert rc == 0, out
    assert "probe physics=1 checked=1 bc=0,0,0,0 bc_checked=1" in out
    assert "OPT=1" in out
    assert "qinit 3 1" in out  # x-aligned shock (OPT = 1): rho = 3 left, 1 right


def test_shallow_water_riemann_test_probes(tmp_path):
    app = os.path.join(REF, "apps", "2d", "shallow_water", "riemann_test")
    exe = _build(tmp_path, "sw_rt",
                 _lib("shallow_water", LIBFILES) + [os.path.join(app, f + ".cpp") for f in
                                                    ("QinitFunc", "SetBndValues", "SourceTermFunc")])
    rc, out = _run(exe, os.path.join(app, "parameters.ini"))
    print(out)
    assert rc == 0, out
    assert "probe physics=2 checked=1 bc=0,0,0,0 bc_checked=1 sw_compat_hl=1" in out
    ini = tmp_path / "with_source.ini"
    txt = open(os.path.join(app, "parameters.ini")).read()
    assert "source_term = false" in txt
    ini.write_text(txt.replace("source_term = false", "source_term = true"))
    rc, out = _run(exe, str(ini))
    print(out)
    assert rc == 0, out
    assert "source_kind=1" in out  # FINESS_SRC_SW_BATHYMETRY
