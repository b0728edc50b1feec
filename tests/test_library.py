"""The C-ABI library: it loads, exports every symbol include/*.h declares,
and its host-side parameter handling follows the reference's schema checks.
No compute call is made (this runs without a GPU)."""
import ctypes
import os
import re

import pytest

import finess_amd
from finess_amd import ERR_CONFIG, OK, Params, default_params, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names.update(re.findall(r"\b(finess_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = load_library()
    names = declared_symbols()
    assert "finess_construct_l" in names and "finess_fin_solve_rk" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(finess_amd.SIGNATURES) == names


def test_defaults_follow_reference_schema():
    p = default_params()
    assert p.weno_version == finess_amd.WENO_JS and p.epsilon == 1e-6 and p.power_param == 2.0
    assert p.alpha_scaling == 1.1 and p.gamma == 1.4 and p.space_order == 5
    assert p.global_alpha == 0 and p.dx == (p.xhigh - p.xlow) / p.mx


@pytest.mark.parametrize("kw", [
    dict(space_order=3),                # GetWenoReconstruct has no third order
    dict(space_order=7),                # needs mbc >= 4 (lib/2d/ConstructL.cpp:31-33)
    dict(global_alpha=1, routing=finess_amd.ROUTE_ORACLE),  # routing has no global-alpha path
    dict(max_cfl=0.3, desired_cfl=0.4),  # finess.py: max_cfl >= desired_cfl
    dict(initial_dt=1.0, max_dt=0.5),    # finess.py: max_dt >= initial_dt
    dict(alpha_scaling=0.9),             # weno.py: alpha_scaling >= 1
    dict(mbc=2),
    dict(nv=0),
])
def test_finalize_rejects(kw):
    lib = load_library()
    p = Params()
    lib.finess_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    assert lib.finess_params_finalize(ctypes.byref(p)) == ERR_CONFIG


def test_lxw_needs_five_ghosts():
    lib = load_library()
    p = Params()
    lib.finess_params_default(ctypes.byref(p))
    p.time_stepping_method = finess_amd.TS_LXW
    assert lib.finess_params_finalize(ctypes.byref(p)) == ERR_CONFIG
    p.mbc = 5
    assert lib.finess_params_finalize(ctypes.byref(p)) == OK


def test_host_boundary_scope():
    """FINESS_BC_HOST (a host SetBndValues before every stage) is accepted for
    the 2D Runge-Kutta schemes of orders 1-3, routed or not, and refused
    elsewhere."""
    lib = load_library()

    def fin(**kw):
        p = Params()
        lib.finess_params_default(ctypes.byref(p))
        p.bc[3] = finess_amd.BC_HOST
        for k, v in kw.items():
            setattr(p, k, v)
        return lib.finess_params_finalize(ctypes.byref(p))

    assert fin() == OK
    assert fin(time_order=1) == OK and fin(precision=finess_amd.PREC_FP32) == OK
    assert fin(time_order=4) == ERR_CONFIG
    assert fin(routing=finess_amd.ROUTE_ORACLE) == OK
    assert fin(time_stepping_method=finess_amd.TS_LXW, mbc=5) == ERR_CONFIG
    assert fin(ndims=1, physics=finess_amd.PHYS_ADVECTION, meqn=1) == ERR_CONFIG
    p = Params()
    lib.finess_params_default(ctypes.byref(p))
    p.bc[0] = 4
    assert lib.finess_params_finalize(ctypes.byref(p)) == ERR_CONFIG
    # the no-alloc entry has no ghosts to take a host rule's values: refused
    # before any device call
    p.bc[0] = finess_amd.BC_HOST
    dummy = ctypes.c_void_p(ctypes.addressof(ctypes.c_double()))  # never dereferenced: the check comes first
    assert lib.finess_construct_l_planes_noalloc(ctypes.byref(p), dummy, dummy, dummy, None) == ERR_CONFIG


def test_version_string():
    assert b"gfx950" in load_library().finess_version()
