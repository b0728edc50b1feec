"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the CPU
emulation of the kernel phase functions (tests/emu/libfiness_emu.so).
Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

from finess_amd import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
EMU = os.path.join(ROOT, "tests", "emu", "libfiness_emu.so")
PD = ctypes.POINTER(ctypes.c_double)
PF = ctypes.POINTER(ctypes.c_float)
PP = ctypes.POINTER(Params)


class OrcStats(ctypes.Structure):
    _fields_ = [("steps_accepted", ctypes.c_longlong), ("steps_rejected", ctypes.c_longlong),
                ("dt_min", ctypes.c_double), ("dt_max", ctypes.c_double), ("last_cfl", ctypes.c_double)]


def _build(path, mkdir):
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", mkdir], check=True)


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(PD)


_oracle = None
_emu = None


def oracle():
    global _oracle
    if _oracle is None:
        _build(ORACLE, os.path.join(ROOT, "oracle"))
        lib = ctypes.CDLL(ORACLE)
        sig = {
            "orc_weno5": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double, PD]),
            "orc_weno5_js_weights": (None, [ctypes.c_double, ctypes.c_double, PD, PD]),
            "orc_weno": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, PD]),
            "orc_weno_supported": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
            "orc_weno_tables": (ctypes.c_int, [ctypes.c_int, PD, PD, PD]),
            "orc_flux": (None, [PP, PD, PD, PD, PD]),
            "orc_project_left": (None, [PP, ctypes.c_int, PD, ctypes.c_int, PD, PD]),
            "orc_project_right": (None, [PP, ctypes.c_int, PD, ctypes.c_int, PD, PD]),
            "orc_wave_speed": (None, [PP, ctypes.c_int, PD, PD, PD, PD, PD, PD]),
            "orc_dflux": (None, [PP, PD, PD]),
            "orc_d2flux": (None, [PP, PD, PD]),
            "orc_set_bnd": (None, [PP, PD, PD]),
            "orc_construct_l": (ctypes.c_int, [PP, PD, PD, PD, PD]),
            "orc_get_cfl": (ctypes.c_int, [PP, ctypes.c_double, ctypes.c_double, PD, PD]),
            "orc_fin_solve_rk": (ctypes.c_int, [PP, PD, PD, PD, ctypes.c_double, PD, ctypes.POINTER(OrcStats)]),
            "orc_fin_solve_lxw": (ctypes.c_int, [PP, PD, PD, PD, ctypes.c_double, PD, ctypes.POINTER(OrcStats)]),
            "orc_fin_solve_rk_pif": (ctypes.c_int, [PP, PD, PD, PD, ctypes.c_double, PD, ctypes.POINTER(OrcStats)]),
            "orc_run_fixed_steps": (ctypes.c_int, [PP, PD, PD, ctypes.c_double, ctypes.c_int, PD]),
            "orc_construct_integrated_r": (ctypes.c_int, [PP, ctypes.c_double, PD, PD, PD, PD]),
            "orc_construct_lxw_l": (ctypes.c_int, [PP, PD, PD, PD, PD, PD, PD]),
            "orc_global_wave_spd": (None, [PP, PD, PD, PD, PD]),
            "orc_lxw_mpp_faces": (ctypes.c_int, [PP, ctypes.c_double, PD, PD, PD, PD, PD, PD, PD, PD]),
            "orc_num_threads": (ctypes.c_int, []),
            "orc_face_features": (None, [PP, ctypes.c_int, PD, PF]),
            "orc_face_departure": (ctypes.c_double, [PP, ctypes.c_int, PD]),
            "orc_face_flux_linear": (ctypes.c_double, [PP, ctypes.c_int, PD, PD]),
            "orc_run_fixed_steps_routed": (ctypes.c_int, [PP, PD, PD, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                                          ctypes.c_int, PD]),
            "orc_construct_l_routed": (ctypes.c_int, [PP, PD, PD, ctypes.c_int, ctypes.c_double, PD, PD,
                                                      ctypes.POINTER(ctypes.c_longlong)]),
            "orc_bf16": (ctypes.c_float, [ctypes.c_float]),
            "orc_mlp_set_rounding": (None, [ctypes.c_int]),
            "orc_mlp_forward": (None, [PF, PF, ctypes.c_int, PF]),
            "orc_mlp_gradient": (ctypes.c_double, [PF, PF, PF, ctypes.c_int, PD]),
            "orc_set_num_threads": (None, [ctypes.c_int]),
        }
        for n, (r, a) in sig.items():
            f = getattr(lib, n)
            f.restype, f.argtypes = r, a
        _oracle = lib
    return _oracle


def emu():
    global _emu
    if _emu is None:
        _build(EMU, os.path.join(ROOT, "tests", "emu"))
        lib = ctypes.CDLL(EMU)
        lib.emu_construct_l.restype = ctypes.c_int
        lib.emu_construct_l.argtypes = [PP, PD, PD, PD]
        lib.emu_construct_integrated_r.restype = ctypes.c_int
        lib.emu_construct_integrated_r.argtypes = [PP, ctypes.c_double, PD, PD, PD]
        lib.emu_run_fixed_steps.restype = ctypes.c_int
        lib.emu_run_fixed_steps.argtypes = [PP, PD, ctypes.c_double, ctypes.c_int]
        lib.emu_face_features.restype = None
        lib.emu_face_features.argtypes = [PP, ctypes.c_int, PD, ctypes.c_int, PF]
        lib.emu_face_departure.restype = ctypes.c_double
        lib.emu_face_departure.argtypes = [PP, ctypes.c_int, PD]
        lib.emu_face_flux_linear.restype = ctypes.c_double
        lib.emu_face_flux_linear.argtypes = [PP, ctypes.c_int, PD, PD]
        _emu = lib
    return _emu


# ---- convenience wrappers ---------------------------------------------------
def construct_l(p, q, aux=None):
    lib = oracle()
    q = np.ascontiguousarray(q, dtype=np.float64)
    aux = np.ones(p.shape_aux()) if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    L = np.zeros(p.shape_q())
    sm = np.zeros(p.shape_smax())
    rc = lib.orc_construct_l(ctypes.byref(p), ptr(q), ptr(aux), ptr(L), ptr(sm))
    assert rc == 0, rc
    return L, sm


def set_bnd(p, q, aux=None):
    aux = np.ones(p.shape_aux()) if aux is None else aux
    oracle().orc_set_bnd(ctypes.byref(p), ptr(q), ptr(aux))
    return q


def fin_solve(p, q, t, tend, dtv, aux=None, lxw=False, pif=False):
    lib = oracle()
    aux = np.ones(p.shape_aux()) if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    tt = ctypes.c_double(t)
    st = OrcStats()
    fn = lib.orc_fin_solve_rk_pif if pif else lib.orc_fin_solve_lxw if lxw else lib.orc_fin_solve_rk
    rc = fn(ctypes.byref(p), ptr(q), ptr(aux), ctypes.byref(tt), float(tend), ptr(dtv), ctypes.byref(st))
    return rc, tt.value, st


def fixed_steps(p, q, dt, n, aux=None):
    aux = np.ones(p.shape_aux()) if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    cfl = ctypes.c_double()
    rc = oracle().orc_run_fixed_steps(ctypes.byref(p), ptr(q), ptr(aux), float(dt), int(n), ctypes.byref(cfl))
    assert rc == 0, rc
    return cfl.value


def integrated_r(p, dt, q, aux=None):
    aux = np.ones(p.shape_aux()) if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    F = np.zeros(p.shape_q())
    G = np.zeros(p.shape_q())
    rc = oracle().orc_construct_integrated_r(ctypes.byref(p), float(dt), ptr(q), ptr(aux), ptr(F), ptr(G))
    assert rc == 0, rc
    return F, G


def global_wave_spd(p, q, aux=None):
    aux = np.ones(p.shape_aux()) if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    a1, a2 = ctypes.c_double(), ctypes.c_double()
    oracle().orc_global_wave_spd(ctypes.byref(p), ptr(q), ptr(aux), ctypes.byref(a1), ctypes.byref(a2))
    return a1.value, a2.value


def face_shapes(p):
    """Shapes of the x-face and y-face flux arrays (ghosts included)."""
    g = p.mbc
    return (p.mx + 1 + 2 * g, p.my + 2 * g, p.meqn), (p.mx + 2 * g, p.my + 1 + 2 * g, p.meqn)


def lxw_mpp_faces(p, dt, q, aux=None):
    """High-order, Lax-Friedrichs and limited interface fluxes of one limited
    LxW step of q (boundary values set): dict of x-face / y-face arrays."""
    aux = np.ones(p.shape_aux()) if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    sx, sy = face_shapes(p)
    out = {k: np.zeros(sx if k[0] == "F" else sy) for k in ("Fh", "Gh", "Fl", "Gl", "Fm", "Gm")}
    rc = oracle().orc_lxw_mpp_faces(ctypes.byref(p), float(dt), ptr(q), ptr(aux), ptr(out["Fh"]), ptr(out["Gh"]),
                                    ptr(out["Fl"]), ptr(out["Gl"]), ptr(out["Fm"]), ptr(out["Gm"]))
    assert rc == 0, rc
    return out


def emu_construct_l(p, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    L = np.zeros(p.shape_q())
    sm = np.zeros(p.shape_smax())
    assert emu().emu_construct_l(ctypes.byref(p), ptr(q), ptr(L), ptr(sm)) == 0
    return L, sm


def emu_integrated_r(p, dt, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    F = np.zeros(p.shape_q())
    G = np.zeros(p.shape_q())
    assert emu().emu_construct_integrated_r(ctypes.byref(p), float(dt), ptr(q), ptr(F), ptr(G)) == 0
    return F, G


def interior(p, a):
    g = p.mbc
    if p.ndims == 1:
        return a[g:g + p.mx]
    return a[g:g + p.mx, g:g + p.my]


def rel_linf(a, b):
    """Relative L-infinity per field (last axis): ||a-b|| / max(||b||, tiny);
    absolute error for an identically-zero field."""
    a = a.reshape(-1, a.shape[-1])
    b = b.reshape(-1, b.shape[-1])
    out = []
    for m in range(a.shape[1]):
        nb = np.max(np.abs(b[:, m]))
        d = np.max(np.abs(a[:, m] - b[:, m]))
        out.append(d / nb if nb > 0 else d)
    return np.array(out)


def emu_fixed_steps(p, q, dt, n):
    assert emu().emu_run_fixed_steps(ctypes.byref(p), ptr(q), float(dt), int(n)) == 0
    return q


def fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PF)


def construct_l_routed(p, q, mode, tau=0.05):
    q = np.ascontiguousarray(q, dtype=np.float64)
    aux = np.zeros(p.shape_aux())
    L = np.zeros(p.shape_q())
    sm = np.zeros(p.shape_smax())
    nt = ctypes.c_longlong()
    assert oracle().orc_construct_l_routed(ctypes.byref(p), ptr(q), ptr(aux), int(mode), float(tau), ptr(L), ptr(sm),
                                           ctypes.byref(nt)) == 0
    return L, sm, nt.value


def face_stencil(p, q, dir_, i, j):
    """States q(i-3..i+2) across x-face (i-1/2, j) or y-face (i, j-1/2), 1-based."""
    g = p.mbc
    if dir_ == 1:
        return np.ascontiguousarray(q[g + i - 4:g + i + 2, g + j - 1, :])
    return np.ascontiguousarray(q[g + i - 1, g + j - 4:g + j + 2, :])


def features(p, dir_, qs):
    f = np.zeros(24, np.float32)
    oracle().orc_face_features(ctypes.byref(p), int(dir_), ptr(np.ascontiguousarray(qs, np.float64)), fptr(f))
    return f


def departure(p, dir_, qs):
    return oracle().orc_face_departure(ctypes.byref(p), int(dir_), ptr(np.ascontiguousarray(qs, np.float64)))


def mlp_forward(P, X):
    X = np.ascontiguousarray(X, np.float32)
    z = np.zeros(len(X), np.float32)
    oracle().orc_mlp_forward(fptr(np.ascontiguousarray(P, np.float32)), fptr(X), len(X), fptr(z))
    return z


def mlp_gradient(P, X, y):
    g = np.zeros(1889)
    loss = oracle().orc_mlp_gradient(fptr(np.ascontiguousarray(P, np.float32)), fptr(np.ascontiguousarray(X, np.float32)),
                                     fptr(np.ascontiguousarray(y, np.float32)), len(y), ptr(g))
    return g, loss


def face_flux_linear(p, dir_, qs):
    fh = np.zeros(5)
    a = oracle().orc_face_flux_linear(ctypes.byref(p), int(dir_), ptr(np.ascontiguousarray(qs, np.float64)), ptr(fh))
    return fh, a


def emu_face_features(p, dir_, qs, fp32=False):
    f = np.zeros(24, np.float32)
    emu().emu_face_features(ctypes.byref(p), int(dir_), ptr(np.ascontiguousarray(qs, np.float64)), int(fp32), fptr(f))
    return f


def emu_face_departure(p, dir_, qs):
    return emu().emu_face_departure(ctypes.byref(p), int(dir_), ptr(np.ascontiguousarray(qs, np.float64)))


def emu_face_flux_linear(p, dir_, qs):
    fh = np.zeros(5)
    a = emu().emu_face_flux_linear(ctypes.byref(p), int(dir_), ptr(np.ascontiguousarray(qs, np.float64)), ptr(fh))
    return fh, a


def fixed_steps_routed(p, q, dt, n, tau, per_step):
    """SSP-RK3 steps with exact-departure routing; q (reference layout) in place."""
    aux = np.zeros(p.shape_aux())
    c = ctypes.c_double()
    assert oracle().orc_run_fixed_steps_routed(ctypes.byref(p), ptr(q), ptr(aux), float(dt), int(n), float(tau),
                                               int(per_step), ctypes.byref(c)) == 0
    return c.value
