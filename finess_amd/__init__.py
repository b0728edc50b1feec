"""finess_amd -- MI355X (gfx950) implementation of FINESS's finite-difference
WENO operator and its time steppers, driven through the C-ABI of
``libfiness_hip.so`` (declared in ``include/finess_hip.h``).

This module is the Python mirror of the reference's operator interface:
``Solver.construct_l`` is ``ConstructL`` (lib/2d/ConstructL.cpp:13,
lib/1d/ConstructL.cpp:16), ``Solver.fin_solve_rk`` is ``FinSolveRK``
(lib/2d/FinSolveRK.cpp:13) and ``Solver.fin_solve_lxw`` is ``FinSolveLxW``
(lib/2d/FinSolveLxW.cpp:16), with the reference's argument meaning: state
arrays in the reference tensor layout with ghost cells, ``dtv`` a 3-vector
whose entry 1 is the in/out time step and entry 2 the maximum time step.

Errors: the reference exits (``exit(1)``) or throws; here every failing call
raises :class:`FinessError` carrying the C status code.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FINESS_HIP_LIB") or os.path.join(_HERE, "libfiness_hip.so")

# status codes (include/finess_hip.h)
OK, ERR_CONFIG, ERR_NONFINITE, ERR_NV, ERR_HIP, ERR_DT, ERR_CFL, ERR_ARG, ERR_NODEVICE = 0, -1, -2, -3, -4, -5, -6, -7, -8
WENO_JS, WENO_FD, WENO_Z = 0, 1, 2
TS_RK, TS_LXW, TS_PIF = 0, 1, 2
PHYS_ADVECTION, PHYS_EULER, PHYS_SHALLOW_WATER = 0, 1, 2
BC_EXTRAP, BC_PERIODIC, BC_REFLECT, BC_HOST = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
ROUTE_NONE, ROUTE_LEARNED, ROUTE_ORACLE = 0, 1, 2
SRC_ZERO, SRC_SW_BATHYMETRY, SRC_DAMPING = 0, 1, 2

_ERRNAMES = {
    ERR_CONFIG: "unsupported configuration",
    ERR_NONFINITE: "non-finite state",
    ERR_NV: "exceeded allowed # of time steps (nv)",
    ERR_HIP: "HIP runtime error",
    ERR_DT: "dt is out of bounds (dt > max_dt)",
    ERR_CFL: "CFL number is out of bounds",
    ERR_ARG: "bad argument",
    ERR_NODEVICE: "no GPU device",
}


class FinessError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__("finess error %d (%s)%s" % (code, _ERRNAMES.get(code, "?"), (": " + msg) if msg else ""))


class Params(ctypes.Structure):
    """Mirror of ``finess_params`` (include/finess_hip.h)."""

    _fields_ = [
        ("ndims", ctypes.c_int),
        ("mx", ctypes.c_int), ("my", ctypes.c_int), ("mbc", ctypes.c_int),
        ("xlow", ctypes.c_double), ("xhigh", ctypes.c_double),
        ("ylow", ctypes.c_double), ("yhigh", ctypes.c_double),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("meqn", ctypes.c_int), ("maux", ctypes.c_int),
        ("space_order", ctypes.c_int), ("time_order", ctypes.c_int),
        ("time_stepping_method", ctypes.c_int),
        ("initial_dt", ctypes.c_double), ("max_dt", ctypes.c_double),
        ("desired_cfl", ctypes.c_double), ("max_cfl", ctypes.c_double),
        ("nv", ctypes.c_int),
        ("global_alpha", ctypes.c_int), ("source_term", ctypes.c_int), ("verbosity", ctypes.c_int),
        ("weno_version", ctypes.c_int),
        ("epsilon", ctypes.c_double), ("power_param", ctypes.c_double), ("alpha_scaling", ctypes.c_double),
        ("physics", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("sw_compat_hl", ctypes.c_int),
        ("bc", ctypes.c_int * 4),
        ("precision", ctypes.c_int),
        ("routing", ctypes.c_int),
        ("route_tau", ctypes.c_double),
        ("tile_i", ctypes.c_int), ("tile_j", ctypes.c_int),
        ("route_per_step", ctypes.c_int),
        ("host_controller", ctypes.c_int),
        ("mpp_limiter", ctypes.c_int),
        ("source_kind", ctypes.c_int),
        ("source_coef", ctypes.c_double),
    ]

    def copy(self):
        p = Params()
        ctypes.pointer(p)[0] = self
        return p

    def shape_q(self):
        """Shape of a reference-layout state array (ghosts included)."""
        if self.ndims == 1:
            return (self.mx + 2 * self.mbc, self.meqn)
        return (self.mx + 2 * self.mbc, self.my + 2 * self.mbc, self.meqn)

    def shape_aux(self):
        na = max(1, self.maux)
        if self.ndims == 1:
            return (self.mx + 2 * self.mbc, na)
        return (self.mx + 2 * self.mbc, self.my + 2 * self.mbc, na)

    def shape_smax(self):
        if self.ndims == 1:
            return (self.mx + 2 * self.mbc,)
        return (self.mx + 1 + 2 * self.mbc, self.my + 1 + 2 * self.mbc, 2)


class Stats(ctypes.Structure):
    _fields_ = [
        ("steps_accepted", ctypes.c_longlong),
        ("steps_rejected", ctypes.c_longlong),
        ("dt_min", ctypes.c_double),
        ("dt_max", ctypes.c_double),
        ("last_cfl", ctypes.c_double),
        ("t", ctypes.c_double),
    ]


_lib = None
_PD = ctypes.POINTER(ctypes.c_double)
# finess_bnd_fn: int (*)(void* user, double t, double* q)
BND_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double))
_PF = ctypes.POINTER(ctypes.c_float)
MLP_NPARAM = 1889

# every symbol include/finess_hip.h declares, with its ctypes signature
SIGNATURES = {
    "finess_params_default": (None, [ctypes.POINTER(Params)]),
    "finess_params_finalize": (ctypes.c_int, [ctypes.POINTER(Params)]),
    "finess_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)]),
    "finess_destroy": (None, [ctypes.c_void_p]),
    "finess_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "finess_set_state": (ctypes.c_int, [ctypes.c_void_p, _PD, _PD, ctypes.c_double]),
    "finess_get_state": (ctypes.c_int, [ctypes.c_void_p, _PD, _PD]),
    "finess_construct_l": (ctypes.c_int, [ctypes.c_void_p, _PD, _PD, _PD, _PD]),
    "finess_fin_solve_rk": (ctypes.c_int, [ctypes.c_void_p, _PD, _PD, _PD, ctypes.c_double, _PD]),
    "finess_fin_solve_lxw": (ctypes.c_int, [ctypes.c_void_p, _PD, _PD, _PD, ctypes.c_double, _PD]),
    "finess_set_step_log": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "finess_set_bnd_callback": (ctypes.c_int, [ctypes.c_void_p, BND_FN, ctypes.c_void_p]),
    "finess_read_step_log": (ctypes.c_int, [ctypes.c_void_p, _PD, ctypes.c_int]),
    "finess_fin_solve_rk_pif": (ctypes.c_int, [ctypes.c_void_p, _PD, _PD, _PD, ctypes.c_double, _PD]),
    "finess_apply_mpp_limiter_2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _PD, _PD, _PD, _PD, _PD]),
    "finess_construct_integrated_r": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _PD, _PD, _PD, _PD]),
    "finess_advance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _PD, ctypes.POINTER(Stats)]),
    "finess_advance_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _PD, ctypes.POINTER(Stats)]),
    "finess_run_fixed_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _PD]),
    "finess_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "finess_last_kernel_time": (ctypes.c_int, [ctypes.c_void_p, _PD, ctypes.POINTER(ctypes.c_longlong)]),
    "finess_version": (ctypes.c_char_p, []),
    "finess_device_count": (ctypes.c_int, []),
    "finess_set_indicator": (ctypes.c_int, [ctypes.c_void_p, _PF, ctypes.c_float]),
    "finess_route_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                                          ctypes.POINTER(ctypes.c_longlong)]),
    "finess_mlp_create": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_longlong, ctypes.POINTER(ctypes.c_void_p)]),
    "finess_mlp_destroy": (None, [ctypes.c_void_p]),
    "finess_mlp_harvest": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_longlong)]),
    "finess_mlp_get_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong, _PF, _PF, _PF]),
    "finess_mlp_set_data": (ctypes.c_int, [ctypes.c_void_p, _PF, _PF, ctypes.c_longlong]),
    "finess_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, _PF]),
    "finess_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _PF]),
    "finess_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _PF]),
    "finess_mlp_gradient": (ctypes.c_int, [ctypes.c_void_p, _PF, _PF, ctypes.c_int, _PF, _PF]),
    "finess_mlp_eval": (ctypes.c_int, [ctypes.c_void_p, _PF, ctypes.c_longlong, ctypes.c_int, _PF]),
    "finess_batch_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_void_p)]),
    "finess_batch_destroy": (None, [ctypes.c_void_p]),
    "finess_batch_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "finess_batch_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _PD, ctypes.c_double, ctypes.c_double]),
    "finess_batch_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _PD, _PD, _PD,
                                              ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "finess_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]),
    "finess_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _PD]),
    "finess_physics_eig": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_int, _PD, _PD, _PD]),
    "finess_physics_wave_speed": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_int, _PD, _PD, _PD, _PD, _PD]),
    "finess_physics_source": (ctypes.c_int, [ctypes.POINTER(Params), _PD, _PD, _PD]),
    "finess_construct_l_planes_noalloc": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_void_p, ctypes.c_void_p]),
}


def load_library(path=None):
    """Loads ``libfiness_hip.so``.  Raises if it has not been built: there is
    no CPU fallback for any operator of this package."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FinessError(ERR_ARG, "%s is missing; build it with `python -m finess_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def default_params(**kw):
    """``finess_params_default`` then keyword overrides, then finalize."""
    lib = load_library()
    p = Params()
    lib.finess_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k == "bc":
            for i, b in enumerate(v):
                p.bc[i] = b
        else:
            setattr(p, k, v)
    rc = lib.finess_params_finalize(ctypes.byref(p))
    if rc:
        raise FinessError(rc, "parameters rejected")
    return p


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)


def _fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PF)


class Solver:
    """One device context (``finess_ctx``) bound to a parameter set."""

    def __init__(self, params):
        self.lib = load_library()
        self.p = params.copy()
        h = ctypes.c_void_p()
        rc = self.lib.finess_create(ctypes.byref(self.p), ctypes.byref(h))
        if rc:
            raise FinessError(rc, "finess_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.finess_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc:
            msg = self.lib.finess_last_error(self.h)
            raise FinessError(rc, "%s: %s" % (what, msg.decode() if msg else ""))

    # ---- state -------------------------------------------------------
    def set_state(self, q, aux=None, t=0.0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if aux is not None:
            aux = np.ascontiguousarray(aux, dtype=np.float64)
        self._check(self.lib.finess_set_state(self.h, _ptr(q), _ptr(aux), float(t)), "set_state")

    def get_state(self):
        q = np.zeros(self.p.shape_q())
        t = ctypes.c_double()
        self._check(self.lib.finess_get_state(self.h, _ptr(q), ctypes.byref(t)), "get_state")
        return q, t.value

    # ---- drop-in operators ---------------------------------------------
    def construct_l(self, q, aux=None, smax=None):
        """ConstructL: returns (Lstar, smax) in the reference layout."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        if aux is None:
            aux = np.ones(self.p.shape_aux())
        aux = np.ascontiguousarray(aux, dtype=np.float64)
        L = np.zeros(self.p.shape_q())
        sm = np.zeros(self.p.shape_smax()) if smax is None else np.ascontiguousarray(smax, dtype=np.float64)
        self._check(self.lib.finess_construct_l(self.h, _ptr(q), _ptr(aux), _ptr(L), _ptr(sm)), "construct_l")
        return L, sm

    def construct_integrated_r(self, dt, q, aux=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        F = np.zeros(self.p.shape_q())
        G = np.zeros(self.p.shape_q())
        self._check(self.lib.finess_construct_integrated_r(self.h, float(dt), _ptr(q), None, _ptr(F), _ptr(G)),
                    "construct_integrated_r")
        return F, G

    def set_bnd_callback(self, fn):
        """SetBndValues for the FINESS_BC_HOST sides: ``fn(t, q)`` receives the
        reference-layout state with ghosts (interior cells within mbc of the
        boundary filled, the rest NaN) and fills its ghost cells in place."""
        shape = self.p.shape_q()
        n = int(np.prod(shape))

        def tramp(_user, t, qp):
            try:
                fn(float(t), np.ctypeslib.as_array(qp, shape=(n,)).reshape(shape))
                return 0
            except Exception:
                return 1

        self._bnd = BND_FN(tramp) if fn is not None else None  # kept alive while installed
        self._check(self.lib.finess_set_bnd_callback(self.h, self._bnd if fn is not None else BND_FN(), None),
                    "set_bnd_callback")

    def set_step_log(self, on=True):
        self._check(self.lib.finess_set_step_log(self.h, 1 if on else 0), "set_step_log")

    def read_step_log(self, max_records=1 << 20):
        """The conservation records of the accepted steps since the last
        read: array (n, meqn + 1) of (t, sum dx dy q per field)."""
        buf = np.zeros((max_records, self.p.meqn + 1))
        n = self.lib.finess_read_step_log(self.h, _ptr(buf), max_records)
        if n < 0:
            self._check(n, "read_step_log")
        return buf[:n].copy()

    def apply_mpp_limiter_2d(self, dt, q, fLF, gLF, fHat, gHat):
        """ApplyMPPLimiter2D: limits fHat, gHat (reference layouts with
        ghosts) in place toward fLF, gLF for the update of q by dt."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (q, fLF, gLF)]
        for a in (fHat, gHat):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        self._check(self.lib.finess_apply_mpp_limiter_2d(self.h, float(dt), _ptr(arrs[0]), _ptr(arrs[1]),
                                                         _ptr(arrs[2]), _ptr(fHat), _ptr(gHat)),
                    "apply_mpp_limiter_2d")

    def fin_solve_rk(self, q, t, tend, dtv, aux=None):
        """FinSolveRK: advances q in place from t to tend; returns t."""
        return self._fin_solve(self.lib.finess_fin_solve_rk, q, t, tend, dtv, aux)

    def fin_solve_lxw(self, q, t, tend, dtv, aux=None):
        return self._fin_solve(self.lib.finess_fin_solve_lxw, q, t, tend, dtv, aux)

    def fin_solve_rk_pif(self, q, t, tend, dtv, aux=None):
        return self._fin_solve(self.lib.finess_fin_solve_rk_pif, q, t, tend, dtv, aux)

    def _fin_solve(self, fn, q, t, tend, dtv, aux):
        assert q.dtype == np.float64 and q.flags["C_CONTIGUOUS"]
        assert dtv.dtype == np.float64 and dtv.shape == (3,)
        tt = ctypes.c_double(t)
        if aux is not None:
            aux = np.ascontiguousarray(aux, dtype=np.float64)
        self._check(fn(self.h, _ptr(q), _ptr(aux), ctypes.byref(tt), float(tend), _ptr(dtv)), "fin_solve")
        return tt.value

    # ---- device-resident stepping ---------------------------------------
    def advance(self, tend, dtv):
        st = Stats()
        self._check(self.lib.finess_advance(self.h, float(tend), _ptr(dtv), ctypes.byref(st)), "advance")
        return st

    def advance_steps(self, nsteps, dtv):
        st = Stats()
        self._check(self.lib.finess_advance_steps(self.h, int(nsteps), _ptr(dtv), ctypes.byref(st)), "advance_steps")
        return st

    def run_fixed_steps(self, dt, nsteps):
        cfl = ctypes.c_double()
        self._check(self.lib.finess_run_fixed_steps(self.h, float(dt), int(nsteps), ctypes.byref(cfl)),
                    "run_fixed_steps")
        return cfl.value

    def sync(self):
        self._check(self.lib.finess_sync(self.h), "sync")

    def last_kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(self.lib.finess_last_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)), "last_kernel_time")
        return ms.value, n.value

    # ---- troubled-interface indicator ------------------------------------
    def set_indicator(self, params, threshold):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.shape == (MLP_NPARAM,)
        self._check(self.lib.finess_set_indicator(self.h, _fptr(params), float(threshold)), "set_indicator")

    def route_stats(self):
        t = ctypes.c_longlong()
        n = ctypes.c_longlong()
        self._check(self.lib.finess_route_stats(self.h, ctypes.byref(t), ctypes.byref(n)), "route_stats")
        return t.value, n.value


class Mlp:
    """The 24-32-32-1 troubled-interface network (``finess_mlp``): a device
    training set harvested from solver states, Adam training and evaluation
    on the matrix cores."""

    def __init__(self, seed=1, capacity=1 << 22):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.finess_mlp_create(int(seed), int(capacity), ctypes.byref(h))
        if rc:
            raise FinessError(rc, "finess_mlp_create")
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.finess_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise FinessError(rc, what)

    def harvest(self, solver, stride=1, tau=0.05):
        n = ctypes.c_longlong()
        self._check(self.lib.finess_mlp_harvest(self.h, solver.h, int(stride), float(tau), ctypes.byref(n)), "harvest")
        self.n = n.value
        return self.n

    def get_data(self, first=0, n=None):
        n = self.n - first if n is None else n
        X = np.zeros((n, 24), np.float32)
        y = np.zeros(n, np.float32)
        d = np.zeros(n, np.float32)
        self._check(self.lib.finess_mlp_get_data(self.h, int(first), int(n), _fptr(X), _fptr(y), _fptr(d)), "get_data")
        return X, y, d

    def set_data(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        self._check(self.lib.finess_mlp_set_data(self.h, _fptr(X), _fptr(y), len(y)), "set_data")
        self.n = len(y)

    def train(self, epochs=4, batch=4096, lr=2e-3):
        loss = np.zeros(epochs, np.float32)
        self._check(self.lib.finess_mlp_train(self.h, int(epochs), int(batch), float(lr), _fptr(loss)), "train")
        return loss

    @property
    def params(self):
        P = np.zeros(MLP_NPARAM, np.float32)
        self._check(self.lib.finess_mlp_get_params(self.h, _fptr(P)), "get_params")
        return P

    @params.setter
    def params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        self._check(self.lib.finess_mlp_set_params(self.h, _fptr(P)), "set_params")

    def gradient(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.finess_mlp_gradient(self.h, _fptr(X), _fptr(y), len(y), _fptr(g), _fptr(loss)), "gradient")
        return g, float(loss[0])

    def logits(self, X, path=0):
        X = np.ascontiguousarray(X, np.float32)
        z = np.zeros(len(X), np.float32)
        self._check(self.lib.finess_mlp_eval(self.h, _fptr(X), len(X), int(path), _fptr(z)), "eval")
        return z


class Batch:
    """Many independent 2D problems (``finess_batch``): same physics, scheme
    and controller settings as ``params``, each on the same domain with its
    own grid; all advanced together by persistent launches captured in a HIP
    graph, every problem under its own CFL controller."""

    def __init__(self, params, sizes):
        self.lib = load_library()
        self.p = params.copy()
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        n = len(self.sizes)
        mx = (ctypes.c_int * n)(*[a for a, _ in self.sizes])
        my = (ctypes.c_int * n)(*[b for _, b in self.sizes])
        h = ctypes.c_void_p()
        rc = self.lib.finess_batch_create(ctypes.byref(self.p), n, mx, my, ctypes.byref(h))
        if rc:
            raise FinessError(rc, "finess_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.finess_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise FinessError(rc, "%s: %s" % (what, self.lib.finess_batch_last_error(self.h).decode()))

    def shape_q(self, k):
        g = self.p.mbc
        mx, my = self.sizes[k]
        return (mx + 2 * g, my + 2 * g, self.p.meqn)

    def set_state(self, k, q, t=0.0, dt=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.shape == self.shape_q(k)
        dt = self.p.initial_dt if dt is None else dt
        self._check(self.lib.finess_batch_set_state(self.h, int(k), _ptr(q), float(t), float(dt)), "set_state")

    def get_state(self, k):
        q = np.zeros(self.shape_q(k))
        t, dt = ctypes.c_double(), ctypes.c_double()
        acc, rej = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.lib.finess_batch_get_state(self.h, int(k), _ptr(q), ctypes.byref(t), ctypes.byref(dt),
                                                    ctypes.byref(acc), ctypes.byref(rej)), "get_state")
        return q, t.value, dt.value, acc.value, rej.value

    def autotune(self, trial_launches=5):
        ti, tj = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.finess_batch_autotune(self.h, int(trial_launches), ctypes.byref(ti), ctypes.byref(tj)),
                    "autotune")
        return ti.value, tj.value

    def run(self, steps):
        sec = ctypes.c_double()
        self._check(self.lib.finess_batch_run(self.h, int(steps), ctypes.byref(sec)), "run")
        return sec.value
