"""finess_construct_l_planes_noalloc, the allocation-free ConstructL on
device field planes, against the context entry point finess_construct_l on
the same state: L of the interior and the two wave-speed maxima GetCFL takes
(lib/2d/GetCFL.cpp:54-58: faces mx+1/2 and my+1/2 excluded).  Both run the
same stage kernel, so the bar is 1e-13 relative L-inf per field (the FP64
parity bar of one operator call is 1e-12).  Two calls in a row share the
device's scratch array; the second is smaller, so stale maxima words of the
first would show in its wave speeds."""
import ctypes

import numpy as np
import pytest

import orc
from finess_amd import BC_PERIODIC, BC_REFLECT, Solver, load_library, problems

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)


def _context_result(p, q):
    orc.set_bnd(p, q)
    with Solver(p) as s:
        L, sm = s.construct_l(q)
    g = p.mbc
    Lp = np.ascontiguousarray(np.moveaxis(orc.interior(p, L), -1, 0))
    sx = sm[g:g + p.mx, g:g + p.my, 0].max()
    sy = sm[g:g + p.mx, g:g + p.my, 1].max()
    return Lp, np.array([sx, sy])


class _Hip:
    """Device buffers through the HIP runtime the library itself links
    (the test allocates; the entry point under test does not)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(d), a.nbytes) == 0
        self.bufs.append(d)
        assert self.rt.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return d

    def download(self, d, shape):
        a = np.empty(shape, dtype=np.float64)
        assert self.rt.hipMemcpy(a.ctypes.data, d, a.nbytes, 2) == 0  # device to host
        return a

    def close(self):
        for d in self.bufs:
            self.rt.hipFree(d)
        self.bufs = []


def _noalloc_result(p, q):
    lib = load_library()
    qp = np.ascontiguousarray(np.moveaxis(orc.interior(p, q), -1, 0))
    hip = _Hip()
    try:
        qd = hip.upload(qp)
        Ld = hip.upload(np.full(qp.shape, np.nan))
        sd = hip.upload(np.full(2, np.nan))
        rc = lib.finess_construct_l_planes_noalloc(ctypes.byref(p), qd, Ld, sd, None)
        assert rc == 0, rc
        assert hip.rt.hipDeviceSynchronize() == 0
        return hip.download(Ld, qp.shape), hip.download(sd, (2,))
    finally:
        hip.close()


def _compare(p, q):
    L0, s0 = _context_result(p, q)
    L1, s1 = _noalloc_result(p, q)
    err = orc.rel_linf(np.moveaxis(L1, 0, -1), np.moveaxis(L0, 0, -1))
    print("noalloc vs construct_l: L rel err", err, "speeds", s1, s0, "identical", np.array_equal(L1, L0))
    assert np.all(np.isfinite(L1))
    assert np.all(err <= 1e-13), err
    assert np.allclose(s1, s0, rtol=1e-15, atol=0.0), (s1, s0)


def test_noalloc_matches_construct_l_two_calls():
    # first call: Euler, 96 x 80, extrapolation, global_alpha on
    p, q, _ = problems.riemann4(mx=96, my=80, global_alpha=1)
    q = q.copy()
    q[..., 0] *= 1.0 + 0.02 * RNG.random(q.shape[:-1])
    _compare(p, q)
    # second call, smaller: Euler, 40 x 36, periodic in x, reflecting in y
    p, q, _ = problems.riemann4(mx=40, my=36, bc=[BC_PERIODIC, BC_PERIODIC, BC_REFLECT, BC_REFLECT])
    q = q.copy()
    q[..., 0] *= 1.0 + 0.02 * RNG.random(q.shape[:-1])
    _compare(p, q)


def test_noalloc_shallow_water_fp32():
    from finess_amd import PREC_FP32
    p, q, _ = problems.sw_bump(nx=45, ny=29, precision=PREC_FP32)
    q = q.copy()
    q[..., 1] = 0.1 * RNG.random(q.shape[:-1])
    q[..., 2] = 0.1 * RNG.random(q.shape[:-1])
    _compare(p, q)


def test_noalloc_rejects_what_it_does_not_run():
    lib = load_library()
    p, _, _ = problems.riemann4(mx=32, my=32, space_order=7, mbc=4)
    z = ctypes.c_void_p(0)
    # a null pointer and an unsupported order are refused before any launch
    assert lib.finess_construct_l_planes_noalloc(ctypes.byref(p), z, z, z, z) == -7
    dummy = ctypes.c_void_p(16)
    assert lib.finess_construct_l_planes_noalloc(ctypes.byref(p), dummy, dummy, dummy, z) == -1
