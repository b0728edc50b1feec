"""Initial data of the workloads (the apps' QinitFunc), sampled at cell
centres x_i = xlow + (i - 1/2) dx exactly as SampleFunction does
(lib/2d/SampleFunction.cpp, lib/2d/ConstructL.cpp:98-99), returned as
reference-layout arrays with every ghost cell (corners included) filled by
sampling the same function -- RunFinpack samples QinitFunc on the whole
grid including ghosts (lib/2d/RunFinpack.cpp:77-80).

None of these problems exists in the reference; they are new apps written to
the reference's app interface.  Seeded randomness uses splitmix64 so that the
CPU oracle and the GPU path see bit-identical inputs.
"""
import math

import numpy as np

from . import (BC_EXTRAP, BC_PERIODIC, BC_REFLECT, PHYS_ADVECTION, PHYS_EULER, PHYS_SHALLOW_WATER, TS_LXW, TS_RK,
               default_params)

GAMMA = 1.4


def splitmix64(seed, n):
    """n uniform doubles in [0,1) from splitmix64 (top 53 bits)."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cell_centres(p):
    g = p.mbc
    xs = p.xlow + (np.arange(-g, p.mx + g) + 0.5) * p.dx
    if p.ndims == 1:
        return xs
    ys = p.ylow + (np.arange(-g, p.my + g) + 0.5) * p.dy
    return np.meshgrid(xs, ys, indexing="ij")


def euler_conserved(rho, u, v, pr, gamma=GAMMA):
    E = pr / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, np.zeros_like(rho), E], axis=-1)


# ---------------------------------------------------------------------------
# workload 1: 1D linear advection, sin(2 pi x), periodic
# ---------------------------------------------------------------------------
def advection_1d(mx=200, cfl=0.4, **kw):
    args = dict(ndims=1, mx=mx, my=1, mbc=3, xlow=0.0, xhigh=1.0, meqn=1, maux=1, physics=PHYS_ADVECTION,
                bc=[BC_PERIODIC] * 4, time_order=3, desired_cfl=cfl, max_cfl=max(cfl, 0.45),
                initial_dt=cfl / mx, max_dt=1.0, nv=10 ** 7)
    args.update(kw)
    p = default_params(**args)
    x = cell_centres(p)
    q = np.sin(2 * math.pi * x)[:, None].copy()
    aux = np.ones((p.mx + 2 * p.mbc, 1))
    return p, q, aux


# ---------------------------------------------------------------------------
# workload 2: 2D Euler four-quadrant Riemann problem, transmissive
# ---------------------------------------------------------------------------
QUADRANTS = {  # (rho, u, v, p)
    "NE": (1.5, 0.0, 0.0, 1.5),
    "NW": (0.5323, 1.206, 0.0, 0.3),
    "SW": (0.138, 1.206, 1.206, 0.029),
    "SE": (0.5323, 0.0, 1.206, 0.3),
}


def riemann4(mx=1024, my=None, cfl=0.4, **kw):
    my = mx if my is None else my
    args = dict(ndims=2, mx=mx, my=my, mbc=3, meqn=5, maux=1, physics=PHYS_EULER, gamma=GAMMA,
                bc=[BC_EXTRAP] * 4, time_order=3, desired_cfl=cfl, max_cfl=0.45,
                initial_dt=cfl * (1.0 / mx) / 2.1, max_dt=1.0, nv=10 ** 7)
    args.update(kw)
    p = default_params(**args)
    X, Y = cell_centres(p)
    rho = np.empty_like(X); u = np.empty_like(X); v = np.empty_like(X); pr = np.empty_like(X)
    for key, (r0, u0, v0, p0) in QUADRANTS.items():
        m = ((Y >= 0.5) if key[0] == "N" else (Y < 0.5)) & ((X >= 0.5) if key[1] == "E" else (X < 0.5))
        rho[m], u[m], v[m], pr[m] = r0, u0, v0, p0
    return p, np.ascontiguousarray(euler_conserved(rho, u, v, pr)), np.zeros(p.shape_aux())


# ---------------------------------------------------------------------------
# smooth 2D Euler density wave (exact solution: translation), periodic
# ---------------------------------------------------------------------------
def density_wave(mx=64, my=None, u0=1.0, v0=0.5, amp=0.2, cfl=0.4, lxw=False, **kw):
    my = mx if my is None else my
    args = dict(ndims=2, mx=mx, my=my, mbc=5 if lxw else 3, meqn=5, maux=1, physics=PHYS_EULER, gamma=GAMMA,
                bc=[BC_PERIODIC] * 4, time_order=3, time_stepping_method=TS_LXW if lxw else TS_RK,
                desired_cfl=cfl, max_cfl=max(cfl, 0.45), initial_dt=cfl / mx / 3.0, max_dt=1.0, nv=10 ** 7)
    args.update(kw)
    p = default_params(**args)
    X, Y = cell_centres(p)
    return p, density_wave_state(X, Y, 0.0, u0, v0, amp), np.zeros(p.shape_aux())


def density_wave_state(X, Y, t, u0=1.0, v0=0.5, amp=0.2):
    rho = 1.0 + amp * np.sin(2 * math.pi * (X - u0 * t + Y - v0 * t))
    return np.ascontiguousarray(euler_conserved(rho, np.full_like(X, u0), np.full_like(X, v0), np.ones_like(X)))


# ---------------------------------------------------------------------------
# workload 3: 2D Euler Kelvin-Helmholtz, periodic, Lax-Wendroff (mbc = 5)
# ---------------------------------------------------------------------------
def kelvin_helmholtz(mx=2048, my=None, seed=20240611, cfl=0.4, lxw=True, **kw):
    my = mx if my is None else my
    args = dict(ndims=2, mx=mx, my=my, mbc=5, meqn=5, maux=1, physics=PHYS_EULER, gamma=GAMMA,
                bc=[BC_PERIODIC] * 4, time_order=3, time_stepping_method=TS_LXW if lxw else TS_RK,
                desired_cfl=cfl, max_cfl=max(cfl, 0.45), initial_dt=cfl / mx / 2.5, max_dt=1.0, nv=10 ** 7)
    args.update(kw)
    p = default_params(**args)
    X, Y = cell_centres(p)
    phi = 2 * math.pi * splitmix64(seed, 8)
    band = np.abs(Y - 0.5) < 0.25
    rho = np.where(band, 2.0, 1.0)
    u = np.where(band, 0.5, -0.5)
    pert = np.zeros_like(X)
    for k in range(1, 9):
        pert += np.sin(2 * math.pi * k * X + phi[k - 1])
    v = 0.01 * pert * np.exp(-(((np.abs(Y - 0.5) - 0.25) / 0.05) ** 2))
    return p, np.ascontiguousarray(euler_conserved(rho, u, v, np.full_like(X, 2.5))), np.zeros(p.shape_aux())


# 2D Euler near-vacuum test for the positivity-preserving limiter: the four
# quadrants move apart diagonally (rho = 1, p = 0.4, |u| = |v| = u0), the 2D
# counterpart of the 1D double rarefaction; Lax-Wendroff (mbc = 5)
# ---------------------------------------------------------------------------
def double_rarefaction(mx=64, my=None, u0=2.0, cfl=0.3, **kw):
    my = mx if my is None else my
    args = dict(ndims=2, mx=mx, my=my, mbc=5, meqn=5, maux=1, physics=PHYS_EULER, gamma=GAMMA,
                bc=[BC_EXTRAP] * 4, time_order=3, time_stepping_method=TS_LXW, desired_cfl=cfl,
                max_cfl=max(cfl, 0.45), initial_dt=cfl / max(mx, my) / (u0 + 1.0) / 2.0, max_dt=1.0, nv=10 ** 7)
    args.update(kw)
    p = default_params(**args)
    X, Y = cell_centres(p)
    u = np.where(X < 0.5, -u0, u0)
    v = np.where(Y < 0.5, -u0, u0)
    return p, np.ascontiguousarray(euler_conserved(np.ones_like(X), u, v, np.full_like(X, 0.4))), np.zeros(p.shape_aux())


# 2D Euler Sedov blast as apps/2d/euler/sedov_blast sets it up
# (parameters.ini, QinitFunc.cpp, AfterQinit.cpp): [0, 1.1]^2, rho = 1, zero
# velocity, energy 1e-12, and 0.979264 / (dx dy) in the corner cell (1, 1);
# reflecting left and bottom, extrapolation right and top; Lax-Wendroff with
# the positivity-preserving limiter
# ---------------------------------------------------------------------------
def sedov_blast(mx=160, my=None, cfl=0.4, mpp_limiter=1, **kw):
    my = mx if my is None else my
    args = dict(ndims=2, mx=mx, my=my, mbc=5, meqn=5, maux=1, physics=PHYS_EULER, gamma=GAMMA,
                xlow=0.0, xhigh=1.1, ylow=0.0, yhigh=1.1, bc=[BC_REFLECT, BC_EXTRAP, BC_REFLECT, BC_EXTRAP],
                time_order=3, time_stepping_method=TS_LXW, desired_cfl=cfl, max_cfl=0.5, initial_dt=1e-5,
                max_dt=1.0e10, nv=100000, mpp_limiter=mpp_limiter)
    args.update(kw)
    p = default_params(**args)
    q = np.zeros(p.shape_q())
    q[..., 0] = 1.0
    q[..., 4] = 1.0e-12
    g = p.mbc
    q[g, g, 4] = 0.979264 / (p.dx * p.dy)
    return p, q, np.zeros(p.shape_aux())


# ---------------------------------------------------------------------------
# workload 4 member: 2D shallow water Gaussian bump, periodic (g = 1)
# ---------------------------------------------------------------------------
def sw_bump(nx=64, ny=64, A=0.3, cx=0.5, cy=0.5, s=0.1, cfl=0.4, **kw):
    args = dict(ndims=2, mx=nx, my=ny, mbc=3, meqn=3, maux=1, physics=PHYS_SHALLOW_WATER,
                bc=[BC_PERIODIC] * 4, time_order=3, desired_cfl=cfl, max_cfl=max(cfl, 0.45),
                initial_dt=cfl / max(nx, ny) / 2.0, max_dt=1.0, nv=10 ** 7)
    args.update(kw)
    p = default_params(**args)
    X, Y = cell_centres(p)
    h = 1.0 + A * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / s ** 2)
    q = np.stack([h, np.zeros_like(h), np.zeros_like(h)], axis=-1)
    return p, np.ascontiguousarray(q), np.zeros(p.shape_aux())


def sw_batch_members(n=512, seed=7, sizes=(64, 96, 128, 192, 256, 384)):
    """Workload-4 draws: (nx, ny, A, cx, cy, s) per member, seeded."""
    u = splitmix64(seed, 6 * n).reshape(n, 6)
    out = []
    for k in range(n):
        nx = sizes[int(u[k, 0] * len(sizes))]
        ny = sizes[int(u[k, 1] * len(sizes))]
        A = 0.1 + 0.4 * u[k, 2]
        cx = 0.3 + 0.4 * u[k, 3]
        cy = 0.3 + 0.4 * u[k, 4]
        s = 0.05 + 0.10 * u[k, 5]
        out.append((nx, ny, A, cx, cy, s))
    return out


# ---------------------------------------------------------------------------
# workload 5: 2D Euler Mach-3 shock / light bubble on [0,4]x[0,1]
# ---------------------------------------------------------------------------
def shock_bubble(mx=4096, my=1024, seed=99, cfl=0.4, **kw):
    args = dict(ndims=2, mx=mx, my=my, mbc=3, xlow=0.0, xhigh=4.0, ylow=0.0, yhigh=1.0, meqn=5, maux=1,
                physics=PHYS_EULER, gamma=GAMMA, bc=[BC_EXTRAP, BC_EXTRAP, BC_REFLECT, BC_REFLECT],
                time_order=3, desired_cfl=cfl, max_cfl=0.45, initial_dt=cfl * (4.0 / mx) / 8.0, max_dt=1.0,
                nv=10 ** 7)
    args.update(kw)
    p = default_params(**args)
    X, Y = cell_centres(p)
    g = GAMMA
    rho2 = (g + 1) * 9.0 / ((g - 1) * 9.0 + 2.0)
    p2 = 1.0 + 2 * g / (g + 1) * (9.0 - 1.0)
    S = 3.0 * math.sqrt(g)
    u2 = S * (1.0 - 1.0 / rho2)
    rho = np.ones_like(X)
    u = np.zeros_like(X)
    pr = np.ones_like(X)
    post = X < 0.5
    rho[post], u[post], pr[post] = rho2, u2, p2
    bub = (X - 1.0) ** 2 + (Y - 0.5) ** 2 < 0.2 ** 2
    rho[bub] = 0.1
    noise = splitmix64(seed, X.size).reshape(X.shape)
    rho = rho * (1.0 + 0.02 * (noise - 0.5))
    return p, np.ascontiguousarray(euler_conserved(rho, u, np.zeros_like(X), pr)), np.zeros(p.shape_aux())
