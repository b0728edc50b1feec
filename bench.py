#!/usr/bin/env python3
"""Benchmark of the FINESS WENO path on one MI355X: learned+FP32 against
plain FP64 at equal L1 density error.

Workload (one GPU): 2D Euler four-quadrant Riemann problem on [0,1]^2,
1024 x 1024 cells, gamma 1.4, characteristic-wise WENO5-JS (eps 1e-6, p 2)
with local Lax-Friedrichs splitting, SSP-RK3, a-posteriori CFL controller at
CFL 0.4, transmissive boundaries, t_final 0.3.  Two forms of the same solver:

  plain FP64    full characteristic WENO at every interface, FP64 throughout;
  learned+FP32  the troubled-interface MLP (24-32-32-1, BF16 on the matrix
                cores, FP32 accumulation; trained on the GPU before the timed
                region on interfaces harvested from this problem's run and a
                smooth density wave) routes smooth interfaces to the
                component-wise linear FD5 flux and the rest to full
                characteristic WENO; FP32 stencil arithmetic, FP64 state,
                stage accumulation and pressure recovery.  The indicator runs
                at the first stage of every fourth step; the stages in between
                reuse its routing record.

A "step" is one accepted time step: three stages (plain: one fused kernel
each; learned: the troubled-face pass and the lookup stage kernel, with the
indicator kernel in front at every fourth step) and the time-step
controller, which runs on the device after each step (in the learned form
inside the next step's troubled-face pass, or on its own before an
indicator step).  Both forms are timed over
K steps from the state at t = 0.15 (the middle of the run) after W warm-up
steps; inputs are resident in HBM before the timed region starts.  The L1
density error of each form at t_final is measured against a 2048 x 2048
plain-FP64 run averaged onto the 1024 grid.

Prints one JSON line.  `value` is the learned+FP32 Mcell-updates/s
(cells * steps / seconds); the plain FP64 figures and the error ratio sit
beside it.  `roofline` refers to the stage of the learned form (the kernels
of one stage together, averaged over the stages of a 4-step indicator
interval): achieved = algorithmic HBM bytes per stage / average stage
duration, measured with HIP events on the stream the kernels run on.  `cpu_baseline` is the CPU restatement of the
reference (oracle/) timed on this host's cores on a bounded sample of the
same workload.

Two further workloads of the same solver are reported beside the headline
(skip them with --no-extra): the third-order Lax-Wendroff form on the
periodic Kelvin-Helmholtz problem at 2048 x 2048 (FP64 against the FP32
stencil form), and the Mach-3 shock-bubble at 4096 x 1024 (learned+FP32 with
the same indicator, which never saw this problem, against plain FP64).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 320.0   # SSP-RK3: 8 state passes x 5 fields x 8 B (FP64 state in both forms)
STAGES_PER_STEP = 3
# HBM bytes per stage launch at 1024x1024 measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs, calibrated against a 256 MiB device copy (commands and raw
# counts in profiles/k_stage2d_lookup_euler_fp32_riemann4_1024.md and
# profiles/k_stage2d_euler_fp64_riemann4_1024.md)
TRAFFIC_ROUTED_1024 = 134.0e6   # mean over the 12 stages of a 4-step indicator interval
TRAFFIC_PLAIN_1024 = 114.1e6
# VALU instructions per launch at 1024x1024 (rocprofv3 SQ_INSTS_VALU and the
# SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters, same profile files): a wave64
# VALU instruction holds a SIMD for 4 cycles, so at 100% issue a launch takes
# N x 4 / (1024 SIMDs x 2.4 GHz).  The bench divides that by the measured
# launch time.
VALU_PLAIN_1024 = 40.55e6        # k_stage2d, FP64
VALU_F64_PLAIN_1024 = 31.13e6    # of which FP64 arithmetic
SIMDS, CLOCK_HZ = 1024, 2.4e9
T_FINAL = 0.3
T_TIMED = 0.15
# the indicator runs at the first stage of every ROUTE_INTERVAL-th step; the
# record serves the later stages and the following steps (waves move at most
# 0.45 cells per step, so a record is at most 1.8 cells stale)
ROUTE_INTERVAL = 4


def host_cpu():
    """The host's CPU as lscpu reports it: model, sockets, physical cores, and
    the cores this process may run on."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    sockets = int(info.get("Socket(s)", "1") or 1)
    per = int(info.get("Core(s) per socket", "0") or 0)
    return {"model": info.get("Model name", "unknown"), "sockets": sockets, "cores_per_socket": per,
            "physical_cores": sockets * per if per else None,
            "threads_per_core": int(info.get("Thread(s) per core", "1") or 1),
            "usable_cpus": len(os.sched_getaffinity(0))}


def _oracle(threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc  # the oracle loader lives with the tests
    orc.oracle().orc_set_num_threads(threads)
    return orc


def _time_fixed(orc, p, q, dt, steps, aux=None):
    t0 = time.perf_counter()
    orc.fixed_steps(p, q, dt, steps, aux=aux)
    return time.perf_counter() - t0


def cpu_baseline(mx, steps, threads, others=True):
    """Times the oracle's fixed-step loop (no controller) on `steps` steps of
    the plain FP64 form of the headline workload at full size with `threads`
    OpenMP threads, and on a one-thread sample; with `others`, bounded samples
    of the other workloads beside it."""
    from finess_amd import TS_LXW, problems
    orc = _oracle(threads)
    p, q, _ = problems.riemann4(mx=mx)
    dt = 0.4 / mx / 2.1
    el = _time_fixed(orc, p, q, dt, steps, aux=np.zeros(p.shape_aux()))
    out = {
        "value": round(mx * mx * steps / el / 1e6, 3),
        "unit": "Mcell-updates/s",
        "cores": int(orc.oracle().orc_num_threads()),
        "kind": "port",
        "sample": "%d SSP-RK3 step(s) of the %dx%d four-quadrant problem, plain FP64, fixed dt, %d OpenMP "
                  "threads, %.1f s" % (steps, mx, mx, threads, el),
        "host": host_cpu(),
    }
    orc.oracle().orc_set_num_threads(1)
    p, q, _ = problems.riemann4(mx=mx)
    el1 = _time_fixed(orc, p, q, dt, 2, aux=np.zeros(p.shape_aux()))
    out["one_thread"] = {"value": round(mx * mx * 2 / el1 / 1e6, 3), "sample": "2 steps, 1 thread, %.1f s" % el1}
    if others:
        o = {}
        # workload 1: the whole 500-step run (one thread: 200 cells)
        p, q, aux = problems.advection_1d(mx=200)
        t0 = time.perf_counter()
        orc.fin_solve(p, q, 0.0, 1.0, np.array([0.0, p.initial_dt, p.max_dt]), aux=aux)
        e = time.perf_counter() - t0
        o["advection_1d_200"] = {"value": round(200 * 500 / e / 1e6, 3), "threads": 1,
                                 "sample": "the full run, 500 steps, %.3f s" % e}
        orc.oracle().orc_set_num_threads(threads)
        p, q, _ = problems.kelvin_helmholtz(mx=2048)
        e = _time_fixed(orc, p, q, 0.3 / 2048 / 2.37, 2)
        o["kelvin_helmholtz_lxw_2048"] = {"value": round(2048 * 2048 * 2 / e / 1e6, 3), "threads": threads,
                                          "sample": "2 Lax-Wendroff steps at 2048^2, %.1f s" % e}
        draws = problems.sw_batch_members(n=512)[:16]
        cells, e = 0, 0.0
        for nx, ny, A, cx, cy, sg in draws:
            p, q, _ = problems.sw_bump(nx=nx, ny=ny, A=A, cx=cx, cy=cy, s=sg)
            e += _time_fixed(orc, p, q, p.initial_dt, 20)
            cells += nx * ny
        o["batched_shallow_water"] = {"value": round(cells * 20 / e / 1e6, 3), "threads": threads,
                                      "sample": "16 of the 512 members, 20 steps each, one after another, %.1f s" % e}
        p, q, _ = problems.shock_bubble(mx=4096, my=1024)
        e = _time_fixed(orc, p, q, 0.4 / 1024 / 4.6, 2)
        o["shock_bubble_4096x1024"] = {"value": round(4096 * 1024 * 2 / e / 1e6, 3), "threads": threads,
                                       "sample": "2 steps, plain FP64, %.1f s" % e}
        out["other_workloads"] = o
    return out


def make_solver(form, mx, indicator=None):
    from finess_amd import PREC_FP32, ROUTE_LEARNED, Solver, problems
    kw = {} if form == "plain" else dict(precision=PREC_FP32, routing=ROUTE_LEARNED, route_per_step=ROUTE_INTERVAL)
    p, q, _ = problems.riemann4(mx=mx, **kw)
    s = Solver(p)
    s.set_state(q)
    if indicator is not None:
        indicator.install(s)
    return p, s


def full_run(form, mx, indicator=None):
    """Runs to t_final; returns the interior density, wall seconds and steps."""
    p, s = make_solver(form, mx, indicator)
    dtv = np.array([0.0, p.initial_dt, p.max_dt])
    s.sync()
    t0 = time.perf_counter()
    st = s.advance(T_FINAL, dtv)
    s.sync()
    el = time.perf_counter() - t0
    route = s.route_stats() if indicator is not None else None
    q, _ = s.get_state()
    s.close()
    g = p.mbc
    return q[g:g + mx, g:g + mx, 0].copy(), el, int(st.steps_accepted), route


def timed_steps(form, mx, steps, warmup, indicator=None):
    p, s = make_solver(form, mx, indicator)
    dtv = np.array([0.0, p.initial_dt, p.max_dt])
    s.advance(T_TIMED, dtv)
    s.advance_steps(warmup, dtv)
    s.sync()
    t0 = time.perf_counter()
    st = s.advance_steps(steps, dtv)
    s.sync()
    el = time.perf_counter() - t0
    # dominant kernel: average stage-launch duration over back-to-back steps,
    # HIP events on the context's stream
    s.run_fixed_steps(float(dtv[1]), 30)
    ms, launches = s.last_kernel_time()
    s.close()
    return el, int(st.steps_rejected), ms / 1e3 / launches


def batched_shallow_water(n=512, steps=200):
    """512 independent periodic shallow-water problems (grids drawn from
    {64,...,384}^2), WENO5 SSP-RK3, FP32 stencil + FP64 state, per-problem
    CFL controller, one persistent launch per stage in a HIP graph, tile
    shape from the autotuner.  The same batch in plain FP64 beside it: its
    throughput and the FP32 form's error against it (relative L1 of the
    depth per member after the same accepted steps; the members' final
    times differ by the controllers' dt rounding, also reported)."""
    from finess_amd import PREC_FP32, PREC_FP64, Batch, problems
    draws = problems.sw_batch_members(n=n)
    sizes = [(d[0], d[1]) for d in draws]
    cells = sum(a * c for a, c in sizes)
    res = {}
    for name, prec in (("fp32", PREC_FP32), ("fp64", PREC_FP64)):
        mem = [problems.sw_bump(nx=nx, ny=ny, A=A, cx=cx, cy=cy, s=s, precision=prec)[:2]
               for nx, ny, A, cx, cy, s in draws]
        with Batch(mem[0][0], sizes) as b:
            for k, (p, q) in enumerate(mem):
                b.set_state(k, q, 0.0, p.initial_dt)
            tile = b.autotune(5)
            sec = b.run(steps)
            fin = [b.get_state(k) for k in range(n)]
        g = mem[0][0].mbc
        res[name] = {"sec": sec, "tile": tile, "rej": sum(f[4] for f in fin),
                     "h": [f[0][g:g + sz[0], g:g + sz[1], 0] for f, sz in zip(fin, sizes)],
                     "t": np.array([f[1] for f in fin])}
    rel = [float(np.sum(np.abs(a - b)) / np.sum(np.abs(b))) for a, b in zip(res["fp32"]["h"], res["fp64"]["h"])]
    r32 = res["fp32"]
    return {"value": round(cells * steps / r32["sec"] / 1e6, 1), "unit": "Mcell-updates/s", "problems": n,
            "steps_per_problem": steps, "cells": cells, "seconds": round(r32["sec"], 4), "tile": list(r32["tile"]),
            "steps_rejected": int(r32["rej"]),
            "fp64": {"value": round(cells * steps / res["fp64"]["sec"] / 1e6, 1), "tile": list(res["fp64"]["tile"]),
                     "steps_rejected": int(res["fp64"]["rej"])},
            "fp32_vs_fp64": {"rel_l1_depth_max": float(np.max(rel)), "rel_l1_depth_mean": float(np.mean(rel)),
                             "max_final_time_difference": float(np.max(np.abs(r32["t"] - res["fp64"]["t"])))}}


def _interior_density(p, q):
    g = p.mbc
    return q[g:g + p.mx, g:g + p.my, 0].copy()


def kelvin_helmholtz_lxw(mx=2048, steps=20, warmup=3, t_final=1.0):
    """Third-order Lax-Wendroff, periodic Kelvin-Helmholtz at mx^2: the FP64
    and the FP32-stencil forms (FP64 state and accumulation) timed over
    `steps` accepted steps after `warmup`, from t = 0, and the relative L1
    density difference between the two after those steps."""
    from finess_amd import PREC_FP32, Solver, problems
    out, rho, tt = {}, {}, {}
    for name, kw in (("fp64", {}), ("fp32", dict(precision=PREC_FP32))):
        p, q, _ = problems.kelvin_helmholtz(mx=mx, **kw)
        with Solver(p) as s:
            s.set_state(q)
            dtv = np.array([0.0, p.initial_dt, p.max_dt])
            s.advance_steps(warmup, dtv)
            s.sync()
            t0 = time.perf_counter()
            st = s.advance_steps(steps, dtv)
            s.sync()
            el = time.perf_counter() - t0
            qf, tt[name] = s.get_state()
        rho[name] = _interior_density(p, qf)
        out[name] = {"value": round(mx * mx * steps / el / 1e6, 1), "ms_per_step": round(el / steps * 1e3, 4),
                     "steps_rejected": int(st.steps_rejected)}
    out["speedup_fp32"] = round(out["fp32"]["value"] / out["fp64"]["value"], 4)
    out["l1_rel_density_fp32_vs_fp64"] = float(np.mean(np.abs(rho["fp32"] - rho["fp64"])) / np.mean(np.abs(rho["fp64"])))
    out["t"] = {k: round(v, 6) for k, v in tt.items()}
    out["workload"] = "2D Euler Kelvin-Helmholtz %dx%d periodic, LxW order 3, WENO5-JS, CFL 0.4, %d steps after %d" % (
        mx, mx, steps, warmup)
    if t_final:
        # both forms from t = 0 to t_final (the roll-up): wall time and the
        # density difference between them
        fin = {}
        for name, kw in (("fp64", {}), ("fp32", dict(precision=PREC_FP32))):
            p, q, _ = problems.kelvin_helmholtz(mx=mx, **kw)
            with Solver(p) as s:
                s.set_state(q)
                s.sync()
                t0 = time.perf_counter()
                st = s.advance(t_final, np.array([0.0, p.initial_dt, p.max_dt]))
                s.sync()
                el = time.perf_counter() - t0
                qf, _ = s.get_state()
            rho[name] = _interior_density(p, qf)
            fin[name] = {"seconds": round(el, 2), "steps": int(st.steps_accepted),
                         "value": round(mx * mx * st.steps_accepted / el / 1e6, 1)}
        fin["l1_rel_density_fp32_vs_fp64"] = float(
            np.mean(np.abs(rho["fp32"] - rho["fp64"])) / np.mean(np.abs(rho["fp64"])))
        out["to_t_final"] = dict(t_final=t_final, **fin)
    return out


def shock_bubble_learned(ind, mx=4096, my=1024, steps=40, warmup=5, t_final=0.2):
    """Mach-3 shock-bubble at mx x my: learned+FP32 (routing record every
    ROUTE_INTERVAL steps) against plain FP64, timed over `steps` accepted
    steps after `warmup`, from t = 0; the relative L1 density difference
    between the forms after those steps and the routed fraction."""
    from finess_amd import PREC_FP32, ROUTE_LEARNED, Solver, problems
    out, rho = {}, {}
    for name, kw in (("plain_fp64", {}),
                     ("learned_fp32", dict(precision=PREC_FP32, routing=ROUTE_LEARNED, route_per_step=ROUTE_INTERVAL))):
        p, q, _ = problems.shock_bubble(mx=mx, my=my, **kw)
        with Solver(p) as s:
            s.set_state(q)
            if name == "learned_fp32":
                ind.install(s)
            dtv = np.array([0.0, p.initial_dt, p.max_dt])
            s.advance_steps(warmup, dtv)
            s.sync()
            t0 = time.perf_counter()
            st = s.advance_steps(steps, dtv)
            s.sync()
            el = time.perf_counter() - t0
            route = s.route_stats() if name == "learned_fp32" else None
            qf, t = s.get_state()
        rho[name] = _interior_density(p, qf)
        out[name] = {"value": round(mx * my * steps / el / 1e6, 1), "ms_per_step": round(el / steps * 1e3, 4),
                     "steps_rejected": int(st.steps_rejected), "t": round(t, 6)}
        if route is not None:
            out[name]["troubled_fraction"] = round(route[0] / max(route[1], 1), 4)
    out["speedup_vs_plain_fp64"] = round(out["learned_fp32"]["value"] / out["plain_fp64"]["value"], 4)
    out["l1_rel_density_learned_vs_plain"] = float(
        np.mean(np.abs(rho["learned_fp32"] - rho["plain_fp64"])) / np.mean(np.abs(rho["plain_fp64"])))
    if t_final:
        # both forms from t = 0 to t_final: wall time, the density difference
        # of the learned form from the plain FP64 one, the routed share
        fin = {}
        for name, kw in (("plain_fp64", {}),
                         ("learned_fp32", dict(precision=PREC_FP32, routing=ROUTE_LEARNED,
                                               route_per_step=ROUTE_INTERVAL))):
            p, q, _ = problems.shock_bubble(mx=mx, my=my, **kw)
            with Solver(p) as s:
                s.set_state(q)
                if name == "learned_fp32":
                    ind.install(s)
                s.sync()
                t0 = time.perf_counter()
                st = s.advance(t_final, np.array([0.0, p.initial_dt, p.max_dt]))
                s.sync()
                el = time.perf_counter() - t0
                route = s.route_stats() if name == "learned_fp32" else None
                qf, _ = s.get_state()
            rho[name] = _interior_density(p, qf)
            fin[name] = {"seconds": round(el, 2), "steps": int(st.steps_accepted),
                         "value": round(mx * my * st.steps_accepted / el / 1e6, 1)}
            if route is not None:
                fin[name]["troubled_fraction"] = round(route[0] / max(route[1], 1), 4)
        fin["l1_rel_density_learned_vs_plain"] = float(
            np.mean(np.abs(rho["learned_fp32"] - rho["plain_fp64"])) / np.mean(np.abs(rho["plain_fp64"])))
        out["to_t_final"] = dict(t_final=t_final, **fin)
    out["workload"] = ("2D Euler Mach-3 shock-bubble %dx%d, WENO5-JS + LLF, SSP-RK3, CFL 0.4, %d steps after %d; "
                       "indicator trained on the four-quadrant problem and a density wave" % (mx, my, steps, warmup))
    return out


def at_2048(ind, steps=20, warmup=3, accuracy=True):
    """Both forms on the four-quadrant problem at 2048^2 (timed from t = 0.15
    like the headline): Mcell-updates/s and each stage's fraction of the HBM
    roofline (algorithmic bytes over its event-timed average launch); with
    `accuracy`, the L1 density errors at t_final against a 4096^2 plain FP64
    run averaged 2 x 2."""
    mx = 2048
    cells = mx * mx
    bytes_per_launch = cells * BYTES_PER_CELL_UPDATE / STAGES_PER_STEP
    out = {}
    for form in ("plain", "learned"):
        el, rej, launch = timed_steps(form, mx, steps, warmup, ind if form == "learned" else None)
        out[form] = {"value": round(cells * steps / el / 1e6, 1), "ms_per_step": round(el / steps * 1e3, 4),
                     "steps": steps, "steps_rejected": rej, "avg_stage_launch_us": round(launch * 1e6, 2),
                     "hbm_frac": round(bytes_per_launch / launch / 1e9 / HBM_PEAK_GBS, 4)}
    res = {"plain_fp64": out["plain"], "learned_fp32": out["learned"],
           "speedup_vs_plain_fp64": round(out["learned"]["value"] / out["plain"]["value"], 4),
           "workload": "2D Euler four-quadrant Riemann 2048x2048, char-WENO5-JS + LLF, SSP-RK3"}
    if accuracy:
        ref, _, _, _ = full_run("plain", 2 * mx)
        ref = 0.25 * (ref[0::2, 0::2] + ref[1::2, 0::2] + ref[0::2, 1::2] + ref[1::2, 1::2])
        rp, _, _, _ = full_run("plain", mx)
        rl, _, _, route = full_run("learned", mx, ind)
        l1p, l1l = float(np.mean(np.abs(rp - ref))), float(np.mean(np.abs(rl - ref)))
        res["accuracy"] = {"reference": "plain FP64 at 4096x4096 to t=%.2f, 2x2-averaged" % T_FINAL,
                           "l1_density_plain_fp64": l1p, "l1_density_learned_fp32": l1l,
                           "l1_ratio": round(l1l / l1p, 5), "troubled_fraction": round(route[0] / max(route[1], 1), 4)}
    return res


def advection_1d_full(mx=200):
    """1D advection of sin(2 pi x), 200 cells, WENO5 + SSP-RK3 at CFL 0.4 to
    t = 1 (500 steps, the controller on the device): Mcell-updates/s and the L1
    error against the exact shifted solution."""
    from finess_amd import Solver, problems
    p, q, aux = problems.advection_1d(mx=mx, initial_dt=0.4 / mx)
    with Solver(p) as s:
        dtv = np.array([0.0, p.initial_dt, p.max_dt])
        qw = q.copy()
        s.fin_solve_rk(qw, 0.0, 0.05, dtv.copy(), aux=aux)  # warm-up (first launches)
        t0 = time.perf_counter()
        t = s.fin_solve_rk(q, 0.0, 1.0, dtv, aux=aux)
        el = time.perf_counter() - t0
    g = p.mbc
    x = p.xlow + (np.arange(mx) + 0.5) * p.dx
    l1 = float(np.mean(np.abs(q[g:g + mx, 0] - np.sin(2 * np.pi * (x - t)))))
    steps = int(round(1.0 / (0.4 / mx)))
    return {"value": round(mx * steps / el / 1e6, 2), "unit": "Mcell-updates/s", "seconds": round(el, 4),
            "steps": steps, "l1_error_vs_exact": l1,
            "workload": "1D advection sin(2 pi x), 200 cells, WENO5-JS + LLF, SSP-RK3, CFL 0.4, t = 1, FP64"}


def wide_order(order, version="js", mx=1024, steps=20, warmup=3):
    """The plain FP64 form with another WENO order (the face-tile kernels of
    stage2d_wide.hip) on the four-quadrant problem: Mcell-updates/s over fixed
    steps from t = 0 and the stage's event-timed average."""
    from finess_amd import WENO_FD, WENO_JS, WENO_Z, Solver, problems
    ver = {"js": WENO_JS, "z": WENO_Z, "fd": WENO_FD}[version]
    p, q, _ = problems.riemann4(mx=mx, space_order=order, weno_version=ver, mbc=max(3, (order + 1) // 2))
    with Solver(p) as s:
        s.set_state(q)
        dt = float(p.initial_dt)
        s.run_fixed_steps(dt, warmup)
        s.sync()
        t0 = time.perf_counter()
        s.run_fixed_steps(dt, steps)
        s.sync()
        el = time.perf_counter() - t0
        ms, launches = s.last_kernel_time()
    return {"workload": "2D Euler four-quadrant Riemann %dx%d, WENO-%s%d + LLF, SSP-RK3, plain FP64, fixed dt"
                        % (mx, mx, version.upper(), order),
            "value": round(mx * mx * steps / el / 1e6, 1), "unit": "Mcell-updates/s", "steps": steps,
            "avg_stage_us": round(ms / launches * 1e3, 2)}


def copy_calibration(mib=256, reps=5):
    """A plain device copy of known size, for scaling the FETCH_SIZE and
    WRITE_SIZE counters of a profiled run to bytes."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 8, "reps": reps}), flush=True)


def main():
    global ROUTE_INTERVAL
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mx", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the t_final runs and the L1 errors")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched shallow-water workload")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the Lax-Wendroff Kelvin-Helmholtz and the shock-bubble workloads")
    ap.add_argument("--route-interval", type=int, default=ROUTE_INTERVAL,
                    help="steps between indicator evaluations of the learned form")
    ap.add_argument("--train-epochs", type=int, default=10)
    ap.add_argument("--train-seeds", type=int, default=3,
                    help="networks trained from different initialisations; the one flagging the fewest faces of "
                         "the held-out snapshot at the recall target is kept")
    ap.add_argument("--train-stride", type=int, default=2, help="harvest every n-th interface of each snapshot")
    ap.add_argument("--save-indicator", metavar="PATH",
                    help="write the trained network (the drop-in layer's indicator file) after training")
    ap.add_argument("--indicator", metavar="PATH",
                    help="use this network instead of training one (a file of --save-indicator), so that a "
                         "profiled run times the default run's network without profiling its training")
    ap.add_argument("--wide", type=int, default=0, metavar="ORDER",
                    help="only time the plain FP64 form with WENO order 1, 7, 9 or 11 (face-tile kernels)")
    ap.add_argument("--wide-version", default="js", choices=["js", "z", "fd"])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (profiler counter calibration)")
    ap.add_argument("--only", choices=["kh_lxw", "batch", "advection_1d"],
                    help="only run one of the other workloads (for profiling its kernels)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.only:
        fn = {"kh_lxw": lambda: kelvin_helmholtz_lxw(steps=args.steps, warmup=args.warmup, t_final=0.0),
              "batch": lambda: batched_shallow_water(steps=args.steps),
              "advection_1d": advection_1d_full}[args.only]
        print(json.dumps(fn()), flush=True)
        return
    if args.wide:
        print(json.dumps(wide_order(args.wide, args.wide_version)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: finess_amd runs on one GPU only; --gpus %d is not supported "
              "(the multi-GPU slab decomposition is a design, not code)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from finess_amd.indicator import Indicator, train_indicator

    ROUTE_INTERVAL = args.route_interval
    mx = args.mx
    cells = mx * mx
    if args.indicator:
        ind = Indicator.load(args.indicator)
    else:
        ind = train_indicator(mx=mx, epochs=args.train_epochs, stride=args.train_stride,
                              seeds=tuple(range(3, 3 + args.train_seeds)))
    if args.save_indicator:
        ind.save(args.save_indicator)

    # throughput over K steps from the mid-run state
    el_p, rej_p, launch_p = timed_steps("plain", mx, args.steps, args.warmup)
    el_l, rej_l, launch_l = timed_steps("learned", mx, args.steps, args.warmup, ind)
    v_plain = cells * args.steps / el_p / 1e6
    v_learned = cells * args.steps / el_l / 1e6

    accuracy = None
    if not args.no_accuracy:
        rho_ref, el_r, n_r, _ = full_run("plain", 2 * mx)
        rho_ref = 0.25 * (rho_ref[0::2, 0::2] + rho_ref[1::2, 0::2] + rho_ref[0::2, 1::2] + rho_ref[1::2, 1::2])
        rho_p, el_fp, n_p, _ = full_run("plain", mx)
        rho_l, el_fl, n_l, route = full_run("learned", mx, ind)
        l1_p = float(np.mean(np.abs(rho_p - rho_ref)))
        l1_l = float(np.mean(np.abs(rho_l - rho_ref)))
        accuracy = {
            "reference": "plain FP64 at %dx%d to t=%.2f, 2x2-averaged" % (2 * mx, 2 * mx, T_FINAL),
            "l1_density_plain_fp64": l1_p,
            "l1_density_learned_fp32": l1_l,
            "l1_ratio": round(l1_l / l1_p, 5),
            "within_5_percent": bool(l1_l <= 1.05 * l1_p),
            "l1_learned_vs_plain": float(np.mean(np.abs(rho_l - rho_p))),
            "troubled_fraction": round(route[0] / max(route[1], 1), 4),
            "full_run_mcups": {"plain_fp64": round(cells * n_p / el_fp / 1e6, 1),
                               "learned_fp32": round(cells * n_l / el_fl / 1e6, 1),
                               "steps": [n_p, n_l]},
        }

    bytes_per_launch = cells * BYTES_PER_CELL_UPDATE / STAGES_PER_STEP
    achieved = bytes_per_launch / launch_l / 1e9
    out = {
        "metric": "Mcell-updates/s, 2D Euler WENO5: learned+FP32 vs plain FP64 at equal L1 error",
        "value": round(v_learned, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el_l / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "2D Euler four-quadrant Riemann, %dx%d, char-WENO5-JS + LLF, SSP-RK3, CFL 0.4, "
                        "transmissive, t_final %.1f; learned MLP routing + FP32 stencil vs plain FP64"
                        % (mx, mx, T_FINAL),
            "mx": mx, "my": mx, "meqn": 5, "timed_from_t": T_TIMED, "indicator_every_steps": ROUTE_INTERVAL,
            "steps_rejected": {"plain_fp64": rej_p, "learned_fp32": rej_l},
        },
        "plain_fp64": {"value": round(v_plain, 3), "ms_per_step": round(el_p / args.steps * 1e3, 4),
                       "avg_stage_launch_us": round(launch_p * 1e6, 2),
                       "hbm_frac": round(bytes_per_launch / launch_p / 1e9 / HBM_PEAK_GBS, 4),
                       "traffic": TRAFFIC_PLAIN_1024 if mx == 1024 else None,
                       # the stage is bound by VALU issue, FP64 arithmetic above all
                       "valu": None if mx != 1024 else {
                           "issue_frac": round(VALU_PLAIN_1024 * 4 / SIMDS / CLOCK_HZ / launch_p, 4),
                           "f64_issue_frac": round(VALU_F64_PLAIN_1024 * 4 / SIMDS / CLOCK_HZ / launch_p, 4),
                           "instructions_per_launch": VALU_PLAIN_1024,
                           "f64_instructions_per_launch": VALU_F64_PLAIN_1024}},
        "speedup_vs_plain_fp64": round(v_learned / v_plain, 4),
        "indicator": ind.info,
        "accuracy": accuracy,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": TRAFFIC_ROUTED_1024 if mx == 1024 else None,
            "kernel": "learned+FP32 stage: k_route_full + k_stage2d_lookup, with k_classify in front at the "
                      "first stage of every 4th step; mean over the 12 stages of a 4-step interval",
            "avg_launch_us": round(launch_l * 1e6, 2),
            "bytes_per_launch": bytes_per_launch,
        },
    }
    if not args.no_batch:
        out["batched_shallow_water"] = batched_shallow_water()
    if not args.no_extra:
        w = at_2048(ind, accuracy=not args.no_accuracy)
        out["plain_fp64_2048"] = dict(w["plain_fp64"], workload=w["workload"] + ", plain FP64")
        out["learned_fp32_2048"] = dict(w["learned_fp32"], speedup_vs_plain_fp64=w["speedup_vs_plain_fp64"],
                                        accuracy=w.get("accuracy"), workload=w["workload"] + ", learned+FP32")
        out["advection_1d_200"] = advection_1d_full()
        out["kelvin_helmholtz_lxw_2048"] = kelvin_helmholtz_lxw(t_final=0.0 if args.no_accuracy else 1.0)
        out["shock_bubble_4096x1024"] = shock_bubble_learned(ind, t_final=0.0 if args.no_accuracy else 0.2)
    if not args.no_cpu:
        # the host's physical cores that this process may use
        h = host_cpu()
        threads = min(h["physical_cores"] or h["usable_cpus"], h["usable_cpus"])
        out["cpu_baseline"] = cpu_baseline(mx, args.cpu_steps, threads, others=not args.no_extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
