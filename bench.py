#!/usr/bin/env python3
"""Benchmark of the FINESS WENO path on one MI355X.

Workload (one GPU): 2D Euler four-quadrant Riemann problem, 1024 x 1024 cells,
gamma 1.4, characteristic-wise WENO5-JS (eps 1e-6, p 2) with local
Lax-Friedrichs splitting, SSP-RK3, a-posteriori CFL controller at CFL 0.4,
transmissive boundaries, plain FP64.  A "step" is one accepted time step
(three fused stage launches plus the controller's read-back of the wave-speed
maxima).  Inputs are resident in HBM before the timed region starts.

Prints one JSON line.  `value` is Mcell-updates/s = cells * steps / seconds.
`roofline` refers to the dominant kernel (the fused stage kernel): achieved =
algorithmic HBM bytes per launch / average launch duration, the duration
measured with HIP events on the stream the kernel runs on.  `cpu_baseline` is
the CPU restatement of the reference (oracle/) timed on this host's cores on a
bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 320.0   # SSP-RK3: 8 state passes x 5 fields x 8 B
STAGES_PER_STEP = 3


def cpu_baseline(mx, steps, threads):
    """Times the oracle's fixed-step loop (no controller) on `steps` steps of
    the same workload at full size."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc  # the oracle loader lives with the tests
    from finess_amd import problems
    lib = orc.oracle()
    lib.orc_set_num_threads(threads)
    p, q, aux = problems.riemann4(mx=mx)
    dt = 0.4 / mx / 2.1
    t0 = time.perf_counter()
    orc.fixed_steps(p, q, dt, steps, aux=np.zeros(p.shape_aux()))
    el = time.perf_counter() - t0
    return {
        "value": round(mx * mx * steps / el / 1e6, 3),
        "unit": "Mcell-updates/s",
        "cores": int(lib.orc_num_threads()),
        "kind": "port",
        "sample": "%d SSP-RK3 step(s) of the %dx%d four-quadrant problem, FP64, fixed dt, %.1f s" % (steps, mx, mx, el),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mx", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the multi-GPU slab decomposition is a design, not code)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from finess_amd import Solver, problems
    p, q, _ = problems.riemann4(mx=args.mx)
    cells = p.mx * p.my
    s = Solver(p)
    s.set_state(q)
    dtv = np.array([0.0, p.initial_dt, p.max_dt])
    s.advance_steps(args.warmup, dtv)
    s.sync()
    t0 = time.perf_counter()
    st = s.advance_steps(args.steps, dtv)
    s.sync()
    el = time.perf_counter() - t0
    value = cells * args.steps / el / 1e6

    # dominant kernel: average duration of the fused stage launches, timed
    # with HIP events on the context's stream over back-to-back steps
    dt_fixed = dtv[1]
    s.run_fixed_steps(dt_fixed, 50)
    ms, launches = s.last_kernel_time()
    avg_launch_s = ms / 1e3 / launches
    bytes_per_launch = cells * BYTES_PER_CELL_UPDATE / STAGES_PER_STEP
    achieved = bytes_per_launch / avg_launch_s / 1e9
    s.close()

    out = {
        "metric": "Mcell-updates/s, 2D Euler WENO5: learned+FP32 vs plain FP64 at equal L1 error",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "2D Euler four-quadrant Riemann, %dx%d, char-WENO5-JS + LLF, SSP-RK3, CFL 0.4, "
                        "transmissive, plain FP64" % (p.mx, p.my),
            "mx": p.mx, "my": p.my, "meqn": 5, "form": "plain-fp64",
            "steps_rejected": int(st.steps_rejected), "dt_last": float(dtv[1]),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_stage2d<Euler2D,double,JS5>",
            "avg_launch_us": round(avg_launch_s * 1e6, 2),
            "bytes_per_launch": bytes_per_launch,
        },
    }
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(args.mx, args.cpu_steps, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
