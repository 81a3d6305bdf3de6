#!/usr/bin/env python3
"""Headline benchmark: pair-interactions/s of the 3D dam break on one MI355X.

Workload: tank 1.6 x 0.6 x 0.6 m, water column 0.4 x 0.6 x 0.3 m, dx = 0.005 m
(576,000 particles), h = 1.3 dx, linear EoS with c0 = 10 sqrt(g H), SA-SPH
walls, lattice jitter U(+-0.01 dx) from seed 20260101, leapfrog.  A step is one
whole time step on the device (dt, momentum, velocity smoothing, trajectories,
cell sort, neighbour lists, SA-SPH boundary table, continuity + EoS, pressure
smoothing), inputs resident in HBM.  value = pair interactions / s of the
FP32-mixed path, sum_steps sum_i (nPartIntorno(i)-1) / t; the plain FP64 path
is timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle of one pair pass: each cell stages its 27
# neighbour cells once (record size S per particle) plus the per-particle output
REC_BYTES = {  # pass: (S fp64, out fp64, S fp32, out fp32)
    "continuity": (68, 8, 36, 4),
    "pressure_smoothing": (52, 8, 28, 4),
    "momentum": (84, 24, 44, 12),
    "velocity_smoothing": (68, 24, 36, 12),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(args):
    from sphera_amd.case import DamBreak3D, build
    case = DamBreak3D(tank=(1.6, 0.6, 0.6), column=(0.4, 0.6, 0.3), dx=args.dx, name="dam_break_3D")
    tb = build(case)
    rng = np.random.default_rng(20260101)
    c = tb.particles["coord"]
    coord = c + rng.uniform(-0.01, 0.01, c.shape) * case.dx
    return case, tb, coord


def run_device(tb, coord, precision, steps, warmup, tune=None):
    from sphera_amd import runtime as rt
    d = rt.SphContext(tb, precision=precision, tune=tune)
    d.upload(coord=coord)
    d.preloop()
    if warmup:
        d.step(warmup)
    d.set_timing(True)
    p0 = d.pair_counter()  # synchronises the context stream
    t0 = time.perf_counter()
    d.step(steps)          # enqueues every step, then synchronises
    t1 = time.perf_counter()
    pairs = d.pair_counter() - p0
    timing = d.timing()
    n = d.n
    sc = d.scalars()
    launch = d.launch()
    d.close()
    return dict(seconds=t1 - t0, pairs=pairs, timing=timing, n=n, it=sc["it"], dt=sc["dt"],
                launch=launch)


def cpu_baseline(tb, coord, steps):
    from oracle.pyoracle import Oracle
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    o = Oracle(tb, coord=coord)
    o.preloop()
    o.step(1)
    # Σ over steps of the list pairs built in each step
    pairs = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(1)
        pairs += o.pair_count()
    t1 = time.perf_counter()
    return dict(value=pairs / (t1 - t0), unit="pair-interactions/s", cores=threads,
                kind="port",
                sample=f"oracle FP64 OpenMP restatement, full {len(coord)}-particle case, "
                       f"{steps} whole time steps after 1 untimed ({(t1 - t0):.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.005)
    ap.add_argument("--cpu-steps", type=int, default=int(os.environ.get("SPH_CPU_STEPS", "5")))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU decomposition "
              "(every run uses one MI355X); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from sphera_amd import runtime as rt
    case, tb, coord = make_case(args)
    log(f"[bench] case n={len(coord)} grid={tuple(tb.grid.ncd)} steps={args.steps} warmup={args.warmup}")
    res = {}
    tunes = {}
    for name, prec in (("fp32_mixed", rt.FP32_MIXED), ("fp64", rt.FP64)):
        tune = None
        if not args.no_autotune:
            from sphera_amd.autotune import autotune
            tune, _ = autotune(tb, coord, prec, log=log)
            log(f"[bench] {name}: autotuned launch {tune}")
        tunes[name] = tune
        log(f"[bench] running {name}")
        res[name] = run_device(tb, coord, prec, args.steps, args.warmup, tune)
        r = res[name]
        log(f"[bench] {name}: {r['seconds'] / args.steps * 1e3:.3f} ms/step, "
            f"{r['pairs'] / r['seconds'] / 1e9:.3f} G pairs/s")
    main_r = res["fp32_mixed"]
    value = main_r["pairs"] / main_r["seconds"]
    n = main_r["n"]
    # dominant kernel of the FP32-mixed step and its roofline
    pair_k = {k: v for k, v in main_r["timing"].items() if k in REC_BYTES}
    dom = max(pair_k, key=lambda k: pair_k[k][0])
    ms_tot, calls = pair_k[dom]
    avg_s = ms_tot / calls / 1e3
    S, out = REC_BYTES[dom][2], REC_BYTES[dom][3]
    bytes_alg = n * (27 * S + out)
    achieved = bytes_alg / avg_s / 1e9
    kern = {k: round(v[0] / v[1], 4) for k, v in main_r["timing"].items()}
    kern64 = {k: round(v[0] / v[1], 4) for k, v in res["fp64"]["timing"].items()}
    pair_time = sum(v[0] for k, v in main_r["timing"].items() if k in REC_BYTES) / 1e3
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        log(f"[bench] cpu baseline ({args.cpu_steps} steps)")
        cpu = cpu_baseline(tb, coord, args.cpu_steps)
    out_line = {
        "metric": "pair-interactions/s mixed vs FP64 (3D dam break); p50/p99 ensemble latency",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp64 positions/integration)",
        "data": "synthetic: jittered lattice dam break, seed 20260101",
        "config": {"workload": "3D dam break, tank 1.6x0.6x0.6 m, column 0.4x0.6x0.3 m, "
                               f"dx={case.dx} m, h/dx=1.3, SA-SPH walls", "particles": n,
                   "grid_cells": int(tb.grid.nmax)},
        "fp64": {"value": res["fp64"]["pairs"] / res["fp64"]["seconds"],
                 "ms_per_step": res["fp64"]["seconds"] / args.steps * 1e3,
                 "kernel_ms": kern64, "launch": res["fp64"]["launch"]},
        "launch": main_r["launch"],
        "mixed_over_fp64": (value / (res["fp64"]["pairs"] / res["fp64"]["seconds"])),
        "pair_kernels_only_value": main_r["pairs"] / pair_time if pair_time > 0 else None,
        "kernel_ms": kern,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_alg_per_launch": bytes_alg,
                     "avg_launch_ms": avg_s * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out_line), flush=True)


if __name__ == "__main__":
    main()
