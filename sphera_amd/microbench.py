"""Neighbour-search and pair-pass microbenchmark on a uniform random cloud.

N particles uniform in [0,1)^3 (numpy PCG64, seed 12345), the smoothing length
chosen for ~60 neighbours in the 2h sphere, 2h = (3*60/(4*pi*N))^(1/3)
(h = 1.3 dx), velocities N(0, 0.1) per component, rho = 1000 (1 + U(-0.01,
0.01)), p from the linear EoS, m = rho0/N.  No SA-SPH frontiers.  One context
runs the cell sort, the neighbour lists, and whole steps; the per-kernel
times of the sort, the lists, one continuity and one momentum pass are
reported (HIP events on the context stream).

    python -m sphera_amd.microbench [--n 16777216] [--precision fp32_mixed|fp64]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np


def cloud(n, seed=12345):
    from .case import DamBreak3D, build
    two_h = (3.0 * 60.0 / (4.0 * math.pi * n)) ** (1.0 / 3.0)
    dx = two_h / 2.0 / 1.3
    case = DamBreak3D(tank=(1.0, 1.0, 1.0), column=(1.0, 1.0, 1.0), dx=dx, name="cloud")
    tb = build(case, particles=False, incidence=False)
    rng = np.random.default_rng(seed)
    coord = rng.random((n, 3))
    vel = rng.normal(0.0, 0.1, (n, 3))
    dens = case.rho0 * (1.0 + rng.uniform(-0.01, 0.01, n))
    pres = case.bulk / case.rho0 * (dens - case.rho0) + case.p_ref
    mass = np.full(n, case.rho0 / n)
    return tb, dict(coord=coord, vel=vel, dens=dens, pres=pres, mass=mass)


def run(n, precision="fp32_mixed", steps=3, log=print):
    from . import runtime as rt
    prec = rt.FP32_MIXED if precision == "fp32_mixed" else rt.FP64
    t = time.perf_counter()
    tb, st = cloud(n)
    log(f"[micro] n={n} grid={tuple(int(v) for v in tb.grid.ncd)} set-up {time.perf_counter() - t:.1f} s")
    d = rt.SphContext(tb, precision=prec, capacity=n, semi=False)
    try:
        d.upload(**st)
        d.preloop()
        d.step(1)
        d.sync()
        d.set_timing(True)
        p0 = d.pair_counter()
        d.step(steps)
        pairs = (d.pair_counter() - p0) / steps
        tm = d.timing()
    finally:
        d.close()
    ms = {k: v[0] / v[1] for k, v in tm.items()}
    sort = sum(ms.get(k, 0.0) for k in ("sort_count", "sort_scan", "sort_scatter", "sort_gather"))
    return {"n": n, "precision": precision, "grid": [int(v) for v in tb.grid.ncd],
            "pairs_per_particle": pairs / n, "sort_ms": sort, "neighbours_ms": ms.get("neighbours"),
            "continuity_ms": ms.get("continuity"), "momentum_ms": ms.get("momentum"),
            "pairs_per_s_continuity": pairs / (ms["continuity"] / 1e3),
            "pairs_per_s_momentum": pairs / (ms["momentum"] / 1e3),
            "kernel_ms": {k: round(v, 4) for k, v in ms.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=16777216)
    ap.add_argument("--precision", choices=("fp32_mixed", "fp64"), default="fp32_mixed")
    ap.add_argument("--steps", type=int, default=3)
    a = ap.parse_args()
    out = run(a.n, a.precision, a.steps, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
