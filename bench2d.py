"""The 2D dam break on the device: 3.22 x 1.8 m tank, 1.0 x 0.5 m water column,
dx = 0.01 m (5,000 particles), SA-SPH walls, 1,000 leapfrog steps (the case
the reference runs on its CPU path in its -DSPACE_2D build).

Runs the device in FP64 and FP32-mixed (one replayed graph per step), checks
both against the FP64 oracle after the same steps, and times the reference's
own 2D build (flang OpenMP, through the dump harness) on the same case on the
host cores.  One JSON line on stdout.

    python bench2d.py [--steps 1000]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from sphera_amd import case2d  # noqa: E402
from sphera_amd import runtime as rt  # noqa: E402

FIELDS = ("coord", "vel", "dens", "pres")


def _device(tb, precision, steps):
    d = rt.SphContext(tb, precision=precision)
    try:
        d.upload()
        d.preloop()
        d.step_graph(1)  # capture outside the timed region
        d.sync()
        p0 = d.pair_counter()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            d.step_graph(1)
        d.sync()
        t1 = time.perf_counter()
        pairs = d.pair_counter() - p0
        return d.state(), (t1 - t0) / (steps - 1), pairs / (t1 - t0)
    finally:
        d.close()


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--ref-steps", type=int, default=200,
                    help="steps of the reference's timed sample (difference of two runs)")
    a = ap.parse_args(argv)
    from oracle.pyoracle import Oracle
    case = case2d.DamBreak2D()
    tb = case2d.build(case)
    n = len(tb.particles["coord"])
    print(f"[bench2d] n={n} steps={a.steps}", file=sys.stderr, flush=True)
    o = Oracle(tb)
    o.preloop()
    t0 = time.perf_counter()
    o.step(a.steps)
    t_or = time.perf_counter() - t0
    ref = o.state()
    out = dict(workload=f"2D dam break 3.22 x 1.8 m tank, 1.0 x 0.5 m column, dx 0.01 m, {n} particles",
               steps=a.steps, oracle_ms_per_step=t_or / a.steps * 1e3)
    for name, prec in (("fp64", rt.FP64), ("fp32_mixed", rt.FP32_MIXED)):
        st, s_per_step, pps = _device(tb, prec, a.steps)
        err = {f: _rel(st[f], ref[f]) for f in FIELDS}
        out[name] = dict(ms_per_step=s_per_step * 1e3, pair_interactions_per_s=pps,
                         rel_err_vs_oracle=err, time_s=st["time"], oracle_time_s=ref["time"])
        print(f"[bench2d] {name}: {s_per_step * 1e3:.3f} ms/step, max rel err {max(err.values()):.2e}",
              file=sys.stderr, flush=True)
    # the reference itself on the host cores: two runs, their difference
    from oracle.decks import HARNESS2D, run_harness
    if HARNESS2D.exists():
        threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
        times = []
        for k in (1, 1 + a.ref_steps):
            with tempfile.TemporaryDirectory(prefix="sph_ref2d_") as wd:
                t0 = time.perf_counter()
                run_harness(case, k, every=k, lists=False, threads=threads, workdir=wd)
                times.append(time.perf_counter() - t0)
        out["reference_cpu"] = dict(ms_per_step=(times[1] - times[0]) / a.ref_steps * 1e3, cores=threads,
                                    sample=f"{a.ref_steps} steps: runs of 1 and {1 + a.ref_steps} steps took "
                                           f"{times[0]:.2f} s and {times[1]:.2f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
