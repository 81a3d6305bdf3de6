"""Launch-shape autotuner for the particle kernels.

For one case and precision it times a few whole time steps for each
workgroup size (and, for FP32-mixed, each list-walk chunk) on a scratch
context and returns the fastest as the ``tune`` dict that ``SphContext`` /
``sph_params`` accept.  The chunk trades registers (occupancy) against loads
in flight per lane.
"""
from __future__ import annotations

import time


def candidates(tb, precision):
    chunks = (4, 8) if precision == 1 else (8,)
    return [dict(wg_size=w, chunk=c) for w in (64, 128, 256) for c in chunks]


def time_config(tb, coord, precision, tune, steps=4, warmup=2):
    from . import runtime as rt
    d = rt.SphContext(tb, precision=precision, tune=tune)
    try:
        d.upload(coord=coord)
        d.preloop()
        d.step(warmup)
        d.sync()
        t0 = time.perf_counter()
        d.step(steps)
        t1 = time.perf_counter()
        return (t1 - t0) / steps
    finally:
        d.close()


def autotune(tb, coord, precision, steps=4, log=None):
    """Returns (best tune dict, list of (tune, seconds per step))."""
    results = []
    for cfg in candidates(tb, precision):
        try:
            sec = time_config(tb, coord, precision, cfg, steps=steps)
        except Exception as e:  # a shape the device rejects is skipped
            if log:
                log(f"[autotune] {cfg}: rejected ({e})")
            continue
        results.append((cfg, sec))
        if log:
            log(f"[autotune] precision={precision} {cfg} -> {sec * 1e3:.3f} ms/step")
    best = min(results, key=lambda r: r[1])[0]
    return best, results
