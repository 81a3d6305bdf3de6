"""Launch-shape autotuner for the particle kernels.

For one case and precision it times a few whole time steps for each
workgroup size on a scratch context and returns the fastest as the ``tune``
dict that ``SphContext`` / ``sph_params.wg_size`` accept.  (Staging the
neighbour cells' particles in LDS was measured against plain L2-cached
gathers during development and lost at every block size and tile on MI355X,
so the only launch knob left is the workgroup size.)
"""
from __future__ import annotations

import time


def candidates(tb, precision):
    return [dict(wg_size=w) for w in (64, 128, 192, 256)]


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
