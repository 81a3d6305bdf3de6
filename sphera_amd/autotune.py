"""Launch-shape autotuner for the particle kernels.

For one case and precision it picks the workgroup size of the particle
kernels and, for FP32-mixed, the list-walk chunk of each pair pass (4 or 8
entries gathered per step: fewer registers, more waves in flight, or more
loads in flight per lane).  For each workgroup size two scratch contexts run
a few steps with every pass at chunk 4 and at chunk 8, with per-kernel HIP
event timing; each pass takes its faster chunk, and the workgroup size with
the smallest resulting step wins.  The neighbour search's launch shape
(workgroup size x LDS candidate tile, ``nb_shape``) is then chosen the same
way, from its kernel time in each shape.  The result is the ``tune`` dict that
``SphContext`` / ``sph_params`` accept.

The grid's cell size is not a tunable: it is the reference's (CreaGrid), and
it fixes the neighbour-list order the device reproduces.
"""
from __future__ import annotations

import time

# neighbour-search launch shapes (workgroup size x LDS tile): 256 x 4096 FP16,
# 256 x 3072 FP32, 128 x 1536 FP32, 256 x 2816 FP16 held to 64 registers,
# 256 x 4096 in 16-bit fixed point over the cell
NB_SHAPES = (0, 1, 2, 3, 4)
PASSES = {"momentum": "chunk", "velocity_smoothing": "chunk_v", "continuity": "chunk_c",
          "pressure_smoothing": "chunk_p"}


def _run(tb, coord, precision, tune, steps, warmup):
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
        d.set_timing(True)
        d.step(steps)
        ms = {k: v[0] / v[1] for k, v in d.timing().items()}
        return (t1 - t0) / steps, ms
    finally:
        d.close()


def autotune(tb, coord, precision, steps=4, warmup=2, log=None):
    """Returns (best tune dict, list of (tune, seconds per step))."""
    results = []
    best, best_s = None, float("inf")
    for wg in (64, 128, 256):
        if precision == 1:  # FP32-mixed: per-pass chunk from the kernel times
            s8, ms8 = _run(tb, coord, precision, dict(wg_size=wg, chunk=8), steps, warmup)
            s4, ms4 = _run(tb, coord, precision, dict(wg_size=wg, chunk=4), steps, warmup)
            tune = dict(wg_size=wg)
            est = s8
            for k, field in PASSES.items():
                c = 4 if ms4.get(k, 1e9) < ms8.get(k, 1e9) else 8
                tune[field] = c
                if c == 4:
                    est -= (ms8[k] - ms4[k]) / 1e3
            results += [(dict(wg_size=wg, chunk=8), s8), (dict(wg_size=wg, chunk=4), s4)]
        else:
            tune = dict(wg_size=wg, chunk=8)
            est, _ = _run(tb, coord, precision, tune, steps, warmup)
            results.append((tune, est))
        if log:
            log(f"[autotune] precision={precision} {tune} -> {est * 1e3:.3f} ms/step")
        if est < best_s:
            best, best_s = tune, est
    nb_ms = {}
    for shape in NB_SHAPES:
        t = dict(best, nb_shape=shape)
        _, ms = _run(tb, coord, precision, t, steps, warmup)
        nb_ms[shape] = ms.get("neighbours", float("inf"))
        if log:
            log(f"[autotune] nb_shape={shape}: neighbours {nb_ms[shape]:.4f} ms")
    best = dict(best, nb_shape=min(nb_ms, key=nb_ms.get))
    return best, results
