"""Long-horizon comparison of the FP32-mixed path with the FP64 path.

Both precisions run the same case from the same state side by side, on two
contexts of one device.  Every `every` steps the two states are downloaded
and compared in pg order:

* per field (coord, vel, acc, dens, pres, dden, vpres): the largest
  |FP32 - FP64| over the particles relative to the field's largest |FP64|
  ("max"), and the RMS difference relative to the RMS of FP64 ("rms"); for
  the field with the largest "max", the particle where it occurs (position,
  number of particles within 2h, its |FP64 value| relative to the field's
  largest);
* the dam-break front x_front = max x over the particles, and the water depth
  at gauges (highest particle within dx of the gauge's x, as the monitoring
  lines of calc_pelo.f90 read it), in both precisions;
* simulated time (dt comes from each run's own velocities) and global mass.

`run(..., extra=f)` calls f(step, fp32_state, fp64_state) at every
comparison and adds the dict it returns to the row (a checker can compare
both states with a third run there).

    python -m sphera_amd.longrun [--steps 2000] [--every 100] [--dx 0.005]
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

GAUGES = (0.1, 0.4, 0.8, 1.2)
FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")


def field_errors(a, b, fields=FIELDS):
    """max and RMS of |a - b| per field, relative to b's largest |value| and RMS;
    also the index of the particle with the largest difference."""
    out = {}
    for f in fields:
        d = np.abs(np.asarray(a[f], np.float64) - b[f])
        d1 = d.reshape(len(d), -1).max(1) if d.ndim > 1 else d
        out[f + "_max_rel"] = float(d.max() / max(np.abs(b[f]).max(), 1e-300))
        out[f + "_rms_rel"] = float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(b[f] * b[f])), 1e-300))
        out[f + "_argmax"] = int(np.argmax(d1)) if len(d1) else -1
    return out


def _worst(errs, b, h):
    """The particle of the field with the largest relative max error."""
    f = max(FIELDS, key=lambda k: errs[k + "_max_rel"])
    i = errs[f + "_argmax"]
    x = b["coord"]
    r2 = np.sum((x - x[i]) ** 2, axis=1)
    v = np.abs(b[f][i]).max() / max(np.abs(b[f]).max(), 1e-300)
    return dict(worst_field=f, worst_index=i, worst_coord=[float(c) for c in x[i]],
                worst_neighbours=int(np.count_nonzero(r2 <= (2.0 * h) ** 2)), worst_value_rel=float(v))


def _surface(st, dx):
    x, z = st["coord"][:, 0], st["coord"][:, 2]
    g = []
    for xg in GAUGES:
        m = np.abs(x - xg) <= dx
        g.append(float(z[m].max()) if m.any() else 0.0)
    return float(x.max()), g


def make_case(dx=0.005):
    """The 3D dam break (1.6 x 0.6 x 0.6 m tank, 0.4 x 0.6 x 0.3 m column) with
    the jittered lattice of seed 20260101; returns (case tables, positions)."""
    from .case import DamBreak3D, build
    case = DamBreak3D(tank=(1.6, 0.6, 0.6), column=(0.4, 0.6, 0.3), dx=dx, name="dam_break_3D")
    tb = build(case)
    rng = np.random.default_rng(20260101)
    coord = tb.particles["coord"] + rng.uniform(-0.01, 0.01, tb.particles["coord"].shape) * dx
    return tb, coord


def run(steps=2000, every=100, dx=0.005, log=print, extra=None):
    from . import runtime as rt
    tb, coord = make_case(dx)
    ctx = {}
    for name, prec in (("fp64", rt.FP64), ("fp32_mixed", rt.FP32_MIXED)):
        d = rt.SphContext(tb, precision=prec)
        d.upload(coord=coord)
        d.preloop()
        ctx[name] = d
    rows = []
    try:
        for it in range(every, steps + 1, every):
            for d in ctx.values():
                d.step(every)
            a, b = ctx["fp32_mixed"].state(), ctx["fp64"].state()
            fa, ga = _surface(a, dx)
            fb, gb = _surface(b, dx)
            ma, mb = ctx["fp32_mixed"].global_quantities(), ctx["fp64"].global_quantities()
            errs = field_errors(a, b)
            r = dict(it=it, time_fp64=b["time"], time_fp32=a["time"], n=len(a["coord"]),
                     front_fp64=fb, front_fp32=fa, gauges_fp64=gb, gauges_fp32=ga,
                     mass_fp64=mb["mass"], mass_fp32=ma["mass"], **errs, **_worst(errs, b, tb.case.h))
            if extra is not None:
                r.update(extra(it, a, b))
            rows.append(r)
            log(json.dumps(r))
    finally:
        for d in ctx.values():
            d.close()
    return rows


def markdown(rows, dx):
    head = ("| step | t (s) | front FP64 / FP32 (m) | " +
            " | ".join(f"{f} max / rms" for f in FIELDS) + " | worst particle |")
    out = [head, "|" + "---|" * (4 + len(FIELDS))]
    for r in rows:
        w = (f"{r['worst_field']} at ({', '.join(f'{c:.3f}' for c in r['worst_coord'])}), "
             f"{r['worst_neighbours']} within 2h, |value| {r['worst_value_rel']:.1e} of max")
        out.append(f"| {r['it']} | {r['time_fp64']:.5f} | {r['front_fp64']:.4f} / {r['front_fp32']:.4f} | " +
                   " | ".join(f"{r[f + '_max_rel']:.1e} / {r[f + '_rms_rel']:.1e}" for f in FIELDS) +
                   f" | {w} |")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--dx", type=float, default=0.005)
    a = ap.parse_args()
    rows = run(a.steps, a.every, a.dx, log=lambda s: print(s, file=sys.stderr, flush=True))
    print(markdown(rows, a.dx), flush=True)


if __name__ == "__main__":
    main()
