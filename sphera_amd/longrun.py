"""Long-horizon comparison of the FP32-mixed path with the FP64 path.

Both precisions run the same case from the same state side by side, on two
contexts of one device.  Every `every` steps the two states are downloaded
and compared in pg order:

* per field (coord, vel, dens, pres): max |FP32 - FP64| relative to the
  field's max |FP64|, and the RMS difference relative to the RMS of FP64;
* the dam-break front x_front = max x over the particles, and the water depth
  at gauges (highest particle within dx of the gauge's x, as the monitoring
  lines of calc_pelo.f90 read it), in both precisions;
* simulated time (dt comes from each run's own velocities) and global mass.

    python -m sphera_amd.longrun [--steps 2000] [--every 100] [--dx 0.005]
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

GAUGES = (0.1, 0.4, 0.8, 1.2)


def _metrics(a, b):
    out = {}
    for f in ("coord", "vel", "dens", "pres"):
        d = np.abs(a[f] - b[f])
        out[f + "_max_rel"] = float(d.max() / max(np.abs(b[f]).max(), 1e-300))
        out[f + "_rms_rel"] = float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(b[f] * b[f])), 1e-300))
    return out


def _surface(st, dx):
    x, z = st["coord"][:, 0], st["coord"][:, 2]
    g = []
    for xg in GAUGES:
        m = np.abs(x - xg) <= dx
        g.append(float(z[m].max()) if m.any() else 0.0)
    return float(x.max()), g


def run(steps=2000, every=100, dx=0.005, log=print):
    from . import runtime as rt
    from .case import DamBreak3D, build
    case = DamBreak3D(tank=(1.6, 0.6, 0.6), column=(0.4, 0.6, 0.3), dx=dx, name="dam_break_3D")
    tb = build(case)
    rng = np.random.default_rng(20260101)
    coord = tb.particles["coord"] + rng.uniform(-0.01, 0.01, tb.particles["coord"].shape) * dx
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
            r = dict(it=it, time_fp64=b["time"], time_fp32=a["time"], n=len(a["coord"]),
                     front_fp64=fb, front_fp32=fa, gauges_fp64=gb, gauges_fp32=ga,
                     mass_fp64=mb["mass"], mass_fp32=ma["mass"], **_metrics(a, b))
            rows.append(r)
            log(json.dumps(r))
    finally:
        for d in ctx.values():
            d.close()
    return rows


def markdown(rows, dx):
    head = ("| step | t FP64 (s) | t FP32 (s) | front FP64 (m) | front FP32 (m) | coord max rel | "
            "vel rms rel | dens max rel | pres rms rel | gauges FP64 (m) | gauges FP32 (m) |")
    out = [head, "|" + "---|" * 11]
    for r in rows:
        out.append(f"| {r['it']} | {r['time_fp64']:.5f} | {r['time_fp32']:.5f} | {r['front_fp64']:.4f} | "
                   f"{r['front_fp32']:.4f} | {r['coord_max_rel']:.2e} | {r['vel_rms_rel']:.2e} | "
                   f"{r['dens_max_rel']:.2e} | {r['pres_rms_rel']:.2e} | "
                   f"{', '.join(f'{g:.4f}' for g in r['gauges_fp64'])} | "
                   f"{', '.join(f'{g:.4f}' for g in r['gauges_fp32'])} |")
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
