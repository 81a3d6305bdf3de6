#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ by running the flang build of the
SPHERA v10 reference (oracle/_ref/harness3d.x, built by oracle/Makefile.ref
from /root/reference/src) on project-authored decks.

    make -f oracle/Makefile.ref -j8 && python tests/golden/make_golden.py

Fixtures (numpy .npz, inputs and reference outputs only):
  ref_tiny.npz          800 particles, lattice IC; state + neighbour lists +
                        SA-SPH boundary table after the pre-loop and after 5
                        steps; set-up tables of the header
  ref_small_jitter.npz  1440 particles, lattice + U(+-0.01 dx) jitter from
                        seed 20260101; state after 0 and 10 steps
  ref_removal.npz       800 particles of which 6 start outside the grid (three
                        holes, three at the tail): OrdGrid1's swap-with-last
                        order; state after the pre-loop and after 6 steps
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle.decks import read_dump, run_harness  # noqa: E402
from sphera_amd.case import DamBreak3D, build  # noqa: E402

TINY = dict(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny")
SMALL = dict(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01, name="small")
FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")


def removal_slots(n):
    """Particles moved outside the grid in the removal fixture."""
    return (3, 17, 200, n - 4, n - 2, n - 1)


def jitter(coord, dx, seed=20260101, amp=0.01):
    rng = np.random.default_rng(seed)
    return coord + rng.uniform(-amp, amp, coord.shape) * dx


def pack_step(prefix, st, out, lists):
    out[prefix + "it"] = st["it"]
    out[prefix + "time"] = st["time"]
    out[prefix + "dt"] = st["dt"]
    out[prefix + "dt_average"] = st["dt_average"]
    for f in FIELDS + ("laminar_flag",):
        out[prefix + f] = st[f]
    if lists:
        out[prefix + "nbr_count"] = st["nbr_count"]
        out[prefix + "nbr"] = st["nbr"]
        out[prefix + "ncbf"] = st["ncbf"]
        bt = st["btab"]
        out[prefix + "bt_face"] = bt["face"]
        out[prefix + "bt_loc"] = bt["loc"]
        out[prefix + "bt_bi"] = bt["bi"]
        out[prefix + "bt_ig"] = bt["ig"]


def main():
    # tiny: lists and boundary table
    case = DamBreak3D(**TINY)
    hdr, steps = read_dump(run_harness(case, nsteps=5, every=5))
    out = {}
    for k in ("nag", "nmaxpartj", "ncd", "extr_lo", "dcd", "h", "coefke", "coefkacl", "eta",
              "eta2", "celerita", "kin_visc", "kerneltab", "ktdelta", "bitab", "gcbf_count",
              "gcbf_start", "gcbf_vector"):
        out["hdr_" + k] = np.asarray(hdr[k])
    fr = [f for f in hdr["faces"] if f["tipo"] != "peri"]
    out["hdr_face_id"] = np.array([f["nf"] for f in fr])
    out["hdr_face_T"] = np.array([f["T"] for f in fr])
    out["hdr_face_area"] = np.array([f["area"] for f in fr])
    out["hdr_face_LX"] = np.array([[nd[2][:2] for nd in f["node"][:4]] for f in fr])
    for st in steps:
        pack_step(f"s{st['it']}_", st, out, lists=True)
    np.savez_compressed(HERE / "ref_tiny.npz", **out)
    # small, jittered: fields over 10 steps
    case = DamBreak3D(**SMALL)
    tb = build(case)
    ic = jitter(tb.particles["coord"], case.dx)
    hdr, steps = read_dump(run_harness(case, nsteps=10, every=10, lists=False, ic_coord=ic))
    out = {"ic_coord": ic}
    for st in steps:
        pack_step(f"s{st['it']}_", st, out, lists=False)
    np.savez_compressed(HERE / "ref_small_jitter.npz", **out)
    # removal order: particles placed outside the grid
    case = DamBreak3D(**TINY)
    tb = build(case)
    ic = tb.particles["coord"].copy()
    n = len(ic)
    for k in removal_slots(n):
        ic[k] = (50.0, 50.0, 50.0)
    hdr, steps = read_dump(run_harness(case, nsteps=6, every=6, lists=False, ic_coord=ic))
    out = {"ic_coord": ic}
    for st in steps:
        out[f"s{st['it']}_nag"] = st["nag"]
        pack_step(f"s{st['it']}_", st, out, lists=False)
    np.savez_compressed(HERE / "ref_removal.npz", **out)
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
