"""The 3D dam break with DB-SPH walls on the device: the 1.6 x 0.6 x 0.6 m tank
as a surface mesh (sphera_amd.dbsph.tank_mesh), the 0.4 x 0.6 x 0.3 m column,
dx = 0.01 m, explicit RK1 (Domain%tipo = "bsph", the reference's DB-SPH
branch), FP64.

Times the device's step (one replayed graph per step) and its kernels (HIP
events of the context's timing mode, a separate run), checks the state
against the FP64 oracle after the same steps, and times the reference's own
DB-SPH build (flang OpenMP, through the dump harness) on the same case on
the host cores.  One JSON line on stdout.

    python bench_dbsph.py [--steps 200] [--oracle-steps 20] [--ref-steps 10] [--moving]

--moving: the tank moves (kinematics records: a translation ramping up to
0.5 m/s along x and a rotation ramping up to 1 rad/s about the vertical
through the floor's centre).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from sphera_amd import runtime as rt  # noqa: E402
from sphera_amd.case import DamBreak3D, build  # noqa: E402
from sphera_amd.dbsph import DBSPHParams  # noqa: E402

FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")


MOVING = dict(kinematics=((0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0), (0.01, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0),
                          (100.0, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0)),
              rotation_centre=(0.8, 0.3, 0.0))


def _case(dx, moving=False):
    return DamBreak3D(tank=(1.6, 0.6, 0.6), column=(0.4, 0.6, 0.3), dx=dx, name="dam_break_dbsph",
                      time_split=0, dbsph=DBSPHParams(**(MOVING if moving else {})))


def _ctx(tb, coord):
    d = rt.SphContext(tb, precision=rt.FP64)
    d.upload(coord=coord)
    d.preloop()
    return d


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--dx", type=float, default=0.01)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--oracle-steps", type=int, default=20)
    ap.add_argument("--ref-steps", type=int, default=10,
                    help="steps of the reference's timed sample (difference of two runs)")
    ap.add_argument("--no-ref", action="store_true")
    ap.add_argument("--moving", action="store_true", help="moving tank (kinematics records)")
    a = ap.parse_args(argv)
    case = _case(a.dx, a.moving)
    t0 = time.perf_counter()
    tb = build(case)
    t_pre = time.perf_counter() - t0
    rng = np.random.default_rng(20261208)
    c = tb.particles["coord"]
    coord = c + rng.uniform(-0.01, 0.01, c.shape) * case.dx
    n, nw = len(coord), len(tb.walls)
    print(f"[bench_dbsph] n={n} walls={nw} (wall pre-processing {t_pre:.1f} s)", file=sys.stderr, flush=True)
    out = dict(workload=f"3D dam break with DB-SPH walls: 1.6 x 0.6 x 0.6 m tank mesh ({nw} wall elements"
                        f"{', moving' if a.moving else ''}), "
                        f"0.4 x 0.6 x 0.3 m column, dx {case.dx} m, {n} particles, explicit RK1",
               dtype="f64", n_particles=n, n_walls=nw, steps=a.steps)
    # parity: the device against the oracle after the same steps
    from oracle.pyoracle import Oracle
    o = Oracle(tb, coord=coord)
    o.preloop()
    t0 = time.perf_counter()
    o.step(a.oracle_steps)
    t_or = time.perf_counter() - t0
    ref, rdb = o.state(), o.dbsph_state()
    d = _ctx(tb, coord)
    d.step(a.oracle_steps)
    st, db = d.state(), d.dbsph_state()
    exact = [f for f in FIELDS if np.array_equal(st[f], ref[f])]
    exact_db = [f for f in ("gamma", "sigma", "rho_sph", "dshep", "fs", "wall_dens", "wall_pres")
                if np.array_equal(db[f], rdb[f])]
    d.close()
    out["parity"] = dict(steps=a.oracle_steps, bit_identical_fields=exact, bit_identical_dbsph=exact_db,
                         all_bit_identical=len(exact) == len(FIELDS) and len(exact_db) == 7)
    out["cpu_baseline_oracle"] = dict(ms_per_step=t_or / a.oracle_steps * 1e3, cores=os.cpu_count(),
                                      kind="port", sample=f"{a.oracle_steps} steps of the oracle")
    print(f"[bench_dbsph] parity after {a.oracle_steps} steps: {out['parity']['all_bit_identical']}",
          file=sys.stderr, flush=True)
    # the timed device run
    d = _ctx(tb, coord)
    d.step_graph(1)  # capture outside the timed region
    d.sync()
    p0 = d.pair_counter()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        d.step_graph(1)
    d.sync()
    t1 = time.perf_counter()
    pairs = d.pair_counter() - p0
    d.close()
    out["ms_per_step"] = (t1 - t0) / a.steps * 1e3
    out["pair_interactions_per_s"] = pairs / (t1 - t0)
    out["particle_steps_per_s"] = n * a.steps / (t1 - t0)
    print(f"[bench_dbsph] {out['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    # kernel times (events around each launch: eager steps)
    d = _ctx(tb, coord)
    d.set_timing(True)
    d.step(20)
    d.sync()
    tm = d.timing()
    d.close()
    tot = sum(v[0] for v in tm.values())
    out["kernels"] = {k: dict(ms_per_call=v[0] / max(v[1], 1), share=v[0] / tot) for k, v in
                      sorted(tm.items(), key=lambda kv: -kv[1][0])}
    # the reference itself on the host cores: two runs, their difference
    from oracle.decks import HARNESS, run_harness
    if HARNESS.exists() and not a.no_ref:
        threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
        times = []
        for k in (1, 1 + a.ref_steps):
            with tempfile.TemporaryDirectory(prefix="sph_refdb_") as wd:
                t0 = time.perf_counter()
                run_harness(case, k, every=k, lists=False, threads=threads, workdir=wd, ic_coord=coord,
                            timeout=3600)
                times.append(time.perf_counter() - t0)
        out["reference_cpu"] = dict(ms_per_step=(times[1] - times[0]) / a.ref_steps * 1e3, cores=threads,
                                    kind="reference",
                                    sample=f"{a.ref_steps} steps: runs of 1 and {1 + a.ref_steps} steps took "
                                           f"{times[0]:.2f} s and {times[1]:.2f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
