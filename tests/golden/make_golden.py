#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ by running the flang build of the
SPHERA v10 reference (oracle/_ref/harness3d.x, built by oracle/Makefile.ref
from /root/reference/src) on project-authored decks.

    make -f oracle/Makefile.ref -j8 && python tests/golden/make_golden.py [fixture ...]

With names (tiny, small_jitter, removal, pelo, 2d, nearwall, nearwall_2d, dclamp,
rk1, viscous, wallfn, wallfn_2d, edges, outlet, outlet_2d, dbsph) only those
fixtures are written.

Fixtures (numpy .npz, inputs and reference outputs only):
  ref_tiny.npz          800 particles, lattice IC; state + neighbour lists +
                        SA-SPH boundary table after the pre-loop and after 5
                        steps; set-up tables of the header
  ref_small_jitter.npz  1440 particles, lattice + U(+-0.01 dx) jitter from
                        seed 20260101; state after 0 and 10 steps
  ref_removal.npz       800 particles of which 6 start outside the grid (three
                        holes, three at the tail): OrdGrid1's swap-with-last
                        order; state after the pre-loop and after 6 steps
  ref_pelo.npz          800 particles, jittered, with a block lifted above a
                        gap and one lone particle; four monitoring lines (two
                        partly outside the grid): the reference's free-surface
                        monitor (.plb, .zlft at g14.7) after 0, 3 and 6 steps
  ref_2d.npz            the 2D build (oracle/_ref/harness2d.x): 600 particles
                        in the x-z plane, x and z jittered U(+-0.01 dx) from
                        seed 20260303; boundary sides of the header; state,
                        neighbour lists and 2D boundary table after the
                        pre-loop and after 10 steps
  ref_nearwall.npz      800 particles; the layers next to the floor and to the
                        x = 0 wall moved to 0.02-0.22 dx from it and given
                        velocities into it (seed 20261203): the elastic wall
                        reaction (below 0.25 dx, approaching) and the zpmin
                        clamp of the volume integrals (below 0.05 h) run;
                        state, lists and boundary table after 0, 4 and 8 steps
  ref_nearwall_2d.npz   the same for the 2D build (600 particles, floor and
                        x = 0 side), after 0, 4 and 8 steps
  ref_viscous.npz       800 particles, dynamic viscosity 20 Pa s and a shear
                        flow u = 3 z: the laminar flag of inter_EqMoto.f90:
                        269-283 is set for hundreds of particles and feeds the
                        wall shear term; state after 0, 3 and 6 steps
  ref_wallfn.npz        800 particles, jittered lattice with the near-wall
                        layers of ref_nearwall.npz and a tangential velocity
                        (v_y = 0.2 m/s, v_x += 2 z); the frontier zone has
                        slip_coefficient_mode = 2 with eps_sigma_z = 0.0005,
                        H_sgr = 0.0015 (roughness length 0.0005 m): the SA-SPH
                        wall function on both of its branches, in the momentum
                        equation and the velocity smoothing; state after 0, 4
                        and 8 steps
  ref_wallfn_2d.npz     the same for the 2D build (600 particles, v_x = 0.2 +
                        2 z, v_z += 0.2)
  ref_edges.npz         800 particles; a block on the floor across the tank
                        (0.15 < x < 0.2 m, top at z = 0.04 m) adds three
                        frontier faces and two convex edges
                        (FindBoundaryConvexEdges3D); the top layer of the
                        column (80 particles) is moved next to the block, in
                        four groups of 20 moving into it: beyond the first
                        edge (within 0.16 dx of it, outside both faces), in
                        front of the upstream face, above the top face and
                        beyond the second edge; convex-edge and face
                        reactions run; state after 0, 4 and 8 steps (one
                        OpenMP thread: the reference's edge order is the
                        threads' arrival order)
  ref_outlet.npz        800 particles; the tank's x = 0.3 m wall is an open
                        face (zone "open"); the jittered column shifted to
                        0.224 < x < 0.3 m and moving towards it at
                        U(2, 4) m/s (seed 20261205), the 10 top particles
                        next to it lifted above the wall's top edge, five
                        particles just below the grid's top moving up:
                        particles crossing the face are removed
                        (CancelOutgoneParticles_3D, then
                        OrdGrid1's swap-with-last), those crossing its
                        plane above it stay, those leaving the grid are
                        removed; state (with the reference's
                        particle order) after 0, 8 and 16 steps,
                        leapfrog, and the same with explicit RK1
                        (ref_outlet_rk1.npz)
  ref_dbsph.npz         DB-SPH walls (Domain%tipo = "bsph", RK1): the tiny
                        tank as a surface mesh (sphera_amd.dbsph.tank_mesh:
                        3,800 triangles), 700 particles on the DB-SPH
                        lattice, U(+-0.01 dx) jitter, v_x U(0, 1) m/s,
                        v_z -U(0, 0.5) m/s (seed 20261207); the wall elements
                        of the header; state, DB-SPH state (Gamma, sigma,
                        rhoSPH_new, dShep, MUSCL gradients, FS, the wall
                        elements' density, pressure and wet flags) and the
                        fluid-wall lists after 0, 10 and 20 steps (fluid
                        lists after 0); ref_dbsph_opts.npz the same with the
                        Gamma limiter, negative wall pressures allowed and no
                        free-surface detection
  ref_outlet_2d.npz     the 2D build (600 particles) with the tank's x = 0.6 m
                        side open (zone 3 "open"): the jittered column
                        shifted by 0.3035 m and moving towards it at
                        U(2, 4) m/s (seed 20261206), the top four next to it
                        lifted above the side's end, three just below the
                        grid's top moving up (CancelOutgoneParticles_2D);
                        state after 0, 8 and 16 steps, leapfrog, and the
                        same with explicit RK1 (ref_outlet_2d_rk1.npz)
  ref_rk1.npz           1440 particles, jittered lattice (seed 20260101), the
                        explicit RK1 branch (time_split = 0, RKscheme = 1:
                        Euler); state after 0, 6 and 12 steps
  ref_dclamp.npz        1440 particles with density_thresholds = 1 and a
                        velocity field converging on the left half of the
                        column and diverging on the right (150/s, jittered
                        lattice): the 0.9/1.1 rho0 density clamp runs; state
                        after 0, 3 and 6 steps
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle.decks import read_dump, read_pelo, run_harness  # noqa: E402
from sphera_amd.case import DamBreak3D, build, control_line_points  # noqa: E402
from sphera_amd import case2d  # noqa: E402

TINY = dict(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny")
SMALL = dict(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01, name="small")
FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")
PELO_LINES = (("v_gap", (0.015, 0.05, 0.0), (0.015, 0.05, 0.19), 20),
              ("v_out", (0.071, 0.021, -0.05), (0.071, 0.021, 0.25), 31),
              ("h_out", (-0.02, 0.05, 0.05), (0.32, 0.05, 0.05), 35),
              ("diag", (0.0, 0.0, 0.0), (0.3, 0.1, 0.2), 41))


def pelo_ic(coord, dx):
    """Jittered lattice; particles with x < 0.03 and z > 0.06 lifted by 0.06
    (fluid above an empty gap) and one particle moved alone to the top."""
    ic = jitter(coord, dx, seed=20260202)
    m = (ic[:, 0] < 0.03) & (ic[:, 2] > 0.06)
    ic[m, 2] += 0.06
    ic[7] = (0.0148, 0.0497, 0.1852)
    return ic


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


def make_tiny():
    """lists and boundary table"""
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


def make_small_jitter():
    """small, jittered: fields over 10 steps"""
    case = DamBreak3D(**SMALL)
    tb = build(case)
    ic = jitter(tb.particles["coord"], case.dx)
    hdr, steps = read_dump(run_harness(case, nsteps=10, every=10, lists=False, ic_coord=ic))
    out = {"ic_coord": ic}
    for st in steps:
        pack_step(f"s{st['it']}_", st, out, lists=False)
    np.savez_compressed(HERE / "ref_small_jitter.npz", **out)


def make_removal():
    """removal order: particles placed outside the grid"""
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


def make_pelo():
    """free-surface monitor lines"""
    case = DamBreak3D(**TINY, control_lines=PELO_LINES)
    tb = build(case)
    ic = pelo_ic(tb.particles["coord"], case.dx)
    dump = run_harness(case, nsteps=6, every=3, lists=False, ic_coord=ic,
                       extra_env={"SPH_HARNESS_PELO": "1"})
    t, pel, zl = read_pelo(dump.parent, case)
    icont, pts = control_line_points(PELO_LINES)
    np.savez_compressed(HERE / "ref_pelo.npz", ic_coord=ic, icont=icont, points=pts,
                        time=t, pelolib=pel, zlft=zl)


CASE2D = dict(tank=(0.6, 0.4), column=(0.3, 0.2), dx=0.01, name="t2d")


def jitter2d(coord, dx, seed=20260303, amp=0.01):
    """U(+-amp dx) on x and z only (the 2D particles stay at y = 0)."""
    rng = np.random.default_rng(seed)
    out = coord.copy()
    j = rng.uniform(-amp, amp, (len(coord), 2)) * dx
    out[:, 0] += j[:, 0]
    out[:, 2] += j[:, 1]
    return out


def make_2d():
    """2D build: sides, lists, 2D boundary table and fields"""
    case = case2d.DamBreak2D(**CASE2D)
    tb = case2d.build(case)
    ic = jitter2d(tb.particles["coord"], case.dx)
    hdr, steps = read_dump(run_harness(case, nsteps=10, every=10, ic_coord=ic))
    out = {"ic_coord": ic}
    for k in ("nag", "nmaxpartj", "ncd", "extr_lo", "dcd", "h", "coefke", "coefkacl", "eta",
              "eta2", "celerita", "kin_visc"):
        out["hdr_" + k] = np.asarray(hdr[k])
    sd = hdr["sides"]
    out["hdr_side_stretch"] = np.array([x["stretch"] for x in sd])
    out["hdr_side_tipo"] = np.array([x["tipo"] for x in sd])
    out["hdr_side_previous"] = np.array([x["previous"] for x in sd])
    for k in ("length", "angle", "T", "RN", "v1", "v2"):
        out["hdr_side_" + k] = np.array([x[k] for x in sd])
    for st in steps:
        pre = f"s{st['it']}_"
        out[pre + "it"] = st["it"]
        out[pre + "time"] = st["time"]
        out[pre + "dt"] = st["dt"]
        out[pre + "dt_average"] = st["dt_average"]
        for f in FIELDS + ("var", "laminar_flag"):
            out[pre + f] = st[f]
        out[pre + "nbr_count"] = st["nbr_count"]
        out[pre + "nbr"] = st["nbr"]
        out[pre + "ncbs"] = st["ncbs"]
        out[pre + "intncbs"] = st["intncbs"]
        out[pre + "bt_side"] = st["btab"]["face"]
        out[pre + "bt_loc"] = st["btab"]["loc"]
        out[pre + "bt_bi"] = st["btab"]["bi"]
    np.savez_compressed(HERE / "ref_2d.npz", **out)


def nearwall_ic(coord, dx, seed=20261203, dim=3):
    """Particles of the layer at the floor (z < dx) and of the layer at the x = 0
    wall (x < dx) moved to U(0.02, 0.22) dx from the wall, with velocities
    U(0, 0.05) m/s into it."""
    rng = np.random.default_rng(seed)
    c = coord.copy()
    v = np.zeros_like(c)
    bot = c[:, 2] < dx
    side = c[:, 0] < dx
    c[bot, 2] = rng.uniform(0.02, 0.22, bot.sum()) * dx
    c[side, 0] = rng.uniform(0.02, 0.22, side.sum()) * dx
    v[bot, 2] = -rng.uniform(0.0, 0.05, bot.sum())
    v[side, 0] = -rng.uniform(0.0, 0.05, side.sum())
    return c, v


def dclamp_vel(coord, rate=150.0):
    """Converging on the left half of the column (x below its mean), diverging
    on the right: |div v| = 3 rate."""
    ctr = coord.mean(axis=0)
    d = coord - ctr
    sgn = np.where(d[:, 0] < 0.0, -1.0, 1.0)[:, None]
    return sgn * rate * d


def _pack_steps(steps, out, lists):
    for st in steps:
        pre = f"s{st['it']}_"
        out[pre + "nag"] = st["nag"]
        pack_step(pre, st, out, lists=lists)


def make_nearwall():
    """3D near-wall: elastic reaction and zpmin clamp"""
    case = DamBreak3D(**TINY)
    tb = build(case)
    ic, iv = nearwall_ic(tb.particles["coord"], case.dx)
    hdr, steps = read_dump(run_harness(case, nsteps=8, every=4, ic_coord=ic, ic_vel=iv))
    out = {"ic_coord": ic, "ic_vel": iv}
    _pack_steps(steps, out, lists=True)
    np.savez_compressed(HERE / "ref_nearwall.npz", **out)


def make_nearwall_2d():
    """2D near-wall: elastic reaction of the 2D build"""
    case = case2d.DamBreak2D(**CASE2D)
    tb = case2d.build(case)
    ic, iv = nearwall_ic(tb.particles["coord"], case.dx, dim=2)
    hdr, steps = read_dump(run_harness(case, nsteps=8, every=4, lists=False, ic_coord=ic, ic_vel=iv))
    out = {"ic_coord": ic, "ic_vel": iv}
    for st in steps:
        pre = f"s{st['it']}_"
        for k in ("it", "time", "dt", "dt_average", "nag"):
            out[pre + k] = st[k]
        for f in FIELDS + ("var", "laminar_flag"):
            out[pre + f] = st[f]
    np.savez_compressed(HERE / "ref_nearwall_2d.npz", **out)


OBSTACLE = (0.15, 0.2, 0.04)


def obstacle_ic(coord, dx, seed=20261204):
    """The column's top layer (80 particles) moved next to the block x0 < x <
    x1, z < zt, in four groups of 20 spread over y, each moving into the block
    at U(0.02, 0.08) m/s per component: beyond the edge (x0, zt) at
    U(0.03, 0.16) dx from both planes, in front of the upstream face
    (z < zt), above the top face (x0 < x < x1), beyond the edge (x1, zt)."""
    rng = np.random.default_rng(seed)
    x0, x1, zt = OBSTACLE
    c = coord.copy()
    v = np.zeros_like(c)
    top = np.flatnonzero(c[:, 2] > c[:, 2].max() - 0.5 * dx)
    assert len(top) == 80, len(top)
    a = rng.uniform(0.03, 0.16, (80, 2)) * dx
    sp = rng.uniform(0.02, 0.08, (80, 2))
    y = np.tile(0.0025 + 0.005 * np.arange(20), 4) + rng.uniform(-0.001, 0.001, 80)
    for g in range(4):
        idx = top[20 * g:20 * (g + 1)]
        s = slice(20 * g, 20 * (g + 1))
        if g == 0:    # beyond the first edge, moving towards it
            c[idx] = np.stack([x0 - a[s, 0], y[s], zt + a[s, 1]], 1)
            v[idx] = np.stack([sp[s, 0], 0 * y[s], -sp[s, 1]], 1)
        elif g == 1:  # in front of the upstream face
            c[idx] = np.stack([x0 - a[s, 0], y[s], rng.uniform(0.3, 0.8, 20) * zt], 1)
            v[idx] = np.stack([sp[s, 0], 0 * y[s], 0 * y[s]], 1)
        elif g == 2:  # above the top face
            c[idx] = np.stack([rng.uniform(x0 + 0.1 * (x1 - x0), x1 - 0.1 * (x1 - x0), 20), y[s],
                               zt + a[s, 1]], 1)
            v[idx] = np.stack([0 * y[s], 0 * y[s], -sp[s, 1]], 1)
        else:         # beyond the second edge (the reference returns at the first edge)
            c[idx] = np.stack([x1 + a[s, 0], y[s], zt + a[s, 1]], 1)
            v[idx] = np.stack([-sp[s, 0], 0 * y[s], -sp[s, 1]], 1)
    return c, v


def make_edges():
    """a block on the floor: convex-edge elastic reaction"""
    case = DamBreak3D(**TINY, obstacle=OBSTACLE)
    tb = build(case)
    ic, iv = obstacle_ic(tb.particles["coord"], case.dx)
    # one thread: FindBoundaryConvexEdges3D.f90:69-146 appends the edges it
    # finds from an OpenMP loop in thread-arrival order, and the reaction
    # loop returns at the first edge that does not apply, so with several
    # threads the reacting particles change from run to run
    hdr, steps = read_dump(run_harness(case, nsteps=8, every=4, lists=False, ic_coord=ic, ic_vel=iv,
                                       threads=1))
    out = {"ic_coord": ic, "ic_vel": iv}
    _pack_steps(steps, out, lists=False)
    np.savez_compressed(HERE / "ref_edges.npz", **out)


def outlet_ic(coord, dx, seed=20261205):
    """The jittered column shifted by 0.224 m towards the open x = 0.3 m wall,
    moving into it at U(2, 4) m/s, with U(-0.5, 0.5) m/s vertically."""
    rng = np.random.default_rng(seed)
    c = jitter(coord, dx)
    c[:, 0] += 0.224
    v = np.zeros_like(c)
    v[:, 0] = rng.uniform(2.0, 4.0, len(c))
    v[:, 2] = rng.uniform(-0.5, 0.5, len(c))
    # the top particles of the layer next to the wall lifted above the tank
    # (0.205-0.215 m, the open face ends at 0.2 m): they cross the face's
    # plane outside the face and stay
    zt = c[:, 2].max() - 0.5 * dx
    top = np.flatnonzero((c[:, 0] > c[:, 0].max() - 0.5 * dx) & (c[:, 2] > zt))
    up = np.flatnonzero((c[:, 0] < c[:, 0].min() + 0.5 * dx) & (c[:, 2] > zt))[:5]
    c[top, 2] = rng.uniform(0.205, 0.215, len(top))
    v[top, 2] = 0.0
    # five particles of the column's top layer, at its upstream end, put just
    # below the grid's top (z = 0.2261 m) and moving up: they leave the grid
    # during the run (OrdGrid1 removes them; the next dt excludes them)
    c[up, 2] = rng.uniform(0.2250, 0.2258, len(up))
    v[up] = [0.0, 0.0, 3.0]
    return c, v


def make_outlet():
    """an open face: particles leaving through it are removed"""
    for ts, name in ((1, "ref_outlet.npz"), (0, "ref_outlet_rk1.npz")):
        case = DamBreak3D(**TINY, outlet=True, time_split=ts)
        tb = build(case)
        ic, iv = outlet_ic(tb.particles["coord"], case.dx)
        hdr, steps = read_dump(run_harness(case, nsteps=16, every=8, lists=False, ic_coord=ic, ic_vel=iv))
        out = {"ic_coord": ic, "ic_vel": iv}
        _pack_steps(steps, out, lists=False)
        np.savez_compressed(HERE / name, **out)


def outlet_2d_ic(coord, dx, seed=20261206):
    """The jittered 2D column shifted by 0.3035 m towards the open x = 0.6 m
    side, moving into it at U(2, 4) m/s, with U(-0.5, 0.5) m/s vertically."""
    rng = np.random.default_rng(seed)
    c = jitter2d(coord, dx)
    c[:, 0] += 0.3035
    v = np.zeros_like(c)
    v[:, 0] = rng.uniform(2.0, 4.0, len(c))
    v[:, 2] = rng.uniform(-0.5, 0.5, len(c))
    # the top four particles of the column next to the side lifted above the
    # tank (0.405-0.415 m, the open side ends at 0.4 m): they cross the
    # side's line beyond its end and stay
    right = np.flatnonzero(c[:, 0] > c[:, 0].max() - 0.5 * dx)
    top = right[np.argsort(c[right, 2])[-4:]]
    left = np.flatnonzero(c[:, 0] < c[:, 0].min() + 0.5 * dx)
    up = left[np.argsort(c[left, 2])[-3:]]
    c[top, 2] = rng.uniform(0.405, 0.415, len(top))
    v[top, 2] = 0.0
    # three of the upstream column's top put just below the grid's top
    # (z = 0.4261 m) and moving up: they leave the grid during the run
    c[up, 2] = rng.uniform(0.4250, 0.4258, len(up))
    v[up] = [0.0, 0.0, 3.0]
    return c, v


def make_outlet_2d():
    """2D open side: particles leaving through it are removed"""
    for ts, name in ((1, "ref_outlet_2d.npz"), (0, "ref_outlet_2d_rk1.npz")):
        case = case2d.DamBreak2D(**CASE2D, outlet=True, time_split=ts)
        tb = case2d.build(case)
        ic, iv = outlet_2d_ic(tb.particles["coord"], case.dx)
        hdr, steps = read_dump(run_harness(case, nsteps=16, every=8, lists=False, ic_coord=ic, ic_vel=iv))
        out = {"ic_coord": ic, "ic_vel": iv}
        for st in steps:
            pre = f"s{st['it']}_"
            for k in ("it", "time", "dt", "dt_average", "nag"):
                out[pre + k] = st[k]
            for f in FIELDS + ("var", "laminar_flag"):
                out[pre + f] = st[f]
        np.savez_compressed(HERE / name, **out)


def dbsph_ic(coord, dx, seed=20261207):
    """U(+-0.01 dx) jitter; v_x U(0, 1) m/s, v_z -U(0, 0.5) m/s."""
    rng = np.random.default_rng(seed)
    c = coord + rng.uniform(-0.01, 0.01, coord.shape) * dx
    v = np.zeros_like(c)
    v[:, 0] = rng.uniform(0.0, 1.0, len(c))
    v[:, 2] = -rng.uniform(0.0, 0.5, len(c))
    return c, v


DBSPH_OPTS = {"ref_dbsph.npz": dict(),
              "ref_dbsph_opts.npz": dict(gamma_limiter=True, negative_wall_p_allowed=True,
                                         fs_allowed=False),
              "ref_dbsph_muscl.npz": dict(muscl_boundary=True),
              # the tank moves: a translation ramping up to 0.5 m/s along x and a
              # rotation about the vertical through the tank's floor centre
              # ramping up to 4 rad/s (wall speeds up to about 1 m/s)
              "ref_dbsph_moving.npz": dict(kinematics=((0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
                                                       (0.001, 0.5, 0.0, 0.0, 0.0, 0.0, 4.0),
                                                       (1.0, 0.5, 0.0, 0.0, 0.0, 0.0, 4.0)),
                                           rotation_centre=(0.15, 0.05, 0.0))}


def make_dbsph(names=None):
    """DB-SPH walls: a tank mesh, RK1 (names: a subset of DBSPH_OPTS)"""
    from sphera_amd.dbsph import DBSPHParams
    for name, opts in DBSPH_OPTS.items():
        if names is not None and name not in names:
            continue
        case = DamBreak3D(**TINY, time_split=0, dbsph=DBSPHParams(**opts))
        tb = build(case)
        ic, iv = dbsph_ic(tb.particles["coord"], case.dx)
        hdr, steps = read_dump(run_harness(case, nsteps=20, every=10, lists=True, ic_coord=ic, ic_vel=iv))
        out = {"ic_coord": ic, "ic_vel": iv}
        for k, v in hdr["walls"].items():
            out["wall_" + k] = v
        for st in steps:
            pre = f"s{st['it']}_"
            for k in ("it", "time", "dt", "dt_average", "nag"):
                out[pre + k] = st[k]
            for f in FIELDS + ("var", "laminar_flag", "gamma", "sigma", "rho_sph", "dshep", "drho", "dvel",
                               "fs", "wall_dens", "wall_pres", "wall_wet", "fw_count", "fw"):
                out[pre + f] = st[f]
            if st["it"] == 0:
                out[pre + "nbr_count"] = st["nbr_count"]
                out[pre + "nbr"] = st["nbr"]
        np.savez_compressed(HERE / name, **out)


WALLFN = dict(slip_mode=2, roughness=(0.0005, 0.0015))


def make_wallfn():
    """slip_coefficient_mode = 2: the SA-SPH wall function"""
    case = DamBreak3D(**TINY, **WALLFN)
    tb = build(case)
    ic, iv = nearwall_ic(jitter(tb.particles["coord"], case.dx), case.dx)
    iv[:, 1] += 0.2
    iv[:, 0] += 2.0 * ic[:, 2]
    hdr, steps = read_dump(run_harness(case, nsteps=8, every=4, lists=False, ic_coord=ic, ic_vel=iv))
    out = {"ic_coord": ic, "ic_vel": iv}
    _pack_steps(steps, out, lists=False)
    np.savez_compressed(HERE / "ref_wallfn.npz", **out)


def make_wallfn_2d():
    """2D wall function"""
    case = case2d.DamBreak2D(**CASE2D, **WALLFN)
    tb = case2d.build(case)
    ic, iv = nearwall_ic(jitter2d(tb.particles["coord"], case.dx), case.dx, dim=2)
    iv[:, 0] += 0.2 + 2.0 * ic[:, 2]
    iv[:, 2] += 0.2
    hdr, steps = read_dump(run_harness(case, nsteps=8, every=4, lists=False, ic_coord=ic, ic_vel=iv))
    out = {"ic_coord": ic, "ic_vel": iv}
    for st in steps:
        pre = f"s{st['it']}_"
        for k in ("it", "time", "dt", "dt_average", "nag"):
            out[pre + k] = st[k]
        for f in FIELDS + ("var", "laminar_flag"):
            out[pre + f] = st[f]
    np.savez_compressed(HERE / "ref_wallfn_2d.npz", **out)


def make_dclamp():
    """density_thresholds = 1: the density clamp"""
    case = DamBreak3D(**SMALL, density_thresholds=1)
    tb = build(case)
    ic = jitter(tb.particles["coord"], case.dx)
    iv = dclamp_vel(ic)
    hdr, steps = read_dump(run_harness(case, nsteps=6, every=3, lists=False, ic_coord=ic, ic_vel=iv))
    out = {"ic_coord": ic, "ic_vel": iv}
    _pack_steps(steps, out, lists=False)
    np.savez_compressed(HERE / "ref_dclamp.npz", **out)


def make_viscous():
    """viscous shear flow: laminar flags"""
    case = DamBreak3D(**TINY, mu=20.0)
    tb = build(case)
    c = tb.particles["coord"]
    iv = np.zeros_like(c)
    iv[:, 0] = 3.0 * c[:, 2]
    hdr, steps = read_dump(run_harness(case, nsteps=6, every=3, lists=False, ic_vel=iv))
    out = {"ic_vel": iv}
    _pack_steps(steps, out, lists=False)
    np.savez_compressed(HERE / "ref_viscous.npz", **out)


def make_rk1():
    """explicit RK1 (Euler) time integration"""
    case = DamBreak3D(**SMALL, time_split=0)
    tb = build(case)
    ic = jitter(tb.particles["coord"], case.dx)
    hdr, steps = read_dump(run_harness(case, nsteps=12, every=6, lists=False, ic_coord=ic))
    out = {"ic_coord": ic}
    _pack_steps(steps, out, lists=False)
    np.savez_compressed(HERE / "ref_rk1.npz", **out)


MAKERS = {"tiny": make_tiny, "small_jitter": make_small_jitter, "removal": make_removal,
          "pelo": make_pelo, "2d": make_2d, "nearwall": make_nearwall,
          "nearwall_2d": make_nearwall_2d, "dclamp": make_dclamp, "rk1": make_rk1,
          "viscous": make_viscous, "wallfn": make_wallfn, "wallfn_2d": make_wallfn_2d,
          "edges": make_edges, "outlet": make_outlet, "outlet_2d": make_outlet_2d,
          "dbsph": make_dbsph}


def main():
    for name in sys.argv[1:] or list(MAKERS):
        MAKERS[name]()
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
