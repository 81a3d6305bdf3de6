"""The oracle (oracle/sph_oracle.cpp) against the reference itself.

1. Against the committed golden fixtures written by tests/golden/make_golden.py
   from the flang build of the reference: bit-identical neighbour lists (in
   the reference's order), SA-SPH boundary table and fields.
2. Where oracle/_ref/harness3d.x is built (this container), against a live
   reference run: longer horizons, jittered input.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from sphera_amd.case import DamBreak3D, build

GOLD = Path(__file__).resolve().parent / "golden"
FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")
HARNESS = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "harness3d.x"


def _tiny():
    return DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny")


def _small():
    return DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01, name="small")


def _check_lists(o, counts, flat):
    cnt, ids = o.lists()
    assert np.array_equal(cnt, counts)
    off = np.concatenate([[0], np.cumsum(counts)])
    for i in range(len(cnt)):
        assert np.array_equal(ids[i, :cnt[i]] + 1, flat[off[i]:off[i + 1]]), i


def _check_btab(o, tb, ncbf, face, loc, bi, ig):
    nc, f, l, b, g = o.btab()
    assert np.array_equal(nc, ncbf)
    k = 0
    for i in range(len(nc)):
        for q in range(nc[i]):
            assert tb.faces.face_id[f[i, q]] == face[k]
            assert np.array_equal(l[i, q], loc[k])
            assert np.array_equal(b[i, q], bi[k])
            assert np.array_equal(g[i, q], ig[k])
            k += 1


def test_golden_tiny_bit_identical():
    g = np.load(GOLD / "ref_tiny.npz")
    tb = build(_tiny())
    o = Oracle(tb)
    o.preloop()
    st = o.state()
    assert st["dt"] == g["s0_dt"] and st["dt_average"] == g["s0_dt_average"]
    _check_lists(o, g["s0_nbr_count"], g["s0_nbr"])
    _check_btab(o, tb, g["s0_ncbf"], g["s0_bt_face"], g["s0_bt_loc"], g["s0_bt_bi"], g["s0_bt_ig"])
    o.step(5)
    st = o.state()
    for f in FIELDS:
        assert np.array_equal(st[f], g["s5_" + f]), f
    assert st["dt"] == g["s5_dt"] and st["time"] == g["s5_time"]
    _check_lists(o, g["s5_nbr_count"], g["s5_nbr"])
    _check_btab(o, tb, g["s5_ncbf"], g["s5_bt_face"], g["s5_bt_loc"], g["s5_bt_bi"], g["s5_bt_ig"])


def test_golden_small_jitter_bit_identical():
    g = np.load(GOLD / "ref_small_jitter.npz")
    tb = build(_small())
    o = Oracle(tb, coord=g["ic_coord"])
    o.preloop()
    o.step(10)
    st = o.state()
    for f in FIELDS:
        assert np.array_equal(st[f], g["s10_" + f]), f
    assert np.array_equal(o.state()["coord"], g["s10_coord"])


def test_state_restart_from_reference_step():
    """Set the oracle to the reference's step-0 state, run 5 steps, compare."""
    g = np.load(GOLD / "ref_tiny.npz")
    tb = build(_tiny())
    o = Oracle(tb)
    st = {k: g["s0_" + k] for k in ("coord", "vel", "acc", "dens", "pres", "laminar_flag", "dt",
                                      "dt_average", "time", "it")}
    st["var"] = g["s0_vel"]
    st["nag"] = len(st["coord"])
    st["it"] = int(st["it"])
    o.set_state(st)
    o.call("calcvarlength")
    o.call("boundary_tab")
    o.step(5)
    s = o.state()
    for f in ("coord", "vel", "dens", "pres"):
        assert np.array_equal(s[f], g["s5_" + f]), f


@pytest.mark.skipif(not HARNESS.exists(), reason="reference not built here")
@pytest.mark.parametrize("jit", [False, True])
def test_live_reference_run(jit, tmp_path):
    from oracle.decks import read_dump, run_harness
    case = DamBreak3D(tank=(0.5, 0.2, 0.3), column=(0.15, 0.2, 0.15), dx=0.01, name="live")
    tb = build(case)
    ic = None
    if jit:
        rng = np.random.default_rng(7)
        ic = tb.particles["coord"] + rng.uniform(-0.01, 0.01, tb.particles["coord"].shape) * case.dx
    hdr, steps = read_dump(run_harness(case, nsteps=60, every=30, lists=False, ic_coord=ic,
                                       workdir=tmp_path))
    o = Oracle(tb, coord=ic)
    o.preloop()
    prev = 0
    for r in steps[1:]:
        o.step(r["it"] - prev)
        prev = r["it"]
        st = o.state()
        assert st["dt"] == r["dt"]
        for f in FIELDS:
            assert np.array_equal(st[f], r[f]), (f, r["it"])


@pytest.mark.skipif(not HARNESS.exists(), reason="reference not built here")
def test_live_reference_removal_order(tmp_path):
    """Particles outside the grid are removed by OrdGrid1's swap-with-last
    (OrdGrid1.f90:62-83), which renumbers pg: the oracle keeps the same order
    (tail particles moved into the holes, removed tails dropped)."""
    from oracle.decks import read_dump, run_harness
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="rm")
    tb = build(case)
    ic = tb.particles["coord"].copy()
    n = len(ic)
    for k in (3, 17, 200, n - 4, n - 2, n - 1):  # holes, and removed particles at the tail
        ic[k] = (50.0, 50.0, 50.0)
    hdr, steps = read_dump(run_harness(case, nsteps=6, every=3, lists=False, ic_coord=ic,
                                       workdir=tmp_path))
    o = Oracle(tb, coord=ic)
    o.preloop()
    assert o.n == n - 6 == steps[0]["nag"]
    prev = 0
    for r in steps:
        o.step(r["it"] - prev)
        prev = r["it"]
        st = o.state()
        for f in FIELDS:
            assert np.array_equal(st[f], r[f]), (f, r["it"])


def test_golden_removal_order():
    """Six particles start outside the grid (three holes, three at the tail):
    the oracle's pg order after OrdGrid1's swap-with-last equals the
    reference's, and so do the fields after 6 steps."""
    g = np.load(GOLD / "ref_removal.npz")
    tb = build(_tiny())
    o = Oracle(tb, coord=g["ic_coord"])
    o.preloop()
    assert o.n == int(g["s0_nag"]) == len(g["ic_coord"]) - 6
    st = o.state()
    for f in FIELDS:
        assert np.array_equal(st[f], g["s0_" + f]), ("s0", f)
    o.step(6)
    st = o.state()
    for f in FIELDS:
        assert np.array_equal(st[f], g["s6_" + f]), ("s6", f)


def _g14_close(a, b):
    """Equality to the 7 significant digits of the reference's g14.7 output."""
    return np.allclose(a, b, rtol=1e-6, atol=1e-12)


def test_golden_free_surface():
    """calc_pelo (oracle or_free_surface) against the reference's .plb/.zlft
    output for four monitoring lines: a vertical line through fluid lifted
    above an empty gap and a lone particle on top (zlft stops below the gap,
    pelolib is the lone particle + dx/2), lines partly outside the grid, and a
    diagonal.  The files hold 7 significant digits; picking another particle
    than the two closest moves a value by ~dx/2 = 5e-3."""
    g = np.load(GOLD / "ref_pelo.npz")
    tb = build(_tiny())
    o = Oracle(tb, coord=g["ic_coord"])
    o.preloop()
    prev = 0
    for k, it in enumerate((0, 3, 6)):
        o.step(it - prev)
        prev = it
        pel, zl = o.free_surface(g["icont"], g["points"])
        assert _g14_close(o.state()["time"], g["time"][k])
        assert _g14_close(pel, g["pelolib"][k]), (it, pel, g["pelolib"][k])
        assert _g14_close(zl, g["zlft"][k]), (it, zl, g["zlft"][k])
    # the fixture exercises the gap, the lone particle and a one-particle cell
    assert g["zlft"][0, 0, 2] < 0.07 < 0.19 < g["pelolib"][0, 0, 2]
    assert np.allclose(g["pelolib"][0, 0], g["ic_coord"][7] + 0.005, atol=1e-9)


@pytest.mark.skipif(not HARNESS.exists(), reason="reference not built here")
def test_live_free_surface(tmp_path):
    """A second case through the live reference: 0.005 m spacing, jittered,
    lines crossing the whole tank along x, y and z, over 8 steps."""
    from oracle.decks import read_pelo, run_harness
    from sphera_amd.case import control_line_points
    lines = (("x", (-0.01, 0.03, 0.04), (0.41, 0.03, 0.04), 60),
             ("y", (0.05, -0.02, 0.1), (0.05, 0.22, 0.1), 25),
             ("z", (0.101, 0.151, -0.01), (0.101, 0.151, 0.31), 70))
    case = DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.2, 0.12), dx=0.005, name="live",
                      control_lines=lines)
    tb = build(case)
    rng = np.random.default_rng(11)
    ic = tb.particles["coord"] + rng.uniform(-0.02, 0.02, tb.particles["coord"].shape) * case.dx
    dump = run_harness(case, nsteps=8, every=4, lists=False, ic_coord=ic, workdir=tmp_path,
                       extra_env={"SPH_HARNESS_PELO": "1"})
    t, pel_r, zl_r = read_pelo(dump.parent, case)
    icont, pts = control_line_points(lines)
    o = Oracle(tb, coord=ic)
    o.preloop()
    for k in range(3):
        if k:
            o.step(4)
        pel, zl = o.free_surface(icont, pts)
        assert _g14_close(pel, pel_r[k]) and _g14_close(zl, zl_r[k]), k


def _oracle_steps(o, g, its, lists=False, tb=None):
    prev = 0
    for it in its:
        if it:
            o.step(it - prev)
            prev = it
        st = o.state()
        assert st["nag"] == g[f"s{it}_nag"], it
        assert st["dt"] == g[f"s{it}_dt"] and st["time"] == g[f"s{it}_time"], it
        for f in FIELDS:
            assert np.array_equal(st[f], g[f"s{it}_" + f]), (it, f)
        if lists:
            _check_lists(o, g[f"s{it}_nbr_count"], g[f"s{it}_nbr"])
            _check_btab(o, tb, g[f"s{it}_ncbf"], g[f"s{it}_bt_face"], g[f"s{it}_bt_loc"],
                        g[f"s{it}_bt_bi"], g[f"s{it}_bt_ig"])


def test_golden_nearwall_bit_identical():
    """Particles 0.02-0.22 dx from the floor and the x = 0 wall, moving into
    them: the elastic reaction of AddElasticBoundaryReaction_3D.f90:67-104
    and the zpmin clamp of ComputeBoundaryVolumeIntegrals_P0.f90:84,143 run
    (the oracle's counters say how often), and lists, boundary table and
    fields equal the reference's after 0, 4 and 8 steps."""
    g = np.load(GOLD / "ref_nearwall.npz")
    tb = build(_tiny())
    o = Oracle(tb, coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    _oracle_steps(o, g, (0, 4, 8), lists=True, tb=tb)
    n = o.branch_counts()
    assert n["elastic"] > 50 and n["zpmin"] > 50, n


def test_golden_nearwall_2d_bit_identical():
    """The 2D build with particles close to the floor and the x = 0 side,
    moving into them: AddElasticBoundaryReaction_2D.f90 runs."""
    from sphera_amd.case2d import DamBreak2D, build as build2d
    g = np.load(GOLD / "ref_nearwall_2d.npz")
    tb = build2d(DamBreak2D(tank=(0.6, 0.4), column=(0.3, 0.2), dx=0.01, name="t2d"))
    o = Oracle(tb, coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    _oracle_steps(o, g, (0, 4, 8))
    assert o.branch_counts()["elastic"] > 20, o.branch_counts()


def test_golden_density_clamp_bit_identical():
    """density_thresholds = 1 with a converging and a diverging half: the
    clamp of Leapfrog_continuity.f90:61-68 runs and the fields equal the
    reference's after 3 and 6 steps."""
    g = np.load(GOLD / "ref_dclamp.npz")
    case = _small()
    case.density_thresholds = 1
    tb = build(case)
    o = Oracle(tb, coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    _oracle_steps(o, g, (0, 3, 6))
    assert o.branch_counts()["density_clamp"] > 50, o.branch_counts()


REFERENCE = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "sphera3d.x"


@pytest.mark.skipif(not (HARNESS.exists() and REFERENCE.exists()), reason="reference not built here")
def test_harness_equals_unmodified_reference_program(tmp_path):
    """The dump harness (a project-written time_step_loop linked with the
    reference's objects) against the reference program as shipped: the
    results file that sphera3d.x writes every 5 steps (Memo_Results.f90:
    279-287, restartcode 0) holds the same positions, velocities, pressures,
    densities, times and dt, bit for bit, as the harness's dump."""
    from oracle.decks import read_dump, read_ris, run_harness, run_reference
    case = DamBreak3D(tank=(0.5, 0.2, 0.3), column=(0.15, 0.2, 0.15), dx=0.01, name="ris")
    wd = run_reference(case, itmax=20, results_step=5, workdir=tmp_path / "ref")
    ris = read_ris(wd / "ris.ris")
    assert [r["it"] for r in ris] == [5, 10, 15, 20]
    hdr, steps = read_dump(run_harness(case, nsteps=20, every=5, lists=False,
                                       workdir=tmp_path / "harness"))
    by_it = {s["it"]: s for s in steps}
    for r in ris:
        s = by_it[r["it"]]
        assert r["time"] == s["time"] and r["dt"] == s["dt"], r["it"]
        for f in ("coord", "vel", "pres", "dens", "mass"):
            assert np.array_equal(r[f], s[f]), (r["it"], f)


def test_golden_rk1_bit_identical():
    """The explicit RK1 branch (time_split = 0, RKscheme = 1: Euler.f90:61-241
    then EoS_fluid_loop, with the lists, boundary table and kernel values of
    the stage's start in the pressure smoothing) against the reference's
    fixture: fields, dt and time bit-identical after 6 and 12 steps."""
    g = np.load(GOLD / "ref_rk1.npz")
    case = _small()
    case.time_split = 0
    tb = build(case)
    o = Oracle(tb, coord=g["ic_coord"])
    o.preloop()
    _oracle_steps(o, g, (0, 6, 12))


@pytest.mark.skipif(not (HARNESS.exists() and REFERENCE.exists()), reason="reference not built here")
def test_rk1_harness_equals_unmodified_reference_program(tmp_path):
    """The harness's RK1 branch against the reference program's own loop.  The
    3D program reaches Euler(BC_zmax_flag) through time_integration's call
    without the argument (time_integration.f90:54) and faults where
    velocity_smoothing reads it, i.e. when TetaV > 0; with TetaV = 0 it runs,
    and its results file equals the harness's dump bit for bit."""
    from oracle.decks import read_dump, read_ris, run_harness, run_reference
    case = DamBreak3D(tank=(0.5, 0.2, 0.3), column=(0.15, 0.2, 0.15), dx=0.01, name="rk",
                      time_split=0, teta_v=0.0)
    wd = run_reference(case, itmax=20, results_step=5, workdir=tmp_path / "ref")
    ris = read_ris(wd / "rk.ris")
    assert [r["it"] for r in ris] == [5, 10, 15, 20]
    hdr, steps = read_dump(run_harness(case, nsteps=20, every=5, lists=False,
                                       workdir=tmp_path / "harness"))
    by_it = {s["it"]: s for s in steps}
    for r in ris:
        s = by_it[r["it"]]
        assert r["time"] == s["time"] and r["dt"] == s["dt"], r["it"]
        for f in ("coord", "vel", "pres", "dens", "mass"):
            assert np.array_equal(r[f], s[f]), (r["it"], f)


def test_golden_viscous_laminar_flags_bit_identical():
    """A viscous shear flow (20 Pa s, u = 3 z) in which the laminar flag of
    inter_EqMoto.f90:269-283 is set for hundreds of particles: fields, dt and
    time equal the reference's after 3 and 6 steps."""
    g = np.load(GOLD / "ref_viscous.npz")
    case = _tiny()
    case.mu = 20.0
    tb = build(case)
    o = Oracle(tb, vel=g["ic_vel"])
    o.preloop()
    _oracle_steps(o, g, (0, 3, 6))
    assert int(g["s6_laminar_flag"].sum()) > 100


def _wallfn_case(dim):
    if dim == 2:
        from sphera_amd.case2d import DamBreak2D
        return DamBreak2D(tank=(0.6, 0.4), column=(0.3, 0.2), dx=0.01, name="t2d",
                          slip_mode=2, roughness=(0.0005, 0.0015))
    return DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny",
                      slip_mode=2, roughness=(0.0005, 0.0015))


@pytest.mark.parametrize("dim", [3, 2])
def test_golden_wall_function_bit_identical(dim):
    """slip_coefficient_mode = 2 (roughness length 0.0005 m) with particles
    0.02-0.22 dx from two walls and a tangential velocity: the SA-SPH wall
    function (wall_function_for_SASPH.f90) runs on both its branches in the
    momentum equation (AddBoundaryContributions_to_ME3D.f90:207-227,
    ME2D.f90:314-326) and in the velocity smoothing, and the fields equal the
    reference's after 0, 4 and 8 steps.  The same input with the wall shear
    switched off (mode 0) ends elsewhere, so the comparison sees the term."""
    if dim == 2:
        from sphera_amd.case2d import build as build_d
    else:
        build_d = build
    g = np.load(GOLD / ("ref_wallfn_2d.npz" if dim == 2 else "ref_wallfn.npz"))
    case = _wallfn_case(dim)
    tb = build_d(case)
    fr = tb.sides if dim == 2 else tb.faces
    assert fr.slip_coeff[0] == 2.0 * 0.0005 / 5.0 + 0.0015 / 5.0 and fr.slip_mode[0] == 2
    o = Oracle(tb, coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    _oracle_steps(o, g, (0, 4, 8))
    n = o.branch_counts()
    assert n["wall_fn_near"] > 20 and n["wall_fn_log"] > 20, n
    case.slip_mode = 0
    o0 = Oracle(build_d(case), coord=g["ic_coord"], vel=g["ic_vel"])
    o0.preloop()
    o0.step(8)
    assert not np.array_equal(o0.state()["vel"], g["s8_vel"])
    assert o0.branch_counts()["wall_fn_near"] == 0


def test_golden_convex_edges_bit_identical():
    """A block on the floor adds two convex frontier edges
    (FindBoundaryConvexEdges3D, restated in sphera_amd.case.find_convex_edges);
    particles beyond the first edge moving into it get the edge reaction of
    AddElasticBoundaryReaction_3D.f90:106-152, those beyond the second edge
    none (the subroutine returns at the first edge that does not apply), and
    lists, fields, dt and time equal the reference's (one OpenMP thread) after
    0, 4 and 8 steps."""
    g = np.load(GOLD / "ref_edges.npz")
    case = _tiny()
    case.obstacle = (0.15, 0.2, 0.04)
    tb = build(case)
    e = tb.edges
    assert len(e) == 2
    assert [tuple(tb.faces.face_id[f] for f in pair) for pair in e.face] == [(12, 13), (13, 14)]
    assert np.array_equal(e.gx, [[0.15, 0.0, 0.04], [0.2, 0.0, 0.04]])
    assert np.array_equal(e.component, [[0.0, 0.1, 0.0]] * 2) and np.array_equal(e.length, [0.1, 0.1])
    assert len(build(_tiny()).edges) == 0  # the tank alone: every edge concave
    o = Oracle(tb, coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    _oracle_steps(o, g, (0, 4, 8))
    n = o.branch_counts()
    assert n["convex_edge"] >= 15 and n["elastic"] > 20, n


@pytest.mark.parametrize("ts,name", [(1, "ref_outlet.npz"), (0, "ref_outlet_rk1.npz")])
def test_golden_open_face_bit_identical(ts, name):
    """The tank's x = 0.3 m wall is an open face: the particles whose step
    crosses it are removed (CancelOutgoneParticles_3D marks them, OrdGrid1
    removes them by swap-with-last), the ten that cross its plane above the
    wall stay, and five leave the grid through its top during the run;
    count, particle order, dt and every field equal the reference's after 8
    and 16 steps, leapfrog and explicit RK1.  Without the open face the same
    input keeps all but those five."""
    g = np.load(GOLD / name)
    case = _tiny()
    case.outlet, case.time_split = True, ts
    o = Oracle(build(case), coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    _oracle_steps(o, g, (0, 8, 16))
    assert o.outgone_count() == 90 and int(g["s16_nag"]) == 705  # and 5 left the grid
    assert int((g["s16_coord"][:, 0] > 0.3).sum()) == 10
    case.outlet = False
    o0 = Oracle(build(case), coord=g["ic_coord"], vel=g["ic_vel"])
    o0.preloop()
    o0.step(16)
    assert o0.state()["nag"] == 795 and o0.outgone_count() == 0


@pytest.mark.parametrize("ts,name", [(1, "ref_outlet_2d.npz"), (0, "ref_outlet_2d_rk1.npz")])
def test_golden_open_side_2d_bit_identical(ts, name):
    """The 2D build with the tank's x = 0.6 m side open: the particles whose
    step crosses it are removed (CancelOutgoneParticles_2D), the four that
    cross its line above its end stay, three leave the grid through its top.
    The open side enters the boundary table of the 56 particles next to it
    (SelectCloseBoundarySides2D.f90:99-107), nine of them beside the floor,
    whose integrals then take the two-side form; it adds no wall term.
    Count, particle order, dt and every field equal the reference's after 8
    and 16 steps, leapfrog and explicit RK1.  With a fixed wall in its place
    the same input keeps all but the three."""
    from sphera_amd.case2d import DamBreak2D, build as build2d
    g = np.load(GOLD / name)
    case = DamBreak2D(tank=(0.6, 0.4), column=(0.3, 0.2), dx=0.01, name="t2d", outlet=True,
                      time_split=ts)
    tb = build2d(case)
    assert list(tb.sides.btype) == [0, 0, 6, 6, 6, 6, 2]
    o = Oracle(tb, coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    nc, face, _, _, _ = o.btab()
    on_open = ((face == 6) & (np.arange(face.shape[1]) < nc[:, None])).any(axis=1)
    assert on_open.sum() == 56 and (on_open & (nc == 2)).sum() == 9
    _oracle_steps(o, g, (0, 8, 16))
    assert o.outgone_count() == 16 and int(g["s16_nag"]) == 581  # and 3 left the grid
    assert int((g["s16_coord"][:, 0] > 0.6).sum()) == 4
    case.outlet = False
    o0 = Oracle(build2d(case), coord=g["ic_coord"], vel=g["ic_vel"])
    o0.preloop()
    o0.step(16)
    assert o0.state()["nag"] == 597 and o0.outgone_count() == 0


def _dbsph_case(name):
    from sphera_amd.dbsph import DBSPHParams
    from tests.golden.make_golden import DBSPH_OPTS
    return DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny", time_split=0,
                      dbsph=DBSPHParams(**DBSPH_OPTS[name]))


def test_dbsph_wall_elements_equal_reference():
    """The tank mesh's wall elements (sphera_amd.dbsph: the quadrilaterals
    split into triangles, centroids, normals, areas, edge adjacency through
    the grid cells, depth coefficients k_d, semi-particle volumes and masses)
    equal the reference's pg_w after its own pre-processing
    (Import_ply_surface_meshes, DBSPH_IC_surface_elements,
    DBSPH_find_close_faces, semi_particle_volumes) bit for bit.  Independent
    checks: the areas sum to the tank's wetted surface, the normals are unit
    vectors into the tank, k_d is 1 on flat interior triangles and above 1
    at the concave edges and corners."""
    g = np.load(GOLD / "ref_dbsph.npz")
    W = build(_dbsph_case("ref_dbsph.npz")).walls
    for k in ("coord", "normal", "weight", "volume", "mass", "k_d", "adjacent", "cell"):
        assert np.array_equal(getattr(W, k), g["wall_" + k]), k
    X, Y, Z = 0.3, 0.1, 0.2
    assert abs(W.weight.sum() - (X * Y + 2 * X * Z + 2 * Y * Z)) < 1e-12
    assert np.allclose(np.linalg.norm(W.normal, axis=1), 1.0, atol=1e-15)
    ctr = np.array([X / 2, Y / 2, Z])
    assert (np.einsum("ij,ij->i", W.normal, ctr - W.coord) > 0).all()
    flat = (W.coord[:, 2] > 0.02) & (W.coord[:, 2] < 0.18) & (W.coord[:, 1] > 0.0) & (W.coord[:, 1] < Y) \
        & (W.coord[:, 0] > 0.02) & (W.coord[:, 0] < 0.28)
    assert np.allclose(W.k_d[flat & ((W.adjacent > 0).sum(1) == 3)], 1.0, atol=1e-12)
    edge = (W.coord[:, 2] < 0.004) & ((W.coord[:, 1] < 0.004) | (W.coord[:, 1] > Y - 0.004))
    assert (W.k_d[edge] > 1.2).all()


@pytest.mark.parametrize("name", ["ref_dbsph.npz", "ref_dbsph_opts.npz", "ref_dbsph_muscl.npz",
                                  "ref_dbsph_moving.npz"])
def test_golden_dbsph_bit_identical(name):
    """DB-SPH walls (Domain%tipo = "bsph", RK1): the oracle's restatement of
    the fluid-wall lists and Shepard sums (CalcVarLength.f90:462-540,
    842-877), the MUSCL gradients and partial-Riemann wall state
    (Gradients_to_MUSCL.f90, BC_wall_elements.f90), the DB-SPH momentum
    terms (inter_EqMoto.f90:229-235,286-322, RHS_momentum_equation.f90:
    173-177) and Euler's DB-SPH step (Euler.f90:61-241) equals the
    reference bit for bit after the pre-loop and 10 and 20 steps: particle
    fields, dt and time, Gamma, sigma, rhoSPH_new, dShep, the gradients, the
    free-surface flags, the wall elements' density, pressure and wet flags,
    and the fluid-wall lists in the reference's order (fluid lists at the
    pre-loop).  The second fixture runs the Gamma limiter, negative wall
    pressures and no free-surface detection; the third the wall elements'
    terms in the MUSCL gradients (Gradients_to_MUSCL_boundary.f90), whose
    gradients and wall pressures differ from the first's at every step; the
    fourth a moving tank (DBSPH_kinematics.f90: an interpolated translation
    and rotation), whose fluid-wall lists differ from the first's."""
    g = np.load(GOLD / name)
    o = Oracle(build(_dbsph_case(name)), coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    _check_lists(o, g["s0_nbr_count"], g["s0_nbr"])
    prev = 0
    for it in (0, 10, 20):
        if it:
            o.step(it - prev)
            prev = it
        st, db = o.state(), o.dbsph_state()
        assert st["nag"] == g[f"s{it}_nag"] and st["dt"] == g[f"s{it}_dt"] and st["time"] == g[f"s{it}_time"]
        for f in FIELDS + ("var",):
            assert np.array_equal(st[f], g[f"s{it}_" + f]), (it, f)
        for f in ("gamma", "sigma", "rho_sph", "dshep", "drho", "dvel", "fs", "wall_dens", "wall_pres",
                  "wall_wet"):
            assert np.array_equal(db[f], g[f"s{it}_" + f]), (it, f)
        cw, idw = o.fw_lists()
        assert np.array_equal(cw, g[f"s{it}_fw_count"])
        assert np.array_equal(np.concatenate([idw[i, :cw[i]] for i in range(len(cw))]) + 1, g[f"s{it}_fw"])
    if name == "ref_dbsph.npz":
        assert g["s20_fs"].sum() > 0 and (g["s20_wall_pres"] > 0).sum() > 100
    elif name == "ref_dbsph_opts.npz":
        assert g["s20_gamma"].max() == 1.0 and (g["s20_wall_pres"] < 0).any()
    elif name == "ref_dbsph_moving.npz":  # the moving walls change the lists
        p = np.load(GOLD / "ref_dbsph.npz")
        for it in (10, 20):
            assert (g[f"s{it}_fw_count"] != p[f"s{it}_fw_count"]).sum() > 100, it
    else:  # the wall terms change the gradients and the wall state
        p = np.load(GOLD / "ref_dbsph.npz")
        for it in (10, 20):
            for f in ("drho", "dvel", "wall_pres"):
                assert not np.allclose(g[f"s{it}_" + f], p[f"s{it}_" + f], rtol=1e-3, atol=0), (it, f)


def test_dbsph_kinematics_starting_late_refused():
    """DBSPH_kinematics.f90 reads the record before the first one at or after
    the current time, so a first record after t = 0 is refused."""
    import dataclasses
    case = _dbsph_case("ref_dbsph.npz")
    case.dbsph = dataclasses.replace(case.dbsph, kinematics=((0.1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),))
    with pytest.raises(ValueError, match="first record"):
        Oracle(build(case))
