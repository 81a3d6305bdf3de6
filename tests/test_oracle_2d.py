"""2D form (the reference's -DSPACE_2D build): case set-up and the oracle
against the reference.

1. The set-up tables of sphera_amd/case2d.py (grid with one cell row in y,
   boundary sides with their local axes, mass-force matrices, relative angles,
   lattice IC) against the header the 2D reference harness wrote.
2. The oracle's 2D path against the committed fixture ref_2d.npz (written by
   tests/golden/make_golden.py from oracle/_ref/harness2d.x): neighbour lists,
   2D boundary table (ComputeBoundaryDataTab_2D) and every field, bit for bit.
3. Where oracle/_ref/harness2d.x is built (this container), the full 2D dam
   break (3.22 x 1.8 m tank, 1.0 x 0.5 m column, dx = 0.01 m, 5,000
   particles) against a live reference run.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from sphera_amd import case2d

GOLD = Path(__file__).resolve().parent / "golden"
HARNESS2D = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "harness2d.x"
FIELDS = ("coord", "vel", "var", "acc", "dens", "pres", "dden", "vpres")
CASE2D = dict(tank=(0.6, 0.4), column=(0.3, 0.2), dx=0.01, name="t2d")


def test_case2d_tables_match_reference_header():
    g = np.load(GOLD / "ref_2d.npz")
    tb = case2d.build(case2d.DamBreak2D(**CASE2D))
    s = tb.scalars
    assert int(g["hdr_nag"]) == len(tb.particles["coord"])
    assert int(g["hdr_nmaxpartj"]) == s["nmaxpartj"]
    assert np.array_equal(g["hdr_ncd"], tb.grid.ncd)
    assert np.array_equal(g["hdr_extr_lo"], tb.grid.extr_lo)
    assert np.array_equal(g["hdr_dcd"], tb.grid.dcd)
    for k in ("h", "coefke", "coefkacl", "eta", "eta2", "celerita", "kin_visc"):
        assert float(g["hdr_" + k]) == s[k], k
    sd = tb.sides
    assert np.array_equal(g["hdr_side_stretch"], sd.stretch)
    assert np.array_equal(g["hdr_side_length"], sd.length)
    assert np.array_equal(g["hdr_side_v1"], sd.v1)
    assert np.array_equal(g["hdr_side_v2"], sd.v2)
    acix = np.ix_([0, 2], [0, 2])
    wall = np.array([t != "peri" for t in g["hdr_side_tipo"]])
    assert np.array_equal(sd.btype != case2d.BT_FIXE, ~wall)
    for i in np.nonzero(wall)[0]:
        assert np.array_equal(g["hdr_side_T"][i][acix], sd.T[i][acix]), i
        assert np.array_equal(g["hdr_side_RN"][i][acix], sd.RN[i][acix]), i
        assert g["hdr_side_angle"][i] == sd.angle[i]
        assert g["hdr_side_previous"][i] == sd.previous[i]


def test_case2d_lattice_ic():
    """Column 1.0 x 0.5 m at dx = 0.01 m: 100 x 50 particles at half-dx
    offsets, y = 0, x outer; hydrostatic pressure p = rho0 g (H - z)."""
    case = case2d.DamBreak2D()
    P = case2d.initial_particles(case)
    c = P["coord"]
    assert c.shape == (5000, 3)
    assert np.all(c[:, 1] == 0.0)
    assert np.allclose(c[:50, 0], 0.005) and np.allclose(c[:50, 2], 0.005 + 0.01 * np.arange(50))
    assert np.allclose(P["pres"], 1000.0 * 9.81 * (0.5 - c[:, 2]), rtol=0, atol=1e-9)
    assert np.allclose(P["mass"], 0.1)


def _check_step(o, g, pre):
    st = o.state()
    for f in FIELDS:
        assert np.array_equal(st[f], g[pre + f]), (pre, f)
    assert st["dt"] == g[pre + "dt"]
    cnt, ids = o.lists()
    assert np.array_equal(cnt, g[pre + "nbr_count"])
    off = np.concatenate([[0], np.cumsum(cnt)])
    flat = g[pre + "nbr"]
    for i in range(len(cnt)):
        assert np.array_equal(ids[i, :cnt[i]] + 1, flat[off[i]:off[i + 1]]), i
    nc, side, loc, bi, _ = o.btab()
    assert np.array_equal(nc, g[pre + "intncbs"])
    k = 0
    for i in range(len(nc)):
        for q in range(nc[i]):
            assert side[i, q] + 1 == g[pre + "bt_side"][k]
            assert np.array_equal(loc[i, q], g[pre + "bt_loc"][k])
            assert np.array_equal(bi[i, q], g[pre + "bt_bi"][k])
            k += 1
    return int(nc.sum())


def test_oracle2d_golden_bit_identical():
    g = np.load(GOLD / "ref_2d.npz")
    tb = case2d.build(case2d.DamBreak2D(**CASE2D))
    o = Oracle(tb, coord=g["ic_coord"])
    o.preloop()
    nrec0 = _check_step(o, g, "s0_")
    o.step(10)
    nrec10 = _check_step(o, g, "s10_")
    # the fixture exercises walls and the corner (two-side) integrals
    assert nrec0 > 50 and nrec10 > 50
    assert np.any(g["s10_intncbs"] == 2) and np.any(g["s10_intncbs"] == 1)


@pytest.mark.skipif(not HARNESS2D.exists(), reason="2D reference harness not built here")
def test_oracle2d_full_dam_break_vs_live_reference():
    from oracle.decks import read_dump, run_harness
    case = case2d.DamBreak2D(name="db2d")
    tb = case2d.build(case)
    hdr, steps = read_dump(run_harness(case, nsteps=60, every=20, lists=False))
    o = Oracle(tb)
    o.preloop()
    done = 0
    for st in steps:
        o.step(st["it"] - done)
        done = st["it"]
        mine = o.state()
        for f in FIELDS:
            assert np.array_equal(mine[f], st[f]), (st["it"], f)
        assert mine["dt"] == st["dt"]
    assert done == 60


REFERENCE2D = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "sphera2d.x"


@pytest.mark.skipif(not (HARNESS2D.exists() and REFERENCE2D.exists()), reason="2D reference not built here")
@pytest.mark.parametrize("time_split", [1, 0])
def test_2d_harness_and_oracle_equal_unmodified_reference_program(time_split, tmp_path):
    """The reference's 2D program as shipped (its own time_step_loop, leapfrog
    or explicit RK1 with TetaV = 0.2; in 2D its time_integration calls Euler
    with the matching interface) writes results every 5 steps: the harness's
    dump and the oracle equal them bit for bit (positions, velocities,
    pressures, densities, time, dt)."""
    from oracle.decks import read_dump, read_ris, run_harness, run_reference
    case = case2d.DamBreak2D(**CASE2D, time_split=time_split)
    wd = run_reference(case, itmax=20, results_step=5, workdir=tmp_path / "ref")
    ris = read_ris(wd / "t2d.ris")
    assert [r["it"] for r in ris] == [5, 10, 15, 20]
    hdr, steps = read_dump(run_harness(case, nsteps=20, every=5, lists=False,
                                       workdir=tmp_path / "harness"))
    by_it = {s["it"]: s for s in steps}
    o = Oracle(case2d.build(case))
    o.preloop()
    prev = 0
    for r in ris:
        s = by_it[r["it"]]
        o.step(r["it"] - prev)
        prev = r["it"]
        st = o.state()
        assert r["time"] == s["time"] == st["time"] and r["dt"] == s["dt"] == st["dt"], r["it"]
        for f in ("coord", "vel", "pres", "dens"):
            assert np.array_equal(r[f], s[f]) and np.array_equal(r[f], st[f]), (r["it"], f)
