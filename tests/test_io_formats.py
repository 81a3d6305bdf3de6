"""Host I/O formats of the reference: the v10 deck (writer and parser subset,
sphera_amd/deck.py) and the results file (sphera_amd/results.py).

Without the reference: deck round trips give the same set-up tables, and a
results file round-trips.  Where the flang build of the reference is in
oracle/_ref (this container): the reference's own v10 3D tutorial deck,
parsed here, gives the reference's initial state and a run equal to the
reference's run on that deck; and the results records written here equal,
byte for byte, the ones the unmodified reference program writes."""
import glob
import io
import re
from pathlib import Path

import numpy as np
import pytest

from sphera_amd import deck, results
from sphera_amd.case import DamBreak3D, build

ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"
TUTORIAL = Path("/root/reference/input/tutorials/release_v_10_0_0/42_dam_break_3D_demo/input")


@pytest.mark.parametrize("kw", [dict(), dict(density_thresholds=1), dict(time_split=0),
                                dict(slip_mode=2, roughness=(0.0005, 0.0015)), dict(slip_mode=0),
                                dict(obstacle=(0.15, 0.2, 0.04)), dict(outlet=True),
                                dict(tank=(0.5, 0.3, 0.3), column=(0.2, 0.3, 0.15), dx=0.012)])
def test_deck_round_trip_builds_the_same_tables(kw):
    args = dict(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="rt",
                control_lines=(("L1", (0.1, 0.05, 0.0), (0.1, 0.05, 0.2), 11),))
    args.update(kw)
    case = DamBreak3D(**args)
    d = deck.parse_deck(deck.deck_text(case))
    a, b = build(case), build(d)
    assert a.scalars == b.scalars
    for k in ("coord", "vel", "dens", "pres", "mass"):
        assert np.array_equal(a.particles[k], b.particles[k]), k
    for k in ("extr_lo", "ncd", "dcd"):
        assert np.array_equal(getattr(a.grid, k), getattr(b.grid, k)), k
    for k in ("T", "LX", "GX", "area", "btype", "slip_coeff", "slip_mode"):
        assert np.array_equal(getattr(a.faces, k), getattr(b.faces, k)), k
    assert np.array_equal(a.gcbf_vector, b.gcbf_vector) and np.array_equal(a.gcbf_count, b.gcbf_count)
    for k in ("face", "gx", "component", "length"):
        assert np.array_equal(getattr(a.edges, k), getattr(b.edges, k)), k
    for k in ("rho0", "bulk", "alfa_mon", "mu", "grav", "p_ref", "cfl", "vsc_coeff", "teta_p", "teta_v",
              "density_thresholds", "time_split", "control_lines"):
        assert getattr(case, k) == getattr(d, k), k


def test_parser_refuses_what_it_does_not_read():
    text = deck.deck_text(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01))
    with pytest.raises(ValueError, match="semi"):
        deck.parse_deck(text.replace("\nsemi\n", "\nfull\n"))
    with pytest.raises(ValueError, match="DBSPH"):  # a bsph deck without a surface mesh
        deck.parse_deck(text.replace("\nsemi\n", "\nbsph\n"))
    with pytest.raises(ValueError, match="10.0.0"):
        deck.parse_deck(text.replace("10.0.0", "9.0.0", 1))
    bad = text.replace(".false. .false. .false. .false. .false. 0.614",
                       ".true. .false. .false. .false. .false. 0.614")
    with pytest.raises(ValueError, match="ALE3"):
        deck.parse_deck(bad)


def test_results_round_trip_and_colour():
    rng = np.random.default_rng(0)
    n = 37
    st = dict(coord=rng.random((n, 3)), vel=rng.random((n, 3)), pres=rng.random(n),
              dens=rng.random(n), mass=rng.random(n))
    fh = io.BytesIO()
    results.write_results(fh, 5, 0.25, 1e-4, st, kin_visc=1e-6)
    results.write_results(fh, 10, 0.5, 2e-4, st, kin_visc=1e-6)
    p = Path(__file__).resolve().parent / "_tmp_results.ris"
    try:
        p.write_bytes(fh.getvalue())
        back = results.read_results(p)
    finally:
        p.unlink()
    assert [r["it"] for r in back] == [5, 10] and back[1]["time"] == 0.5
    for k in st:
        assert np.array_equal(back[0][k], st[k]), k
    assert results.zone_colour("009EA8") == 0xA89E00 and results.zone_colour("FF0000") == 0xFF


def _tutorial_text():
    files = sorted(glob.glob(str(TUTORIAL / "*" / "*.inp")))
    return "".join(Path(f).read_text() for f in files)


def _flang_integer_zeros(text):
    """flang's list-directed read rejects "0." for the integer that opens the
    section-flow-rate and substations data lines; the run gets "0"."""
    out, pending = [], False
    for ln in text.splitlines(keepends=True):
        s = ln.strip()
        if s.lower() in ("##### section flow rate #####", "##### substations #####"):
            pending = True
        elif pending and s and not s.startswith("!"):
            ln = re.sub(r"^(\s*)0\.(\s)", r"\g<1>0\g<2>", ln)
            pending = False
        out.append(ln)
    return "".join(out)


@pytest.mark.skipif(not ((REF / "harness3d.x").exists() and TUTORIAL.is_dir()),
                    reason="reference not built here")
def test_reference_tutorial_deck_parsed_and_run(tmp_path):
    """The reference's v10 3D dam-break tutorial deck (its input/*/*.inp
    concatenated as sphera.f90:222 does): the parsed case has the reference's
    1800 particles, grid and hydrostatic state, and the oracle run from it
    equals the reference's run on the deck, bit for bit, over 10 steps."""
    from oracle.decks import read_dump, run_harness
    from oracle.pyoracle import Oracle
    text = _tutorial_text()
    d = deck.parse_deck(text)
    tb = build(d)
    hdr, steps = read_dump(run_harness(d, nsteps=10, every=5, lists=False, workdir=tmp_path,
                                       text=_flang_integer_zeros(text)))
    assert hdr["nag"] == len(tb.particles["coord"]) == 1800
    assert np.array_equal(hdr["ncd"], tb.grid.ncd) and np.array_equal(hdr["dcd"], tb.grid.dcd)
    assert hdr["nmaxpartj"] == tb.scalars["nmaxpartj"]
    for k in ("coord", "dens", "pres", "mass"):
        assert np.array_equal(steps[0][k], tb.particles[k]), k
    o = Oracle(tb)
    o.preloop()
    prev = 0
    for st in steps:
        if st["it"]:
            o.step(st["it"] - prev)
            prev = st["it"]
        s = o.state()
        assert s["dt"] == st["dt"]
        for k in ("coord", "vel", "acc", "dens", "pres", "dden", "vpres"):
            assert np.array_equal(s[k], st[k]), (st["it"], k)


@pytest.mark.skipif(not ((REF / "harness3d.x").exists() and (REF / "sphera3d.x").exists()),
                    reason="reference not built here")
def test_results_records_equal_the_reference_program(tmp_path):
    """The unmodified reference program writes results every 4 steps; the
    records written here from the harness's state at the same steps are the
    same bytes."""
    from oracle.decks import read_dump, run_harness, run_reference
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="io")
    wd = run_reference(case, itmax=8, results_step=4, workdir=tmp_path / "ref")
    recs = results.records(wd / "io.ris")
    heads = [k for k, r in enumerate(recs) if len(r) == 36 and k + 1 < len(recs)
             and len(recs[k + 1]) == 88 * 800]
    assert len(heads) == 2
    hdr, steps = read_dump(run_harness(case, nsteps=8, every=4, lists=False, workdir=tmp_path / "h"))
    by_it = {s["it"]: s for s in steps}
    kin = case.scalars()["kin_visc"]
    for k in heads:
        it = int(np.frombuffer(recs[k][:4], "<i4")[0])
        s = by_it[it]
        fh = io.BytesIO()
        results.write_results(fh, it, s["time"], s["dt"], s, kin_visc=kin)
        mine = results.records_from_bytes(fh.getvalue())
        assert mine[0] == recs[k] and mine[1] == recs[k + 1], it


@pytest.mark.skipif(not ((REF / "harness3d.x").exists() and (REF / "sphera3d.x").exists()),
                    reason="reference not built here")
def test_vtu_equals_the_reference_program(tmp_path):
    """The unmodified reference program writes the particle file at the start
    (block 0) and after its last step (block 1; its loop runs itmax + 1
    steps): the files written here from the harness's state at those steps
    are the same text."""
    from oracle.decks import read_dump, run_harness, run_reference
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="vt")
    wd = run_reference(case, itmax=4, results_step=2, workdir=tmp_path / "ref")
    hdr, steps = read_dump(run_harness(case, nsteps=5, every=5, lists=False, workdir=tmp_path / "h"))
    assert [s["it"] for s in steps] == [0, 5]
    for blk, st in enumerate(steps):
        fh = io.StringIO()
        results.write_vtu(fh, st, case.scalars()["kin_visc"], st["laminar_flag"])
        assert fh.getvalue() == (wd / f"VTKConverter_vt_block_{blk}.vtu").read_text(), blk


def test_dbsph_case_directory_round_trip(tmp_path):
    """A DB-SPH case written as the reference reads it (the bsph deck, the
    tank's .ply mesh and surface_mesh_list.txt) and read back by
    deck.read_case_dir builds the same particles, grid and wall elements,
    with the DBSPH section's flags."""
    from oracle.decks import write_case_dir
    from sphera_amd.dbsph import DBSPHParams, read_ply, tank_mesh
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="rtdb", time_split=0,
                      dbsph=DBSPHParams(gamma_limiter=True, fs_allowed=False, negative_wall_p_allowed=True,
                                        muscl_boundary=True,
                                        kinematics=((0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
                                                    (0.5, 0.25, 0.0, -0.125, 0.0, 0.0, 2.0)),
                                        rotation_centre=(0.15, 0.05, 0.0)))
    write_case_dir(case, tmp_path)
    v, f = read_ply(tmp_path / "input" / "07_DBSPH" / "tank.ply")
    v0, f0 = tank_mesh(case.tank, case.dx)
    assert np.array_equal(v, v0) and np.array_equal(np.array(f), f0)
    d = deck.read_case_dir(tmp_path)
    assert d.dbsph == case.dbsph and d.time_split == 0
    a, b = build(case), build(d)
    for k in ("coord", "dens", "pres", "mass"):
        assert np.array_equal(a.particles[k], b.particles[k]), k
    for k in ("coord", "normal", "weight", "volume", "mass", "k_d", "adjacent", "cell"):
        assert np.array_equal(getattr(a.walls, k), getattr(b.walls, k)), k
