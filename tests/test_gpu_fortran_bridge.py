"""The reference program itself driving the device through the Fortran
bridge (sphera_amd/fortran/): oracle/_ref/harness3d_hip.x is the flang build
of the reference with time_step_loop replaced by a unit that calls the C ABI
via ISO_C_BINDING.  Its FP64 output must equal the reference's own output
(the golden fixtures written from the flang build) bit for bit; its
FP32-mixed output must stay within the per-field tolerances of
test_gpu_parity.py.  Skipped where the harness was not built (it needs the
reference sources at build time)."""
from pathlib import Path

import numpy as np
import pytest

from oracle.decks import HARNESS2D_HIP, HARNESS_HIP, read_dump, run_harness
from sphera_amd.case import DamBreak3D
from tests.test_gpu_parity import FP32_TOL

GOLD = Path(__file__).resolve().parent / "golden"

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not HARNESS_HIP.exists(), reason="drop-in harness not built")]

FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")


def _tiny():
    return DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny")


def _small():
    return DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01, name="small")


def _run(case, nsteps, every, tmp_path, precision=0, phases=0, ic=None):
    print(f"[bridge] {case.name} steps={nsteps} precision={precision} phases={phases}", flush=True)
    hdr, steps = read_dump(run_harness(case, nsteps=nsteps, every=every, lists=False,
                                       ic_coord=ic, workdir=tmp_path, exe=HARNESS_HIP,
                                       timeout=120,
                                       extra_env={"SPH_HIP_PRECISION": str(precision),
                                                  "SPH_HIP_PHASES": str(phases)}))
    return hdr, {r["it"]: r for r in steps}


@pytest.mark.parametrize("phases", [0, 1])
def test_bridge_fp64_equals_reference_tiny(phases, tmp_path):
    g = np.load(GOLD / "ref_tiny.npz")
    _, st = _run(_tiny(), 5, 5, tmp_path, phases=phases)
    assert st[0]["dt"] == g["s0_dt"]
    for f in FIELDS:
        assert np.array_equal(st[0][f], g["s0_" + f]), ("s0", f)
        assert np.array_equal(st[5][f], g["s5_" + f]), ("s5", f)
    assert st[5]["dt"] == g["s5_dt"] and st[5]["time"] == g["s5_time"]


def test_bridge_fp64_equals_reference_small_jitter(tmp_path):
    g = np.load(GOLD / "ref_small_jitter.npz")
    _, st = _run(_small(), 10, 10, tmp_path, ic=g["ic_coord"])
    for f in FIELDS:
        assert np.array_equal(st[10][f], g["s10_" + f]), f
    assert st[10]["dt"] == g["s10_dt"]


def test_bridge_fp32_mixed_within_tolerance(tmp_path):
    g = np.load(GOLD / "ref_small_jitter.npz")
    _, st = _run(_small(), 10, 10, tmp_path, precision=1, ic=g["ic_coord"])
    for f in FIELDS:
        ref = g["s10_" + f]
        err = np.abs(st[10][f] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err < FP32_TOL[f], (f, err)


def test_bridge_free_surface_equals_reference(tmp_path):
    """calc_pelo through the bridge (sph_hip_monitor_lines, sph_hip_calc_pelo):
    the device run writes .plb/.zlft records that read back as the reference's
    own (tests/golden/ref_pelo.npz, same deck and initial positions)."""
    from oracle.decks import read_pelo
    from tests.golden.make_golden import PELO_LINES, TINY
    g = np.load(GOLD / "ref_pelo.npz")
    case = DamBreak3D(**TINY, control_lines=PELO_LINES)
    print("[bridge] free surface tiny steps=6", flush=True)
    dump = run_harness(case, nsteps=6, every=3, lists=False, ic_coord=g["ic_coord"],
                       workdir=tmp_path, exe=HARNESS_HIP, timeout=120,
                       extra_env={"SPH_HIP_PRECISION": "0", "SPH_HARNESS_PELO": "1"})
    t, pel, zl = read_pelo(dump.parent, case)
    assert np.allclose(t, g["time"], rtol=1e-6, atol=1e-12)
    assert np.allclose(pel, g["pelolib"], rtol=1e-6, atol=1e-12)
    assert np.allclose(zl, g["zlft"], rtol=1e-6, atol=1e-12)


@pytest.mark.skipif(not HARNESS2D_HIP.exists(), reason="2D drop-in harness not built")
def test_bridge_2d_fp64_matches_reference(tmp_path):
    """The reference's 2D build (-DSPACE_2D) driving the device through the
    same bridge (sph_ctx_create2d with its BoundarySide tables): fields after
    10 steps against the 2D reference output (tests/golden/ref_2d.npz) within
    the 2D FP64 tolerance of test_gpu_2d.py (the device's trigonometric
    functions in the corner integrals)."""
    from sphera_amd import case2d
    from tests.golden.make_golden import CASE2D
    from tests.test_gpu_2d import FP64_REL_TOL
    g = np.load(GOLD / "ref_2d.npz")
    case = case2d.DamBreak2D(**CASE2D)
    print("[bridge] 2d steps=10", flush=True)
    hdr, steps = read_dump(run_harness(case, nsteps=10, every=10, lists=False,
                                       ic_coord=g["ic_coord"], workdir=tmp_path,
                                       exe=HARNESS2D_HIP, timeout=120,
                                       extra_env={"SPH_HIP_PRECISION": "0"}))
    st = {r["it"]: r for r in steps}
    for f in FIELDS + ("var",):
        ref = g["s10_" + f]
        err = np.abs(st[10][f] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err <= FP64_REL_TOL, (f, err)
    assert abs(st[10]["dt"] - float(g["s10_dt"])) <= 1e-12 * float(g["s10_dt"])


def test_bridge_convex_edges_wall_function_open_face(tmp_path):
    """Decks with a block on the floor (two convex edges, passed as the
    reference's BoundaryConvexEdge table), with slip_coefficient_mode = 2 and
    with an open face (the particles leaving through it are removed, in the
    reference's order): the device run through the bridge keeps the
    reference's particle count and stays within 1e-12 of its fixtures (the
    wall reactions' and the wall function's log() are the device library's).
    One OpenMP thread: the reference orders the convex edges by thread
    arrival."""
    from tests.golden.make_golden import OBSTACLE, TINY, WALLFN
    for name, case, its in (("ref_edges.npz", DamBreak3D(**TINY, obstacle=OBSTACLE), (4, 8)),
                            ("ref_wallfn.npz", DamBreak3D(**TINY, **WALLFN), (4, 8)),
                            ("ref_outlet.npz", DamBreak3D(**TINY, outlet=True), (8, 16))):
        g = np.load(GOLD / name)
        print(f"[bridge] {name} steps={its[1]}", flush=True)
        hdr, steps = read_dump(run_harness(case, nsteps=its[1], every=its[0], lists=False,
                                           ic_coord=g["ic_coord"], ic_vel=g["ic_vel"],
                                           workdir=tmp_path / name[:-4], exe=HARNESS_HIP, timeout=120,
                                           threads=1,
                                           extra_env={"SPH_HIP_PRECISION": "0", "SPH_HIP_PHASES": "0"}))
        st = {r["it"]: r for r in steps}
        for it in its:
            assert st[it]["nag"] == g[f"s{it}_nag"], (name, it)
            for f in FIELDS:
                ref = g[f"s{it}_" + f]
                err = np.abs(st[it][f] - ref).max() / max(np.abs(ref).max(), 1e-300)
                assert err <= 1e-12, (name, it, f, err)


@pytest.mark.skipif(not HARNESS2D_HIP.exists(), reason="2D drop-in harness not built")
def test_bridge_2d_open_side(tmp_path):
    """The reference's 2D build with an open side (tests/golden/
    ref_outlet_2d.npz) driving the device through the bridge: the bridge
    passes the open side with its second vertex, the device removes the
    particles that cross it, and count and fields after 8 and 16 steps match
    the reference's within the 2D FP64 tolerance of test_gpu_2d.py."""
    from sphera_amd import case2d
    from tests.golden.make_golden import CASE2D
    from tests.test_gpu_2d import FP64_REL_TOL
    g = np.load(GOLD / "ref_outlet_2d.npz")
    case = case2d.DamBreak2D(**CASE2D, outlet=True)
    print("[bridge] 2d open side steps=16", flush=True)
    hdr, steps = read_dump(run_harness(case, nsteps=16, every=8, lists=False,
                                       ic_coord=g["ic_coord"], ic_vel=g["ic_vel"],
                                       workdir=tmp_path, exe=HARNESS2D_HIP, timeout=120,
                                       extra_env={"SPH_HIP_PRECISION": "0"}))
    st = {r["it"]: r for r in steps}
    for it in (8, 16):
        assert st[it]["nag"] == g[f"s{it}_nag"], it
        for f in FIELDS + ("var",):
            ref = g[f"s{it}_" + f]
            err = np.abs(st[it][f] - ref).max() / max(np.abs(ref).max(), 1e-300)
            assert err <= FP64_REL_TOL, (it, f, err)


@pytest.mark.parametrize("name", ["ref_dbsph.npz", "ref_dbsph_muscl.npz", "ref_dbsph_moving.npz"])
def test_bridge_dbsph(tmp_path, name):
    """The reference program with DB-SPH walls (Domain%tipo = "bsph", RK1)
    driving the device through the bridge: the bridge passes pg_w, as the
    reference's own pre-processing built it from the tank's .ply mesh, to
    sph_dbsph_walls; count and fields after 10 and 20 steps equal the
    reference's own run (tests/golden/ref_dbsph*.npz) bit for bit: static
    walls, with the MUSCL wall terms, and a moving tank (kinematics records)."""
    from sphera_amd.dbsph import DBSPHParams
    from tests.golden.make_golden import DBSPH_OPTS, TINY
    g = np.load(GOLD / name)
    case = DamBreak3D(**TINY, time_split=0, dbsph=DBSPHParams(**DBSPH_OPTS[name]))
    print(f"[bridge] dbsph {name} steps=20", flush=True)
    hdr, steps = read_dump(run_harness(case, nsteps=20, every=10, lists=False,
                                       ic_coord=g["ic_coord"], ic_vel=g["ic_vel"],
                                       workdir=tmp_path, exe=HARNESS_HIP, timeout=120,
                                       extra_env={"SPH_HIP_PRECISION": "0"}))
    st = {r["it"]: r for r in steps}
    for it in (10, 20):
        assert st[it]["nag"] == g[f"s{it}_nag"] and st[it]["dt"] == g[f"s{it}_dt"], it
        for f in FIELDS:
            assert np.array_equal(st[it][f], g[f"s{it}_" + f]), (it, f)
