"""2D form on the device (sph_ctx_create2d) vs the FP64 oracle, whose 2D path
is pinned bit for bit to the reference's -DSPACE_2D build
(tests/test_oracle_2d.py).

Bars: neighbour lists identical as ordered lists; boundary-table sides and
local coordinates identical; boundary integrals and fields within
FP64_REL_TOL (the corner integrals of BoundaryVolumeIntegrals2D and
ComputeVolumeIntegral_WdV2D call sin, cos, tan and atan2, which the device
evaluates with its own math library); FP32-mixed within FP32_TOL of the FP64
oracle.
"""
import numpy as np
import pytest

from sphera_amd import case2d

pytestmark = pytest.mark.gpu

FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")
FP64_REL_TOL = 1e-11
# largest per-particle |FP32-mixed - FP64| relative to the field's largest |value|
FP32_TOL = {f: 1e-4 for f in FIELDS}
SMALL = dict(tank=(0.6, 0.4), column=(0.3, 0.2), dx=0.01, name="t2d")


def _jitter2d(coord, dx, seed=20260303, amp=0.01):
    rng = np.random.default_rng(seed)
    out = coord.copy()
    j = rng.uniform(-amp, amp, (len(coord), 2)) * dx
    out[:, 0] += j[:, 0]
    out[:, 2] += j[:, 1]
    return out


def _pair(tb, precision, coord=None):
    from oracle.pyoracle import Oracle
    from sphera_amd import runtime as rt
    o = Oracle(tb, coord=coord)
    o.preloop()
    d = rt.SphContext(tb, precision=precision)
    d.upload(coord=coord)
    d.preloop()
    return d, o


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _lists_equal(d, o):
    dc, di = d.lists()
    oc, oi = o.lists()
    assert np.array_equal(dc, oc), "neighbour counts differ"
    for i in range(len(dc)):
        assert np.array_equal(di[i, :dc[i]], oi[i, :oc[i]]), f"list of particle {i} differs"


def _btab_close(d, o):
    dn, dside, dloc, dbi, _ = d.btab()
    on, oside, oloc, obi, _ = o.btab()
    assert np.array_equal(dn, on), "selected side counts differ"
    # the two tables may be read with different slot strides; compare the
    # slots both hold (the selected count is at most 2 in 2D)
    w = min(dside.shape[1], oside.shape[1])
    assert int(dn.max(initial=0)) <= w
    dside, dloc, dbi = dside[:, :w], dloc[:, :w], dbi[:, :w]
    oside, oloc, obi = oside[:, :w], oloc[:, :w], obi[:, :w]
    m = np.arange(w)[None, :] < dn[:, None]
    assert np.array_equal(dside[m], oside[m])
    assert np.array_equal(dloc[m], oloc[m])
    err = np.abs(dbi[m] - obi[m]).max(initial=0.0)
    scale = np.abs(obi[m]).max(initial=1.0)
    assert err <= 1e-13 * scale, f"boundary integrals differ by {err:.3e}"
    return int(dn.sum())


def test_fp64_2d_lists_boundary_table_and_steps_match_oracle():
    from sphera_amd import runtime as rt
    tb = case2d.build(case2d.DamBreak2D(**SMALL))
    coord = _jitter2d(tb.particles["coord"], tb.case.dx)
    print(f"[gpu] 2d n={len(coord)} fp64", flush=True)
    d, o = _pair(tb, rt.FP64, coord)
    _lists_equal(d, o)
    assert _btab_close(d, o) > 50
    assert d.state()["dt"] == o.state()["dt"]
    for k in (1, 4, 10):
        d.step(k)
        o.step(k)
        ds, os_ = d.state(), o.state()
        assert np.array_equal(ds["ids"], np.arange(len(ds["ids"])))
        for f in FIELDS:
            e = _rel(ds[f], os_[f])
            assert e <= FP64_REL_TOL, f"{f} rel err {e:.3e} after {k} more steps"
        _lists_equal(d, o)
        _btab_close(d, o)
    d.close()


def test_fp32_mixed_2d_within_tolerance():
    from sphera_amd import runtime as rt
    tb = case2d.build(case2d.DamBreak2D(**SMALL))
    coord = _jitter2d(tb.particles["coord"], tb.case.dx)
    print(f"[gpu] 2d n={len(coord)} fp32-mixed", flush=True)
    d, o = _pair(tb, rt.FP32_MIXED, coord)
    _lists_equal(d, o)
    d.step(20)
    o.step(20)
    ds, os_ = d.state(), o.state()
    errs = {f: _rel(ds[f], os_[f]) for f in FIELDS}
    print("[gpu] 2d fp32-mixed rel errors after 20 steps:", {k: f"{v:.2e}" for k, v in errs.items()},
          flush=True)
    for f in FIELDS:
        assert errs[f] <= FP32_TOL[f], f"{f} rel err {errs[f]:.3e}"
    assert np.all(ds["coord"][:, 1] == 0.0) and np.all(ds["vel"][:, 1] == 0.0)
    d.close()


def test_full_2d_dam_break_fp64_and_graph():
    """3.22 x 1.8 m tank, 1.0 x 0.5 m column, dx = 0.01 m (5,000 particles):
    200 steps eagerly and as a replayed step graph against the oracle."""
    from sphera_amd import runtime as rt
    tb = case2d.build(case2d.DamBreak2D())
    print(f"[gpu] 2d full n={len(tb.particles['coord'])} fp64 200 steps", flush=True)
    d, o = _pair(tb, rt.FP64)
    g = rt.SphContext(tb, precision=rt.FP64)
    g.upload()
    g.preloop()
    d.step(200)
    o.step(200)
    for _ in range(2):
        g.step_graph(100)
    g.sync()
    ds, gs, os_ = d.state(), g.state(), o.state()
    for f in FIELDS:
        e = _rel(ds[f], os_[f])
        assert e <= 1e-9, f"{f} rel err {e:.3e} after 200 steps"
        assert np.array_equal(ds[f], gs[f]), f"graph replay differs in {f}"
    # the front has left the column and nothing crossed a wall
    assert ds["coord"][:, 0].max() > 0.995
    assert ds["coord"][:, 0].min() > 0.0 and ds["coord"][:, 2].min() > 0.0
    d.close()
    g.close()


def test_nearwall_2d_fixture():
    """The reference's 2D near-wall fixture (tests/golden/ref_nearwall_2d.npz:
    particles 0.02-0.22 dx from the floor and the x = 0 side, moving into
    them): AddElasticBoundaryReaction_2D runs.  Device FP64 within
    FP64_REL_TOL of the reference after 4 and 8 steps, FP32-mixed within
    FP32_TOL."""
    from pathlib import Path
    from sphera_amd import runtime as rt
    g = np.load(Path(__file__).resolve().parent / "golden" / "ref_nearwall_2d.npz")
    tb = case2d.build(case2d.DamBreak2D(**SMALL))
    print(f"[gpu] 2d nearwall n={len(g['ic_coord'])}", flush=True)
    ctx = {}
    for prec in (rt.FP64, rt.FP32_MIXED):
        d = rt.SphContext(tb, precision=prec)
        d.upload(coord=g["ic_coord"], vel=g["ic_vel"])
        d.preloop()
        ctx[prec] = d
    for it in (4, 8):
        for prec, d in ctx.items():
            d.step(4)
            ds = d.state()
            errs = {f: _rel(ds[f], g[f"s{it}_" + f]) for f in FIELDS}
            exact = sum(bool(np.array_equal(ds[f], g[f"s{it}_" + f])) for f in FIELDS)
            print(f"[gpu] 2d nearwall it={it} precision={prec} bit-identical fields {exact}/7 "
                  f"max rel {max(errs.values()):.2e}", flush=True)
            tol = FP64_REL_TOL if prec == rt.FP64 else FP32_TOL["vel"]
            for f in FIELDS:
                assert errs[f] <= tol, (it, prec, f, errs[f])
    for d in ctx.values():
        d.close()


def test_wall_function_2d_fixture():
    """slip_coefficient_mode = 2 in the 2D build (tests/golden/
    ref_wallfn_2d.npz): the wall function of AddBoundaryContributions_to_ME2D.
    f90:314-326 and velocity_smoothing_SA_SPH_2D.f90:93-106.  Device FP64
    within FP64_REL_TOL of the reference after 4 and 8 steps, FP32-mixed
    within FP32_TOL."""
    from pathlib import Path
    from sphera_amd import runtime as rt
    g = np.load(Path(__file__).resolve().parent / "golden" / "ref_wallfn_2d.npz")
    tb = case2d.build(case2d.DamBreak2D(**SMALL, slip_mode=2, roughness=(0.0005, 0.0015)))
    print(f"[gpu] 2d wall function n={len(g['ic_coord'])}", flush=True)
    ctx = {}
    for prec in (rt.FP64, rt.FP32_MIXED):
        d = rt.SphContext(tb, precision=prec)
        d.upload(coord=g["ic_coord"], vel=g["ic_vel"])
        d.preloop()
        ctx[prec] = d
    for it in (4, 8):
        for prec, d in ctx.items():
            d.step(4)
            ds = d.state()
            errs = {f: _rel(ds[f], g[f"s{it}_" + f]) for f in FIELDS}
            exact = sum(bool(np.array_equal(ds[f], g[f"s{it}_" + f])) for f in FIELDS)
            print(f"[gpu] 2d wall function it={it} precision={prec} bit-identical fields {exact}/7 "
                  f"max rel {max(errs.values()):.2e}", flush=True)
            tol = FP64_REL_TOL if prec == rt.FP64 else FP32_TOL["vel"]
            for f in FIELDS:
                assert errs[f] <= tol, (it, prec, f, errs[f])
    for d in ctx.values():
        d.close()


@pytest.mark.parametrize("ts,name", [(1, "ref_outlet_2d.npz"), (0, "ref_outlet_2d_rk1.npz")])
def test_open_side_2d_fixture(ts, name):
    """An open side in the 2D build (tests/golden/ref_outlet_2d*.npz, the
    x = 0.6 m side): it enters the boundary table of the particles next to
    it as in the oracle (SelectCloseBoundarySides2D.f90:99-107: sides and
    local coordinates identical), adds no wall term, and the particles whose
    step crosses it are removed (CancelOutgoneParticles_2D.f90).  Count and
    particle order equal the reference's after 8 and 16 steps, leapfrog and
    explicit RK1; FP64 fields within FP64_REL_TOL; FP32-mixed removes the
    same particles and is within FP32_TOL."""
    from pathlib import Path
    from oracle.pyoracle import Oracle
    from sphera_amd import runtime as rt
    g = np.load(Path(__file__).resolve().parent / "golden" / name)
    tb = case2d.build(case2d.DamBreak2D(**SMALL, outlet=True, time_split=ts))
    print(f"[gpu] 2d open side {name} n={len(g['ic_coord'])}", flush=True)
    o = Oracle(tb, coord=g["ic_coord"], vel=g["ic_vel"])
    o.preloop()
    ctx = {}
    for prec in (rt.FP64, rt.FP32_MIXED):
        d = rt.SphContext(tb, precision=prec)
        d.upload(coord=g["ic_coord"], vel=g["ic_vel"])
        d.preloop()
        ctx[prec] = d
    _btab_close(ctx[rt.FP64], o)
    dn, dside, _, _, _ = ctx[rt.FP64].btab()
    assert ((dside == 6) & (np.arange(dside.shape[1]) < dn[:, None])).any(axis=1).sum() == 56
    for it in (8, 16):
        for prec, d in ctx.items():
            d.step(8)
            ds = d.state()
            assert ds["nag"] == g[f"s{it}_nag"], (it, prec, ds["nag"])
            errs = {f: _rel(ds[f], g[f"s{it}_" + f]) for f in FIELDS}
            print(f"[gpu] 2d open side it={it} precision={prec} n={ds['nag']} "
                  f"max rel {max(errs.values()):.2e}", flush=True)
            tol = FP64_REL_TOL if prec == rt.FP64 else FP32_TOL["vel"]
            for f in FIELDS:
                assert errs[f] <= tol, (it, prec, f, errs[f])
    for d in ctx.values():
        d.close()
