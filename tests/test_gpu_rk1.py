"""Explicit RK1 time integration on the device (time_split = 0, RKscheme = 1:
Euler.f90 through time_integration.f90), against the reference's fixture
tests/golden/ref_rk1.npz (the oracle's RK1 path is pinned to it and to the
reference program's own loop in tests/test_oracle_reference.py).

Bars: FP64 fields, dt and time bit-identical; FP32-mixed within 1e-4 of the
field's largest |value|; a replayed step graph equals eager steps bit for bit;
segments of a batch equal the same solves alone."""
from pathlib import Path

import numpy as np
import pytest

from sphera_amd.case import DamBreak3D, build

pytestmark = pytest.mark.gpu

FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")
GOLD = Path(__file__).resolve().parent / "golden"


def _case(**kw):
    return DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01, name="small",
                      time_split=0, **kw)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _ctx(tb, precision, coord=None):
    from sphera_amd import runtime as rt
    d = rt.SphContext(tb, precision=precision)
    d.upload(coord=coord)
    d.preloop()
    return d


def test_rk1_matches_reference_fixture():
    from sphera_amd import runtime as rt
    g = np.load(GOLD / "ref_rk1.npz")
    tb = build(_case())
    print(f"[gpu] rk1 n={len(g['ic_coord'])}", flush=True)
    d64 = _ctx(tb, rt.FP64, g["ic_coord"])
    d32 = _ctx(tb, rt.FP32_MIXED, g["ic_coord"])
    for it in (6, 12):
        d64.step(6)
        d32.step(6)
        ds, fs = d64.state(), d32.state()
        assert ds["nag"] == g[f"s{it}_nag"]
        assert ds["dt"] == g[f"s{it}_dt"] and ds["time"] == g[f"s{it}_time"], it
        for f in FIELDS:
            assert np.array_equal(ds[f], g[f"s{it}_" + f]), (it, f, _rel(ds[f], g[f"s{it}_" + f]))
        e32 = {f: _rel(fs[f], g[f"s{it}_" + f]) for f in FIELDS}
        print(f"[gpu] rk1 it={it} fp64 bit-identical; fp32 {{{', '.join(f'{k}: {v:.1e}' for k, v in e32.items())}}}",
              flush=True)
        for f in FIELDS:
            assert e32[f] <= 1e-4, (it, f, e32[f])
    d64.close()
    d32.close()


def test_rk1_oracle_longer_and_graph():
    """40 steps of the jittered case against the oracle (FP64 bit-identical),
    and a replayed 10-step graph against eager steps (both precisions)."""
    from oracle.pyoracle import Oracle
    from sphera_amd import runtime as rt
    tb = build(_case())
    rng = np.random.default_rng(3)
    c = tb.particles["coord"]
    coord = c + rng.uniform(-0.01, 0.01, c.shape) * tb.case.dx
    o = Oracle(tb, coord=coord)
    o.preloop()
    o.step(40)
    os_ = o.state()
    for prec in (rt.FP64, rt.FP32_MIXED):
        a = _ctx(tb, prec, coord)
        b = _ctx(tb, prec, coord)
        a.step(40)
        for _ in range(4):
            b.step_graph(10)
        b.sync()
        sa, sb = a.state(), b.state()
        for f in FIELDS:
            assert np.array_equal(sa[f], sb[f]), (prec, f)
            if prec == rt.FP64:
                assert np.array_equal(sa[f], os_[f]), (f, _rel(sa[f], os_[f]))
            else:
                assert _rel(sa[f], os_[f]) <= 1e-4, (f, _rel(sa[f], os_[f]))
        a.close()
        b.close()


def test_rk1_phase_entry_points_refused():
    from sphera_amd import runtime as rt
    d = _ctx(build(_case()), rt.FP64)
    with pytest.raises(rt.SphError) as e:
        d.phase("dt")
    assert e.value.code == 2
    d.step(1)  # the context still steps
    d.close()


def test_rk1_batch_segments_equal_alone():
    from sphera_amd import runtime as rt
    cases = [DamBreak3D(tank=(0.4, 0.2, 0.3), column=(w, 0.1, 0.12), dx=dx, name=f"s{k}", time_split=0)
             for k, (w, dx) in enumerate(((0.12, 0.01), (0.2, 0.012), (0.08, 0.009)))]
    tbs = [build(c) for c in cases]
    for prec in (rt.FP64, rt.FP32_MIXED):
        b = rt.SphBatch(tbs, precision=prec)
        b.upload()
        b.preloop()
        for _ in range(8):
            b.step_graph(1)
        b.sync()
        for k, tb in enumerate(tbs):
            alone = _ctx(tb, prec)
            alone.step(8)
            sa, sb = alone.state(), b.seg_state(k)
            for f in FIELDS:
                assert np.array_equal(sa[f], sb[f]), (prec, k, f)
            alone.close()
        b.close()


def test_rk1_2d_matches_oracle():
    """The 2D form with explicit RK1 (the oracle's 2D RK1 path equals the
    reference's unmodified 2D program, tests/test_oracle_2d.py): FP64 within
    the 2D form's bar (its corner integrals use the device's trigonometry),
    FP32-mixed within 1e-4, after 20 steps of a jittered tank."""
    from oracle.pyoracle import Oracle
    from sphera_amd import case2d
    from sphera_amd import runtime as rt
    from tests.test_gpu_2d import FP64_REL_TOL, _jitter2d
    tb = case2d.build(case2d.DamBreak2D(tank=(0.6, 0.4), column=(0.3, 0.2), dx=0.01, name="t2d",
                                        time_split=0))
    coord = _jitter2d(tb.particles["coord"], tb.case.dx)
    o = Oracle(tb, coord=coord)
    o.preloop()
    o.step(20)
    os_ = o.state()
    print(f"[gpu] rk1 2d n={len(coord)}", flush=True)
    for prec, tol in ((rt.FP64, FP64_REL_TOL), (rt.FP32_MIXED, 1e-4)):
        d = _ctx(tb, prec, coord)
        d.step(20)
        ds = d.state()
        d.close()
        for f in FIELDS:
            assert _rel(ds[f], os_[f]) <= tol, (prec, f, _rel(ds[f], os_[f]))
