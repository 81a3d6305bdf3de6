"""Deck options the device does not build are refused, not run without them.

sph_ctx_create / sph_ctx_create2d / sph_batch_create return status 95 for
ALE3, C1_BE / C1_divu and KTGF_config > 0, and status 2 for a
slip_coefficient_mode outside 0-2 or a convex-edge option that disagrees with
the edge table.  The checks run before any device call, so these tests run
without a GPU.  The density clamp (density_thresholds = 1), the SA-SPH wall
function (slip_coefficient_mode = 2) and convex frontier edges (NumBEdges > 0)
are built and are accepted; their parity tests are with the other parity
tests."""
import ctypes as C

import numpy as np
import pytest

from sphera_amd import runtime as rt
from sphera_amd.case import DamBreak3D, build
from sphera_amd.case2d import DamBreak2D, build as build2d

pytestmark = pytest.mark.skipif(not rt.LIB_PATH.exists(), reason="device library not built")


def _tiny():
    return build(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny"))


def _create(tb, options=0, faces_patch=None):
    L = rt.load()
    p = rt.params_struct(tb)
    p.options = options
    fs, keep = rt.faces_struct(tb)
    if faces_patch:
        faces_patch(fs, keep)
    h = C.c_void_p()
    rc = L.sph_ctx_create(C.byref(p), C.byref(fs), C.byref(h))
    return rc, L.sph_last_error().decode()


@pytest.mark.parametrize("opt,what", [(rt.OPT_ALE3, "ALE3"), (rt.OPT_C1, "C1_BE"),
                                      (rt.OPT_KTGF, "KTGF_config")])
def test_unbuilt_option_refused(opt, what):
    rc, msg = _create(_tiny(), options=opt | rt.OPT_DENSITY_THRESHOLDS)
    assert rc == 95 and what in msg, (rc, msg)


def test_convex_edge_option_matches_the_edge_table():
    """SPH_OPT_CONVEX_EDGES (NumBEdges > 0) is built; it must agree with the
    convex edges the frontier faces carry, and a 2D context has none."""
    tb = _tiny()
    rc, msg = _create(tb, options=rt.OPT_CONVEX_EDGES)
    assert rc == 2 and "SPH_OPT_CONVEX_EDGES" in msg, (rc, msg)
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny",
                      obstacle=(0.15, 0.2, 0.04))
    tbe = build(case)
    assert rt.params_struct(tbe).options & rt.OPT_CONVEX_EDGES
    rc, msg = _create(tbe, options=0)
    assert rc == 2 and "SPH_OPT_CONVEX_EDGES" in msg, (rc, msg)

    def bad_face(fs, keep):
        keep["ef"][:] = 99
    rc, msg = _create(tbe, options=rt.OPT_CONVEX_EDGES, faces_patch=bad_face)
    assert rc == 2 and "out of range" in msg, (rc, msg)
    L = rt.load()
    t2 = build2d(DamBreak2D(tank=(0.6, 0.4), column=(0.2, 0.1), dx=0.01, name="t2d"))
    p = rt.params_struct(t2)
    p.options = rt.OPT_CONVEX_EDGES
    ss, keep = rt.sides_struct(t2)
    h = C.c_void_p()
    assert L.sph_ctx_create2d(C.byref(p), C.byref(ss), C.byref(h)) == 2


def test_unknown_option_bits_rejected():
    rc, msg = _create(_tiny(), options=1 << 10)
    assert rc == 2 and "unknown option" in msg


def test_unknown_slip_mode_rejected_3d():
    def patch(fs, keep):
        keep["mode"][:] = 3
    rc, msg = _create(_tiny(), faces_patch=patch)
    assert rc == 2 and "slip_coefficient_mode must be 0, 1 or 2" in msg, (rc, msg)


def test_unknown_slip_mode_rejected_2d():
    L = rt.load()
    tb = build2d(DamBreak2D(tank=(0.6, 0.4), column=(0.2, 0.1), dx=0.01, name="t2d"))
    p = rt.params_struct(tb)
    ss, keep = rt.sides_struct(tb)
    keep["mode"][:] = -1
    h = C.c_void_p()
    rc = L.sph_ctx_create2d(C.byref(p), C.byref(ss), C.byref(h))
    assert rc == 2 and "slip_coefficient_mode must be" in L.sph_last_error().decode()


def test_batch_refuses_unbuilt_option():
    L = rt.load()
    tbs = [_tiny(), _tiny()]
    ps = (rt.SphParams * 2)(*[rt.params_struct(tb, rt.FP32_MIXED) for tb in tbs])
    ps[1].options = rt.OPT_ALE3
    packed = [rt.faces_struct(tb) for tb in tbs]
    fptrs = (C.POINTER(rt.SphFaces) * 2)(*[C.pointer(f) for f, _ in packed])
    h = C.c_void_p()
    assert L.sph_batch_create(2, ps, fptrs, C.byref(h)) == 95


def test_params_carry_density_thresholds():
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny",
                      density_thresholds=1)
    assert rt.params_struct(build(case)).options == rt.OPT_DENSITY_THRESHOLDS
    assert rt.params_struct(_tiny()).options == 0


def test_rk2_refused_rk1_accepted_in_params():
    rc, msg = _create(_tiny(), options=rt.OPT_RK2)
    assert rc == 95 and "Heun" in msg, (rc, msg)
    rc, msg = _create(_tiny(), options=rt.OPT_RK1 | rt.OPT_RK2)
    assert rc == 2
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny", time_split=0)
    assert rt.params_struct(build(case)).options == rt.OPT_RK1
