"""C ABI: the device library loads and exports every entry point declared in
include/sphera_hip.h (no device calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared():
    txt = (ROOT / "include" / "sphera_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_phase_api():
    names = declared()
    for ph in ("sph_dt", "sph_momentum", "sph_velocity_smoothing", "sph_sort", "sph_neighbours",
               "sph_boundary_tab", "sph_continuity", "sph_pressure_smoothing", "sph_step"):
        assert ph in names


def test_library_exports_every_declared_symbol():
    from sphera_amd import runtime
    if not runtime.LIB_PATH.exists():
        pytest.skip("device library not built")
    lib = ctypes.CDLL(str(runtime.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(runtime.EXPORTS) <= set(declared())


def test_dropin_harness_fails_loudly_without_device(tmp_path):
    """The reference program linked with the Fortran bridge stops with the C
    ABI's status and message when no device is there (no CPU fallback)."""
    import torch
    from oracle.decks import HARNESS_HIP, run_harness
    from sphera_amd.case import DamBreak3D
    if not HARNESS_HIP.exists():
        pytest.skip("drop-in harness not built")
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny")
    with pytest.raises(RuntimeError, match="sph_ctx_create failed, status +1"):
        run_harness(case, 2, lists=False, workdir=tmp_path, exe=HARNESS_HIP, timeout=60)


def test_dropin_harness_2d_fails_loudly_without_device(tmp_path):
    """The same for the reference's 2D build with the bridge (sph_ctx_create2d)."""
    import torch
    from oracle.decks import HARNESS2D_HIP, run_harness
    from sphera_amd.case2d import DamBreak2D
    if not HARNESS2D_HIP.exists():
        pytest.skip("2D drop-in harness not built")
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    case = DamBreak2D(tank=(0.6, 0.4), column=(0.2, 0.1), dx=0.01, name="t2d")
    with pytest.raises(RuntimeError, match="sph_ctx_create2d failed, status +1"):
        run_harness(case, 2, lists=False, workdir=tmp_path, exe=HARNESS2D_HIP, timeout=60)


def test_batch_bytes_layout():
    """Sizing of segmented batches (no device call): a batch of one segment
    is sized as a one-solve context; a batch of several holds at least the sum
    of its segments' particle arrays (slots rounded to whole 256-blocks)."""
    from sphera_amd import runtime as rt
    if not rt.LIB_PATH.exists():
        pytest.skip("device library not built")
    from sphera_amd.case import DamBreak3D, build
    a = build(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="a"))
    b = build(DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.2, 0.15), dx=0.012, name="b"))
    for prec in (rt.FP64, rt.FP32_MIXED):
        one = rt.batch_bytes([a], prec)
        assert one == rt.ctx_bytes(a, prec)
        both = rt.batch_bytes([a, b], prec)
        # the scalar blocks and the list-stride rounding are shared; everything
        # per particle and per cell is each segment's own
        assert both >= rt.ctx_bytes(a, prec) + rt.ctx_bytes(b, prec) - 4096
