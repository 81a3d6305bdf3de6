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
