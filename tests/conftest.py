import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def tiny_case():
    from sphera_amd.case import DamBreak3D
    return DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny")


@pytest.fixture(scope="session")
def small_case():
    from sphera_amd.case import DamBreak3D
    return DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01, name="small")
