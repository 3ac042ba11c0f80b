import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
PKG = REPO / "particle-stabilizer_amd"
for p in (str(REPO), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import ref
    ref.load()
    return ref


@pytest.fixture(scope="session")
def psx_mod():
    import psx
    psx.load()  # raises if the HIP backend is not built: GPU tests never fall back
    return psx


def pytest_report_header(config):
    return f"psx package: {PKG}"
