import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "proton-dose-m1_amd", ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libm1march.so")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs (whole-march goldens at 1 mm)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O

    if not O.LIB_PATH.exists():
        O.build()
    return O
