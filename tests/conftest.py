import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "horton-part_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def require_gpu():
    from horton_part_amd import _native

    n = _native.device_count()
    if n < 1:
        raise RuntimeError("GPU test needs a visible GPU: " + _native.load().hp_last_error().decode())
