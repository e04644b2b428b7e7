import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long-running GPU case")


@pytest.fixture(scope="session")
def hip():
    """The engine, asserted usable: a GPU test never passes without running the device path."""
    from qubitserf_amd import engine

    assert engine.available(), "gfx950 engine unavailable: " + \
        (engine._native.load().qsf_hip_last_error() or b"").decode()
    return engine
