"""Shared fixtures. GPU tests carry the ``gpu`` marker; everything else runs on the CPU."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: minutes of CPU work")


@pytest.fixture(scope="session")
def reference_available():
    from oracle import reference

    if not reference.available():
        pytest.skip("oracle/_ref/libasref.so is not built (make -C oracle ref)")
    return True
