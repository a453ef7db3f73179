import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs libflyehip.so kernels)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """One libflyehip.so context for the GPU tests of the session."""
    from cuflye_amd.overlap import FlyeHip
    ctx = FlyeHip(device=0)
    yield ctx
    ctx.close()
