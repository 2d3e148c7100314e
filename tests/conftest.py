import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libhd.so")


@pytest.fixture(scope="session")
def ctx():
    import hausdorffdistance_amd as hd

    c = hd.Context(0)
    yield c
    c.close()
