import os
import sys
from pathlib import Path

import pytest

# torch ships its own HIP runtime under the same soname as the system one
# (libamdhip64.so.7). Importing it before libhd.so is loaded makes libhd.so
# bind to that runtime, so torch tensors and the library share one HIP
# runtime in this process; the other order would start two runtimes.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

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
