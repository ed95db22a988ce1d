import sys
from pathlib import Path

import pytest

# PyTorch-ROCm ships its own HIP runtime.  Load it before libqcg.so (which links the system ROCm
# runtime) so that tests which use torch device buffers find the GPU: with the opposite load order
# torch's runtime initialises after the system one and reports no devices.
try:
    import torch  # noqa: F401
except ImportError:  # CPU-only environments without torch still run the oracle tests
    torch = None

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
