import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ahotsph-2hot-apps_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def gpu_lib():
    """libpgrav loaded with torch's HIP runtime already in the process."""
    import torch  # noqa: F401  (one HIP runtime per process: torch's)
    import pgrav
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return pgrav
