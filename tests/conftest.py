import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kernelmatrices.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: larger CPU oracle case")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded through its C ABI; skips if no device."""
    import kmhodlr
    lib = kmhodlr.load()
    if lib.device_count() < 1:
        pytest.skip("no HIP device visible")
    return lib
