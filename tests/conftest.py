import os
import sys

import pytest

# numpy and scipy each bundle an OpenBLAS; their thread pools contend when
# calls alternate, so the oracle runs with one BLAS thread (the reference's
# own BLAS.set_num_threads(1), test/runtests.jl:9).
try:
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)
except Exception:  # pragma: no cover
    pass

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
