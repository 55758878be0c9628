import os
import sys

# torch ships its own HIP runtime under the same soname as the system one; if
# torch is loaded first, libngsamd binds to that one and the process has a
# single runtime (needed by the torch-tensor tests of ngsolve_amd.autograd)
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    pass

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: full-size configuration")


@pytest.fixture(scope="session")
def device():
    from ngsolve_amd._native import ensure_device
    return ensure_device(0)
