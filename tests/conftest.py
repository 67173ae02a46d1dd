import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP engine on a real GPU.  Fails (not skips) when the library is
    missing; skips only when no GPU is visible."""
    from hmglib_amd import hip
    import hmglib_amd
    hmglib_amd.lib()  # raises when libhmglib.so is missing
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")
    return hmglib_amd
