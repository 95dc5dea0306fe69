import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsurfacegrowth.so on the device)")
    config.addinivalue_line("markers", "slow: longer CPU test")
    lib = os.path.join(ROOT, "surfacegrowthconsole_amd", "lib", "libsurfacegrowth.so")
    orc = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not (os.path.exists(lib) and os.path.exists(orc)):
        import __graft_entry__

        __graft_entry__.build()


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU gate: a gpu-marked test on a machine without a device is an error, not a pass."""
    from surfacegrowthconsole_amd import _native

    n = _native.lib().sgk_device_count()
    if n == 0:
        pytest.fail("no HIP device visible: gpu tests need an MI355X")
    return n
