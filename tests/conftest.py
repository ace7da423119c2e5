import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qrochet.jl_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqrochet_hip.so on the device)")


@pytest.fixture(scope="session")
def qh():
    """The host package; GPU tests fail (not skip) if the library cannot load."""
    import qrochet_hip

    qrochet_hip.default_context()
    return qrochet_hip
