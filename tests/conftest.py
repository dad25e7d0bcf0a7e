import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "phase-field-project-3d_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle as O

    O.build()
    return O


def gpu_available() -> bool:
    """True on a machine with an AMD GPU driver (/dev/kfd).  Deliberately
    independent of libpfamd.so: on a GPU box a library that fails to load
    must fail the GPU tests, not skip them."""
    return os.path.exists("/dev/kfd")
