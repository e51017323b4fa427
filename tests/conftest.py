import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def native_lib():
    """The built HIP library (built on demand when missing)."""
    from cudalbm_amd import native, build

    if not os.path.exists(native.LIB_PATH):
        build.build()
    return native.lib()
