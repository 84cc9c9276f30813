import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "walk-on-spheres_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def reference():
    import pyoracle
    if not os.path.exists(pyoracle.REF_SO):
        pytest.skip("reference core (oracle/_ref) not built here")
    return pyoracle.Reference()


@pytest.fixture(scope="session")
def gpu_ctx():
    from wos_amd import Context
    return Context(0)
