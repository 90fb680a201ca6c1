import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "icrar-leap-accelerate_amd")
for p in (os.path.join(REPO, "tests"), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle_bindings import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def leap():
    import leap_amd

    leap_amd.lib()  # raises when the library has not been built
    return leap_amd


@pytest.fixture(scope="session")
def ctx(leap):
    c = leap.Context(0)
    yield c
    c.close()
