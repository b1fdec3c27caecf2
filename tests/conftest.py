import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a gfx950 GPU (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def doki():
    """The product `doki` extension; GPU tests only."""
    import doki_amd
    if not doki_amd.built():
        doki_amd_build()
    return doki_amd.doki


def doki_amd_build():
    from doki_amd import build
    build.build()


@pytest.fixture(scope="session")
def capi():
    import doki_amd
    if not doki_amd.built():
        doki_amd_build()
    from doki_amd import capi as A
    A.check(A.lib().dk_init(0))
    return A
