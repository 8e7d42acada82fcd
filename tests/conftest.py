import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as orc
    if not os.path.exists(orc.LIB):
        orc.build()
    orc.lib()
    return orc


@pytest.fixture(scope="session")
def qjx_lib():
    """The product library; built in-tree if it is missing (CPU-side check only)."""
    from os_d_amd import qjx, build
    if not os.path.exists(qjx.LIB_PATH):
        build.build_lib()
    return qjx.load_library()
