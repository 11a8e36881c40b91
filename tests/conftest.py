import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def orc64():
    from oracle.oracle import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def orc32():
    from oracle.oracle import Oracle
    return Oracle(32)


@pytest.fixture(scope="session")
def refsbp():
    from oracle.oracle import RefSBP, ref_available
    if not ref_available():
        pytest.skip("oracle/_ref/libref_sbp.so not built (reference sources absent)")
    return RefSBP()
