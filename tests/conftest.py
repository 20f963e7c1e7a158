import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built liblbm_hip.so")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle.oracle import load_oracle

    return load_oracle()
