import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~10 s")


@pytest.fixture(scope="session")
def ox():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hx():
    """The product package (its name starts with a digit, so import it by name)."""
    import importlib
    return importlib.import_module("2025_12_09_hexat_v069_amd")
