import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def ps():
    """The native library, initialised on device 0 (GPU tests only)."""
    import puzzlesearch_amd as psa
    psa.init(0)
    return psa


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(TESTS, "golden", "reference_vectors.json")) as f:
        return {r["id"]: r for r in json.load(f)}
