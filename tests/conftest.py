"""pytest configuration: the `gpu` marker and import paths.

CPU tests (-m "not gpu") check the oracle against the golden fixtures and the
closed forms, and that libmgv.so loads and exports every symbol of
include/mgv.h.  GPU tests (-m gpu) call libmgv.so through its C ABI and
compare with the oracle on the same seeded inputs.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "multigpu-vector_amd"), os.path.join(ROOT, "tests"),
          os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libmgv.so kernels)")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    import oracle_py
    return oracle_py.Oracle()


@pytest.fixture(scope="session")
def gpu():
    """Initialised libmgv on device 0 (GPU tests only)."""
    import mgv
    mgv.load()
    mgv.init(0)
    return mgv
