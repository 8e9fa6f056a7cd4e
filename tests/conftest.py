import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session", autouse=True)
def _libraries_built():
    """The oracle library is the checker of every test: build it if absent.
    The HIP library is rebuilt if it was built from other sources (the
    loader refuses such a library)."""
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    import __graft_entry__
    __graft_entry__.build_if_stale()
