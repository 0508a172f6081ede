import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def orc():
    """The oracle package (built on first use)."""
    from oracle import oracle as o
    o.oracle_lib()
    return o


REALTIME = {"simulation": {"max_balls": 300, "resolution": 25}}
