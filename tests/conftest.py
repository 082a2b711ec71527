import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Under pytest-xdist every worker would start an OpenMP team as wide as the machine for
# the oracle's parallel runs; their spinning teams then take minutes per test.  Each
# worker gets its share of the cores instead (before the oracle library is loaded).
if "PYTEST_XDIST_WORKER_COUNT" in os.environ and "OMP_NUM_THREADS" not in os.environ:
    _workers = max(1, int(os.environ["PYTEST_XDIST_WORKER_COUNT"]))
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // _workers))

sys.path.insert(0, os.path.join(ROOT, "physics-sim_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
