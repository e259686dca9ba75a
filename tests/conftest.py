import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session", autouse=True)
def _libraries():
    # A library whose sources changed since it was built (content hash
    # recorded by the build) is rebuilt where there is no GPU; on a machine
    # with a GPU the build belongs before the test run, so the run stops
    # instead of testing old kernels.
    from mag2d_amd import build
    stale = build.stale_libraries()
    if stale:
        if os.path.exists("/dev/kfd"):
            pytest.exit(f"stale or missing native libraries {stale}: build them first "
                        "(python -c 'import __graft_entry__ as g; g.build()')", returncode=3)
        build.build_all()
    yield


@pytest.fixture(scope="session")
def ctx():
    from mag2d_amd import abi
    c = abi.Context(0, seed=1234)
    yield c
    c.close()
