import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def _built():
    return all((ROOT / f).exists() for f in ("mag2d_amd/libmag2d_amd.so", "mag2d_amd/host/libmag2d_host.so",
                                              "oracle/liboracle.so"))


@pytest.fixture(scope="session", autouse=True)
def _libraries():
    # load what is built; build only when nothing is
    if not _built():
        from mag2d_amd import build
        build.build_all()
    yield


@pytest.fixture(scope="session")
def ctx():
    from mag2d_amd import abi
    c = abi.Context(0, seed=1234)
    yield c
    c.close()
