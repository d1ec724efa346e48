import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ics-12-overlapped-tiling_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on a machine with one)")


@pytest.fixture(scope="session")
def otlib():
    import otlib as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def oracle():
    import oracle as m
    m.lib()
    return m
