"""Shared test setup: import paths, the `gpu` marker, oracle/library handles."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "b-tree-index-searching_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def bpt():
    import bptree
    bptree.lib()
    if bptree.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")
    return bptree


@pytest.fixture(scope="session")
def wl():
    import workloads
    return workloads
