"""Shared pytest configuration: the `gpu` marker, import paths and graph
fixtures used by the CPU and GPU suites."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "parallelization-artifact-sc23_amd")
for p in (PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ig():
    import indigo2

    indigo2.load()
    return indigo2


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def ctx(ig):
    c = ig.Context(0)
    yield c
    c.close()
