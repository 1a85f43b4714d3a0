"""Shared test setup: import paths, the `gpu` marker, device fixtures."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "gpufinitefieldmatrices.jl_amd")
for p in (REPO, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def gffm():
    import gpufinitefieldmatrices as g
    g.context()  # fails loudly when there is no GPU or no library
    return g


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle
