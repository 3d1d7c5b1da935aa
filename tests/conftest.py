"""Shared test setup: import paths, the `gpu` marker, and library fixtures."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("video-to-3d_amd", "oracle", "tests"):
    path = os.path.join(REPO, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    import vt3d_oracle

    vt3d_oracle.library()
    return vt3d_oracle


@pytest.fixture(scope="session")
def product():
    import vt3d_dense

    vt3d_dense.library()
    return vt3d_dense
