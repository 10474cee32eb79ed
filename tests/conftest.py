"""pytest configuration: import paths and the `gpu` marker.

Tests marked `gpu` run the HIP library on an MI355X; the rest run on the CPU
(the oracle, problem generators, and the C ABI's load/export checks).
"""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "legion-snap_amd")
for p in (PKG, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on the current device; fails (not skips) when the
    library or the GPU is missing, so GPU runs never pass without native code."""
    import torch
    from snap_amd import _lib
    if not torch.cuda.is_available():
        raise RuntimeError("gpu tests need a visible GPU")
    lib = _lib.load()
    torch.cuda.init()
    return lib
