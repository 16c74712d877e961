"""Shared test setup: import paths, the `gpu` marker and common fixtures.

Tests marked `gpu` need an MI355X (gfx950) and the built libwfamd.so; everything else runs on the
CPU. The oracle (oracle/_build/liboracle.so) is the checker.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "graphics-book-tutorials_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def wf():
    import wfamd
    wfamd.lib()
    return wfamd


@pytest.fixture(scope="session")
def po():
    import pyoracle
    pyoracle.oracle()
    return pyoracle


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bitwise(a, b, what=""):
    a, b = bits(a), bits(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a != b)
        raise AssertionError(f"{what}: {bad.size} of {a.size} words differ, first at {bad[:8].tolist()}")
