"""Shared test setup: import paths, the `gpu` marker and common fixtures.

Tests marked `gpu` need an MI355X (gfx950) and the built libwfamd.so; everything else runs on the
CPU. The oracle (oracle/_build/liboracle.so) is the checker.
"""
import hashlib
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


def oracle_threads():
    """Worker threads for the oracle: the CPUs this process may use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


@pytest.fixture(scope="session")
def oracle_ref(po):
    """oracle_ref(scene, spp, seed, depth) -> ((rgb, lum, lum2) film sums, (extension, shadow,
    invalid) ray counts) of the oracle's render of samples [0, spp) of a product scene (a
    wfamd.Scene, any view). Oracle scenes are cached by geometry (their BVH is built once) and
    renders by view, samples, seed and depth."""
    scenes, renders = {}, {}

    def get(scene, spp, seed=1234, depth=8):
        a = scene.arrays()
        g = hashlib.sha1()
        for k in ("spheres", "vertices", "triangles", "materials", "env"):
            g.update(np.ascontiguousarray(a[k]).tobytes())
        geo = g.hexdigest()
        key = (geo, a["camera"].tobytes(), a["width"], a["height"], spp, seed, depth)
        if key not in renders:
            if geo not in scenes:
                scenes[geo] = po.OScene(a)
            o = scenes[geo]
            o.set_view(a["camera"], a["width"], a["height"])
            renders[key] = o.render_sums(0, spp, seed, depth, threads=oracle_threads())
        return renders[key]
    return get


def assert_equals_oracle(sums, counts, want, what=""):
    """Film sums (rgb, lum, lum2) bit for bit and ray counts (extension, shadow, invalid) equal to
    an oracle_ref result."""
    (orgb, olum, olum2), ocnt = want
    for a, b, k in zip(sums, (orgb, olum, olum2), ("rgb", "lum", "lum2")):
        assert_bitwise(a, b, f"{what} {k} vs oracle")
    if counts is not None:
        assert tuple(int(c) for c in counts) == tuple(int(c) for c in ocnt), f"{what}: rays {counts} oracle {ocnt}"


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bitwise(a, b, what=""):
    a, b = bits(a), bits(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a != b)
        raise AssertionError(f"{what}: {bad.size} of {a.size} words differ, first at {bad[:8].tolist()}")
