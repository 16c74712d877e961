"""Integer indexing of the persistent engine's new paths: division by an invariant divisor.

New paths are generated from their work index (job, sample, pixel in 8x8 tiles); the sample index
and the tile band are quotients by per-job constants (width * height, 8 * width), computed as a
multiply-high and shifts (Granlund and Montgomery, 1994). No reference counterpart: the check is
plain integer division, over divisors that stress the method (small, powers of two and their
neighbours, random, the benchmark image sizes) and dividends up to 2^32 - 1.
"""
import numpy as np
import pytest


def cases():
    rs = np.random.default_rng(1)
    ds = list(range(1, 1200)) + [2 ** k for k in range(32)] + [2 ** k + 1 for k in range(31)] + \
        [2 ** k - 1 for k in range(2, 33)] + rs.integers(1, 2 ** 32, 300, dtype=np.uint64).tolist() + \
        [1280 * 720, 8 * 1280, 320 * 180, 8 * 320, 1920 * 1080, 8 * 1920, 3 * 7 * 11 * 13 * 17 * 19]
    n = np.concatenate([np.arange(0, 3000, dtype=np.uint64), rs.integers(0, 2 ** 32, 8000, dtype=np.uint64),
                        np.array([2 ** 32 - 1, 2 ** 32 - 2, 2 ** 31, 2 ** 31 - 1], np.uint64)]).astype(np.uint32)
    return [int(d) for d in ds if d < 2 ** 32], n


def test_divu32_equals_integer_division(wf):
    ds, n = cases()
    for d in ds:
        want = (n.astype(np.uint64) // d).astype(np.uint32)
        got = wf.test_divu32(n, d)
        assert np.array_equal(got, want), d
        # dividends just below and at multiples of d
        m = np.array([d * k - 1 for k in range(1, 4) if d * k - 1 < 2 ** 32] +
                     [d * k for k in range(1, 4) if d * k < 2 ** 32], np.uint64).astype(np.uint32)
        assert np.array_equal(wf.test_divu32(m, d), (m.astype(np.uint64) // d).astype(np.uint32)), d


@pytest.mark.gpu
def test_divu32_device_equals_host(wf):
    ds, n = cases()
    for d in ds[::7]:
        assert np.array_equal(wf.test_divu32(n, d, device=True), wf.test_divu32(n, d)), d
