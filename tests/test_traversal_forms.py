"""The GPU traversal's slab test and visit forms against the reference-order slab test.

The engines test a node with the per-ray-signed slab form (hit_box_signed) whenever a ray has a
finite origin and finite, nonzero 1/d, and fall back to the reference-order form
(include/wf/bvh.hpp:64-78) for the whole wave otherwise. The signed form must give the same
boolean on every box with lo <= hi; these cases aim at the places where the two could part: a
zero-thickness box, an origin exactly on a box plane, a slab bound that overflows to infinity,
signed zeros, and tmin/tmax at the ends of the range. Then the wave-scheduled traversal itself
(all nodes in LDS, the top in LDS and the rest from memory, and the exact fallback) must give the
host's closest and any hit bit for bit, on random and on axis-aligned rays whose 1/d is infinite;
on the reference scenes, the axis-ray results are the reference's own (tests/golden).
"""
import os

import numpy as np
import pytest

from conftest import assert_bitwise

FMAX = np.float32(3.402823466e38)


def slab_cases(seed, n):
    rs = np.random.default_rng(seed)
    f32 = np.float32
    lo = rs.uniform(-10, 10, (n, 3)).astype(f32)
    ext = rs.uniform(0, 5, (n, 3)).astype(f32)
    ext[rs.random((n, 3)) < 0.2] = 0.0                      # zero-thickness slabs
    hi = (lo + ext).astype(f32)
    big = rs.random(n) < 0.05                                # bounds that overflow (b - o) / d
    lo[big] *= f32(1e37)
    hi[big] = np.maximum(hi[big] * f32(1e37), lo[big])
    o = rs.uniform(-15, 15, (n, 3)).astype(f32)
    on_lo = rs.random((n, 3)) < 0.2
    on_hi = (rs.random((n, 3)) < 0.2) & ~on_lo
    o[on_lo] = lo[on_lo]                                     # origin exactly on a box plane
    o[on_hi] = hi[on_hi]
    d = rs.normal(size=(n, 3))
    aim = rs.random(n) < 0.6                                 # most rays aimed at the box
    target = lo.astype(np.float64) + rs.random((n, 3)) * (hi.astype(np.float64) - lo)
    d[aim] = target[aim] - o[aim] + rs.normal(scale=0.05, size=(aim.sum(), 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f32)
    tiny = rs.random((n, 3)) < 0.08
    d[tiny] = np.copysign(rs.choice([1e-30, 1e-37, 1e-20], tiny.sum()), d[tiny]).astype(f32)
    zero = rs.random((n, 3)) < 0.05
    d[zero] = np.copysign(f32(0.0), rs.choice([-1.0, 1.0], zero.sum())).astype(f32)   # +0 and -0
    tmin = rs.choice(np.array([1e-4, 0.0, -1.0, 2.0], f32), n)
    tmax = rs.choice(np.array([FMAX, 1.0, 7.5, 30.0, 1e-4, 0.0], f32), n)
    tmax[rs.random(n) < 0.01] = np.nan
    return np.concatenate([lo, hi, o, d, tmin[:, None], tmax[:, None]], axis=1).astype(f32)


def applies(cases):
    o, d = cases[:, 6:9], cases[:, 9:12]
    with np.errstate(divide="ignore", over="ignore"):
        inv = np.float32(1.0) / d
    return (np.isfinite(inv) & (inv != 0)).all(axis=1) & np.isfinite(o).all(axis=1) & \
        ~np.isnan(cases[:, 12]) & ~np.isnan(cases[:, 13])


def test_signed_slab_equals_reference_order(wf):
    cases = slab_cases(11, 200000)
    out = wf.test_slab(cases)
    ok = applies(cases)
    assert np.array_equal(out[:, 1] >= 0, ok)                # applies exactly where documented
    assert ok.mean() > 0.6
    ref, sig = out[ok, 0], out[ok, 1]
    assert 0.05 < ref.mean() < 0.95                          # both outcomes well represented
    bad = np.nonzero(ref != sig)[0]
    assert len(bad) == 0, cases[ok][bad[:5]]


def test_signed_slab_on_planes_and_flat_boxes(wf):
    # Exhaustive small grid: boxes on {-1, 0, 1} with lo <= hi, origins on the same grid (so on a
    # plane or in a zero-thickness slab whenever coordinates coincide), all 26 sign patterns of
    # direction, tmin 0 / 1e-4, tmax 0.5 / 3.
    g = np.array([-1.0, 0.0, 1.0], np.float32)
    pairs = np.array([(a, b) for a in g for b in g if a <= b], np.float32)
    rs = np.random.default_rng(5)
    rows = []
    for _ in range(20000):
        lh = pairs[rs.integers(0, len(pairs), 3)]
        o = rs.choice(g, 3)
        d = rs.choice(np.array([-1.0, -0.5, 0.25, 1.0], np.float32), 3)
        rows.append([*lh[:, 0], *lh[:, 1], *o, *d, rs.choice([0.0, 1e-4]), rs.choice([0.5, 3.0])])
    cases = np.array(rows, np.float32)
    out = wf.test_slab(cases)
    assert (out[:, 1] >= 0).all()
    assert np.array_equal(out[:, 0], out[:, 1])
    assert 0.05 < out[:, 0].mean() < 0.95


def axis_rays(scene, n, seed):
    """Rays along the axes (two exact zeros in d, either sign) from origins on node planes."""
    nodes, _ = scene.bvh()
    lo = nodes[:, 0:3].view(np.float32)
    hi = nodes[:, 4:7].view(np.float32)
    rs = np.random.default_rng(seed)
    k = rs.integers(0, len(nodes), n)
    o = np.where(rs.random((n, 3)) < 0.5, lo[k], hi[k]) + \
        rs.choice(np.array([0.0, 0.0, 1e-3, -1e-3], np.float32), (n, 3))
    d = np.zeros((n, 3), np.float32)
    axis = rs.integers(0, 3, n)
    d[np.arange(n), axis] = rs.choice([-1.0, 1.0], n)
    zs = rs.random((n, 3)) < 0.5
    d[(d == 0) & zs] = -0.0
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = FMAX
    return rays


def random_rays(scene, n, seed):
    a = scene.arrays()
    pts = a["vertices"] if len(a["vertices"]) else a["spheres"][:, :3]
    c = pts.mean(axis=0)
    rad = float(np.abs(pts - c).max()) + 1.0
    rs = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = c + rs.uniform(-rad, rad, (n, 3))
    d = rs.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = 1e-4
    rays[:, 7] = FMAX
    return rays


REF_SCENES = ["cornell", "divergence", "trianglehv", "furnace"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "intersect.npz")


def golden_axis(name):
    g = np.load(GOLDEN)
    return (g[name + "_axis_rays"], (g[name + "_axis_t"], g[name + "_axis_prim"]),
            (g[name + "_axis_any_t"], g[name + "_axis_any_prim"]))


@pytest.mark.parametrize("name", REF_SCENES)
def test_axis_rays_match_reference_golden(wf, po, name):
    # The reference's own traversal on rays whose 1/d has infinite components, from origins on
    # node planes (0 * inf slab bounds): the host traversal and the oracle give its results.
    rays, closest, anyhit = golden_axis(name)
    s = wf.Scene(name, 1280, 720)
    o = po.OScene(s.arrays())
    for (gt, gp), flag in ((closest, False), (anyhit, True)):
        t, p = wf.test_intersect(s, rays, flag)
        assert_bitwise(t, gt, f"{name} product t anyhit={flag}")
        assert np.array_equal(p, gp)
        ot, op = o.intersect(rays, flag)
        assert_bitwise(ot, gt, f"{name} oracle t anyhit={flag}")
        assert np.array_equal(op, gp)


TRAV_SCENES = [("cornell555", 32, 18), ("sphere_grid", 32, 18), ("heightfield", 32, 18)]


@pytest.mark.gpu
def test_signed_slab_device_equals_host(wf):
    cases = slab_cases(11, 200000)
    assert np.array_equal(wf.test_slab(cases, device=True), wf.test_slab(cases))


@pytest.mark.gpu
@pytest.mark.parametrize("name", REF_SCENES)
@pytest.mark.parametrize("device", [1, 2, 3])
def test_device_traversal_matches_reference_golden_on_axis_rays(wf, name, device):
    print(f"[axis golden] {name} device mode {device}", flush=True)
    rays, closest, anyhit = golden_axis(name)
    s = wf.Scene(name, 1280, 720)
    for (gt, gp), flag in ((closest, False), (anyhit, True)):
        t, p = wf.test_intersect(s, rays, flag, device=device)
        assert_bitwise(t, gt, f"{name} t anyhit={flag}")
        assert np.array_equal(p, gp)


@pytest.mark.gpu
@pytest.mark.parametrize("name,w,h", TRAV_SCENES)
@pytest.mark.parametrize("device", [2, 3])
def test_wave_traversal_equals_host(wf, name, w, h, device):
    print(f"[wave traversal] {name} device mode {device}", flush=True)
    s = wf.Scene(name, w, h)
    rnd, axr = random_rays(s, 4096, 7), axis_rays(s, 4096, 9)
    mixed = np.concatenate([rnd[:2048], axr[:2048]])[np.random.default_rng(1).permutation(4096)]
    for rays, label in ((rnd, "random"), (axr, "axis"), (mixed, "mixed")):
        for anyhit in (False, True):
            td, pd = wf.test_intersect(s, rays, anyhit, device=device)
            th, ph = wf.test_intersect(s, rays, anyhit, device=0)
            assert_bitwise(td, th, f"{name} {label} t anyhit={anyhit}")
            assert np.array_equal(pd, ph), f"{name} {label} prim anyhit={anyhit}"
