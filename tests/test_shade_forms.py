"""The persistent engine's one-pass shading (shade_surface) against the per-material functions.

The persistent shade kernel runs shade_lambert, shade_metal and shade_dielectric
(include/wf/path_core.hpp:192-276) as one function whose shared steps (scatter draws, square root
and sin/cos of the scatter direction, normalisation, state update, roulette) are written once
with per-material operands, so a wave of mixed materials runs them once. Every lane must still
compute exactly what its own material's function computes: one classify + shade step per case,
all 34 output words (the path's next ray, throughput, radiance, pdf, depth, flags, the shadow
ray and its contribution) compared bit for bit, on the host and on the GPU.

Cases are random rays through each scene (their closest hit found on the host) with random path
state: depths on both sides of the roulette's start and of max_depth, specular and MIS flags,
throughputs below the roulette's clamp and above 1; then rays that start just inside the surface
each hit crossed (dielectric exits with total internal reflection, hits from behind)."""
import numpy as np
import pytest

from conftest import assert_bitwise

SCENES = ["cornell", "divergence", "trianglehv", "furnace", "cornell555", "sphere_grid", "heightfield"]
TRACE = np.uint32(0xFFFFFFFE)   # primitive word -2: find the closest hit on the host


def scene_box(s):
    a = s.arrays()
    pts = a["vertices"] if len(a["vertices"]) else a["spheres"][:, :3]
    if len(a["vertices"]) and len(a["spheres"]):
        pts = np.concatenate([pts, a["spheres"][:, :3]])
    c = pts.mean(axis=0)
    return c, float(np.abs(pts - c).max()) + 1.0


def state_words(q, rs):
    n = len(q)
    f = q.view(np.float32)
    thr = rs.uniform(0.0, 1.5, (n, 3))
    thr[rs.random(n) < 0.2] *= 0.02                    # below the roulette's 0.05 clamp
    f[:, 8:11] = thr
    f[:, 11:14] = rs.uniform(0.0, 2.0, (n, 3))
    f[:, 14] = np.where(rs.random(n) < 0.3, 0.0, rs.uniform(0.0, 3.0, n))
    q[:, 15] = rs.integers(0, 9, n)
    q[:, 16] = rs.integers(0, 2, n)
    q[:, 17] = rs.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    q[:, 18] = rs.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    q[:, 19] = rs.choice([1, 3, 8, 16], n)


def cases(wf, s, n, seed):
    """Random rays through the scene, then rays from just past each hit (inside the surface)."""
    rs = np.random.default_rng(seed)
    c, rad = scene_box(s)
    q = np.zeros((n, 20), np.uint32)
    f = q.view(np.float32)
    f[:, 0:3] = c + rs.uniform(-rad, rad, (n, 3))
    d = rs.normal(size=(n, 3))
    sph = s.arrays()["spheres"]
    if len(sph):   # half of the rays aimed at spheres (small next to the scenes' floors and walls)
        k = rs.integers(0, len(sph), n // 2)
        aim = sph[k, :3] + sph[k, 3:4] * rs.uniform(-1.0, 1.0, (len(k), 3))
        d[: n // 2] = aim - f[: n // 2, 0:3]
    f[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    q[:, 7] = TRACE
    state_words(q, rs)
    first = wf.test_shade(s, q, form=0)
    hit = first[:, 31].astype(np.int32) >= 0
    t = first[:, 32].view(np.float32)
    inner = np.zeros((int(hit.sum()), 20), np.uint32)
    g = inner.view(np.float32)
    g[:, 0:3] = f[hit, 0:3] + f[hit, 3:6] * (t[hit, None] + np.float32(1e-3))
    d = rs.normal(size=(len(inner), 3))
    g[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    inner[:, 7] = TRACE
    state_words(inner, rs)
    return np.concatenate([q, inner])


def check_coverage(outs):
    out = np.concatenate(outs)
    mat = out[:, 2].astype(np.int32)
    for m in (-1, 0, 1, 2):
        assert (mat == m).sum() > 50, f"material type {m}: {(mat == m).sum()} cases"
    assert (out[:, 1] == 1).sum() > 100                      # shadow rays
    assert ((out[:, 0] == 0) & (mat >= 0)).sum() > 100        # roulette and metal deaths


@pytest.fixture(scope="module")
def scenes(wf):
    return {name: wf.Scene(name, 64, 36) for name in SCENES}


def test_one_pass_shading_equals_per_material_functions_on_host(wf, scenes):
    outs = []
    for k, name in enumerate(SCENES):
        s = scenes[name]
        q = cases(wf, s, 6000, 100 + k)
        ref = wf.test_shade(s, q, form=0)
        one = wf.test_shade(s, q, form=1)
        assert_bitwise(one.view(np.float32), ref.view(np.float32), f"{name} one-pass vs per-material")
        outs.append(ref)
    check_coverage(outs)


def test_shade_cases_reach_every_branch(wf, scenes):
    # The dielectric exits from inside (f.inside) and the metal scatters below the surface are
    # present among the cases, so the bitwise comparison above covers them.
    s = scenes["sphere_grid"]
    q = cases(wf, s, 6000, 7)
    out = wf.test_shade(s, q, form=0)
    mat = out[:, 2].astype(np.int32)
    dead_metal = (mat == 1) & (out[:, 0] == 0)
    assert dead_metal.sum() > 10
    die = mat == 2
    assert die.sum() > 200 and np.all(out[die, 19] == 1)
    # paths that left a dielectric from inside: their hit was on a sphere they started in
    hit_t = out[:, 32].view(np.float32)
    assert (die & (hit_t < 1.0) & (np.arange(len(q)) >= 6000)).sum() > 50


def test_bad_cases_are_rejected(wf, scenes):
    s = scenes["cornell"]
    q = np.zeros((1, 20), np.uint32)
    q[0, 7] = np.uint32(1 << 20)
    with pytest.raises(Exception):
        wf.test_shade(s, q, form=0)
    with pytest.raises(Exception):
        wf.test_shade(s, np.zeros((1, 20), np.uint32), form=4)


@pytest.mark.gpu
@pytest.mark.parametrize("form", [2, 3])
def test_device_shading_equals_host(wf, scenes, form):
    for k, name in enumerate(SCENES):
        print(f"[shade forms] {name} form {form}", flush=True)
        s = scenes[name]
        q = cases(wf, s, 6000, 100 + k)
        ref = wf.test_shade(s, q, form=0)
        dev = wf.test_shade(s, q, form=form)
        assert_bitwise(dev.view(np.float32), ref.view(np.float32), f"{name} form {form} vs host")
