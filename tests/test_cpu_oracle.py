"""CPU tests: the oracle pinned against the reference's golden vectors and closed forms, the
product's CPU arm and host estimator against the oracle, and the C-ABI boundary.

Reference tests followed: tests/test_math_rng.cpp (RNG), tests/test_bvh.cpp and
tests/catch/scene_bvh.cpp (BVH invariants, brute force), tests/test_path_core.cpp:880-977 and
tests/catch/cpu_render.cpp (closed forms), tests/test_cpu_render.cpp (threads, variance).
"""
import json
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, assert_bitwise, bits

REF_SCENES = ["cornell", "divergence", "trianglehv", "furnace"]
BENCH_SCENES = [("cornell555", 320, 180), ("sphere_grid", 1280, 720), ("heightfield", 1280, 720)]
PUBLISHED_HASH = {"cornell": "2ec1deb9ff15147d", "divergence": "9f8c9de445c74762", "trianglehv": "a489bc1b27791a9d"}
# results/stats/stats.json "bvh" (nodes, leaves, max_depth, sah_cost)
PUBLISHED_BVH = {"cornell": (7, 4, 3, 7.47625002), "divergence": (563, 282, 10, 8.6194229),
                 "trianglehv": (67279, 33640, 18, 29.0322913)}


def fnv64(data):
    h = 14695981039346656037
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


# ---------------------------------------------------------------------------------------------
# boundary
# ---------------------------------------------------------------------------------------------
def test_library_exports_every_declared_symbol(wf):
    header = open(os.path.join(ROOT, "include", "wfamd.h")).read()
    declared = set(re.findall(r"\b(wf_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations found"
    lib = wf.lib()
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, missing
    assert declared <= set(wf.EXPORTS) | {"wf_set_last_error"}


def test_invalid_inputs_raise(wf):
    with pytest.raises(wf.InvalidArgument):
        wf.Scene("no_such_scene", 16, 16)
    with pytest.raises(wf.InvalidArgument):
        wf.Scene("cornell", 0, 16)
    mat = np.zeros((1, 12), np.float32)
    sph = np.zeros((1, 8), np.float32)
    sph[0, :4] = [0, 0, 0, 1]
    sph[0, 4] = np.int32(5).view(np.float32)   # material index out of range
    with pytest.raises(wf.InvalidArgument):
        wf.Scene.from_arrays(sph, np.zeros((0, 3)), np.zeros((0, 4)), mat, np.zeros(20), np.array(
            [0, 0, 5, 0, 0, 0, 0, 1, 0, 40], np.float32), 8, 8)


def test_scene_view_equals_a_fresh_orbited_scene(wf):
    # A job view (Scene.view: copied geometry, own size and orbited camera) is the scene that
    # make_scene + orbit would build, without rebuilding the BVH: same rig, same BVH, same image.
    base = wf.Scene("cornell555", 64, 36)
    for w, h, yaw in ((40, 30, 75.0), (23, 31, None), (64, 36, 359.5)):
        v = base.view(w, h, yaw)
        fresh = wf.Scene("cornell555", w, h)
        if yaw is not None:
            fresh.orbit(yaw)
        assert v.info().width == w and v.info().height == h
        assert np.array_equal(v.camera_rig().view(np.uint32), fresh.camera_rig().view(np.uint32))
        for a, b in zip(v.bvh(), fresh.bvh()):
            assert np.array_equal(a, b)
        p = wf.params(spp=2, threads=4)
        (a, _, _), _ = wf.render_cpu_sums(v, 0, 2, p)
        (b, _, _), _ = wf.render_cpu_sums(fresh, 0, 2, p)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    with pytest.raises(wf.InvalidArgument):
        base.view(0, 10)
    with pytest.raises(wf.InvalidArgument):
        base.view(10, 9000)


def test_scene_info_is_cached_and_follows_orbit(wf):
    # The handle caches its info (BVH statistics and hash walk the whole scene); an orbit must
    # change the cached hash exactly as a freshly orbited scene's.
    s = wf.Scene("cornell555", 32, 18)
    h0 = s.info().hash
    assert s.info().hash == h0
    s.orbit(40.0)
    fresh = wf.Scene("cornell555", 32, 18)
    fresh.orbit(40.0)
    assert s.info().hash != h0
    assert s.info().hash == fresh.info().hash
    assert s.info().n_nodes == fresh.info().n_nodes and s.info().bvh_sah_cost == fresh.info().bvh_sah_cost


# ---------------------------------------------------------------------------------------------
# RNG and trigonometry
# ---------------------------------------------------------------------------------------------
def test_oracle_rng_matches_reference_golden(po):
    g = json.load(open(os.path.join(GOLDEN, "rng.json")))
    for x, want in g["hash_u32"].items():
        assert po.hash_u32(int(x)) == want
    assert po.hash_u32(0) == 0   # tests/test_math_rng.cpp:493
    for seed, pixel, sample, k0, k1 in g["path_key"]:
        assert po.path_key(seed, pixel, sample) == (k0, k1)
    for k0, k1, depth, dim, want in g["rng"]:
        assert int(np.float32(po.rng(k0, k1, depth, dim)).view(np.uint32)) == want


def test_hash_is_a_bijection_on_a_million_inputs(po):
    # tests/test_math_rng.cpp:488-495: lowbias32 is invertible, so distinct inputs stay distinct.
    import ctypes
    f = po.oracle().or_hash_u32
    out = {f(x) for x in range(0, 1 << 20)}
    assert len(out) == 1 << 20


def test_sincos_equals_correctly_rounded_on_all_2pow24_estimator_angles(wf, po):
    # The estimator only evaluates sin/cos at phi = (2*pi) * u with u = k * 2^-24 (binary32 product).
    k = np.arange(1 << 24, dtype=np.uint32)
    u = k.astype(np.float32) * np.float32(1.0 / 16777216.0)
    phi = np.float32(2.0 * np.float32(np.pi)) * u
    assert phi.dtype == np.float32
    s, c = wf.test_sincos(phi)
    ls, lc = po.sincos(phi)   # C library binary64 sin/cos rounded to binary32
    bad = np.flatnonzero((bits(s) != bits(ls)) | (bits(c) != bits(lc)))
    assert bad.size == 0, f"{bad.size} angles differ, first {phi[bad[:4]]}"
    # Same evaluation order through the grid hook (what the device test compares against).
    assert wf.test_sincos_grid_hash(0, 1 << 12) == wf.test_sincos_grid_hash(0, 1 << 12)


def test_sincos_closed_form_values(wf):
    x = np.array([0.0, np.pi / 6, np.pi / 4, np.pi / 2, np.pi, 1.5 * np.pi], np.float32)
    s, c = wf.test_sincos(x)
    np.testing.assert_allclose(s, np.sin(x.astype(np.float64)), atol=1e-7)
    np.testing.assert_allclose(c, np.cos(x.astype(np.float64)), atol=1e-7)
    assert s[0] == 0.0 and c[0] == 1.0


# ---------------------------------------------------------------------------------------------
# scenes and BVH
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", REF_SCENES)
def test_reference_scenes_hash_bvh_and_rig_match_golden(wf, po, name):
    g = json.load(open(os.path.join(GOLDEN, "scenes.json")))[name]
    s = wf.Scene(name, 1280, 720)
    info = s.info()
    assert "%016x" % info.hash == g["hash"]
    if name in PUBLISHED_HASH:
        assert "%016x" % info.hash == PUBLISHED_HASH[name]
    nodes, order = s.bvh()
    assert len(nodes) == g["nodes"]
    assert "%016x" % fnv64(np.ascontiguousarray(nodes[:, :9]).tobytes()) == g["nodes_fnv"]
    assert "%016x" % fnv64(np.ascontiguousarray(order).tobytes()) == g["prim_order_fnv"]
    assert [int(v) for v in bits(s.camera_rig())] == g["camera_rig_bits"]
    o = po.OScene(s.arrays())
    assert "%016x" % o.hash() == g["hash"]
    onodes, oorder = o.bvh()
    assert np.array_equal(onodes[:, :9], nodes[:, :9]) and np.array_equal(oorder, order)
    assert [int(v) for v in bits(o.camera_rig())] == g["camera_rig_bits"]


@pytest.mark.parametrize("name", list(PUBLISHED_BVH))
def test_bvh_statistics_match_published(wf, name):
    info = wf.Scene(name, 1280, 720).info()
    n, leaves, depth, sah = PUBLISHED_BVH[name]
    assert (info.n_nodes, info.bvh_leaves, info.bvh_max_depth) == (n, leaves, depth)
    assert abs(info.bvh_sah_cost - sah) < 1e-5 * sah


def check_bvh_invariants(nodes, order, n_prims):
    # tests/test_bvh.cpp:81-134: preorder, forward skips, full binary, leaf <= 4, coverage.
    assert sorted(order.tolist()) == list(range(n_prims))
    nn = len(nodes)
    skip, start, count = nodes[:, 3], nodes[:, 7], nodes[:, 8]
    lo, hi = nodes[:, 0:3].view(np.float32), nodes[:, 4:7].view(np.float32)
    assert skip[-1] == -1
    assert ((skip == -1) | (skip > np.arange(nn))).all()
    leaf = count > 0
    assert (count[leaf] <= 4).all() and (start[~leaf] == -1).all()
    internal = np.flatnonzero(~leaf)
    right = skip[internal + 1]
    assert (right > internal + 1).all() and (skip[right] == skip[internal]).all()
    for child in (internal + 1, right):
        assert (lo[internal] <= lo[child] + 1e-5).all() and (hi[child] <= hi[internal] + 1e-5).all()


@pytest.mark.parametrize("name,w,h", BENCH_SCENES)
def test_benchmark_scenes_shape_and_bvh(wf, po, name, w, h):
    s = wf.Scene(name, w, h)
    info = s.info()
    a = s.arrays()
    nodes, order = s.bvh()
    check_bvh_invariants(nodes, order, info.n_spheres + info.n_triangles)
    o = po.OScene(a)
    onodes, oorder = o.bvh()
    assert np.array_equal(onodes[:, :9], nodes[:, :9]) and np.array_equal(oorder, order)
    assert "%016x" % o.hash() == "%016x" % info.hash
    mats = a["materials"]
    mtype = mats[:, 0].view(np.int32)
    if name == "cornell555":
        assert info.n_triangles == 36 and info.n_spheres == 0 and info.n_lights == 2
        np.testing.assert_array_equal(mats[:3, 4:7], np.float32([[.73, .73, .73], [.65, .05, .05], [.12, .45, .15]]))
        assert (mats[3, 8:11] == 15).all()
        v = a["vertices"]
        assert v.min() >= -1e-3 and v.max() <= 555 + 1e-3
    elif name == "sphere_grid":
        assert info.n_spheres == 1024 and info.n_triangles == 2 and info.n_lights == 0
        r = a["spheres"][:, 3]
        assert r.min() >= 0.2 and r.max() <= 0.4
        m = a["spheres"][:, 4].view(np.int32)
        types = mtype[m]
        counts = np.bincount(types, minlength=3)
        assert counts.sum() == 1024 and counts.min() > 280          # uniform over three types
        fuzz = mats[m[types == 1], 2]
        assert fuzz.min() >= 0 and fuzz.max() <= 0.3
        assert (mats[m[types == 2], 1] == np.float32(1.5)).all()
        assert a["env"][3].view(np.int32) == 1 and a["env"][7].view(np.int32) == 1   # sky gradient, NEE
    else:
        assert info.n_vertices == 1024 * 1024 and info.n_triangles == 2 * 1023 * 1023 == 2093058
        y = a["vertices"][:, 1]
        assert 0.1 < y.max() - y.min() < 1.5
        assert a["env"][11].view(np.int32) == 1   # sun


# ---------------------------------------------------------------------------------------------
# traversal
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", REF_SCENES)
def test_traversal_matches_reference_golden(wf, po, name):
    g = np.load(os.path.join(GOLDEN, "intersect.npz"))
    rays, t, prim = g[name + "_rays"], g[name + "_t"], g[name + "_prim"]
    s = wf.Scene(name, 1280, 720)
    pt, pp = wf.test_intersect(s, rays)
    assert_bitwise(pt, t, "product t")
    assert np.array_equal(pp, prim)
    o = po.OScene(s.arrays())
    ot, op = o.intersect(rays)
    assert_bitwise(ot, t, "oracle t")
    assert np.array_equal(op, prim)
    bt, bp = o.intersect_brute(rays)        # brute force agrees wherever there is no exact tie
    hit = prim >= 0
    assert np.array_equal(bp >= 0, hit)
    assert_bitwise(bt[hit], t[hit], "brute t")


@pytest.mark.parametrize("name,w,h", BENCH_SCENES[:2])
def test_anyhit_agrees_with_closest_hit_on_occlusion(wf, name, w, h):
    s = wf.Scene(name, w, h)
    rs = np.random.default_rng(3)
    a = s.arrays()
    pts = a["vertices"] if len(a["vertices"]) > 4 else a["spheres"][:, :3]
    c = pts.mean(axis=0)
    rays = np.zeros((2000, 8), np.float32)
    rays[:, 0:3] = c + rs.uniform(-5, 5, (2000, 3))
    d = rs.normal(size=(2000, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = 1e-4
    rays[:, 7] = rs.uniform(0.5, 50, 2000)
    _, p = wf.test_intersect(s, rays, anyhit=False)
    ta, pa = wf.test_intersect(s, rays, anyhit=True)
    assert np.array_equal(p >= 0, pa >= 0)


# ---------------------------------------------------------------------------------------------
# renders: product CPU arm against the oracle, the oracle against the reference
# ---------------------------------------------------------------------------------------------
ALL_SMALL = [(n, 48, 27) for n in REF_SCENES] + [("cornell555", 48, 27), ("sphere_grid", 64, 36),
                                                  ("heightfield", 32, 18)]


@pytest.mark.parametrize("name,w,h", ALL_SMALL)
def test_cpu_arm_equals_oracle_bitwise(wf, po, name, w, h):
    s = wf.Scene(name, w, h)
    spp = 3
    (rgb, lum, lum2), st = wf.render_cpu_sums(s, 0, spp, wf.params(spp=spp, threads=4))
    o = po.OScene(s.arrays())
    (orgb, olum, olum2), cnt = o.render_sums(0, spp, 1234, 8, threads=4)
    assert_bitwise(rgb, orgb, "rgb")
    assert_bitwise(lum, olum, "lum")
    assert_bitwise(lum2, olum2, "lum2")
    assert (st.extension_rays, st.shadow_rays, st.invalid_hits) == cnt


@pytest.mark.parametrize("name,w,h", ALL_SMALL)
def test_oracle_wavefront_restatement_equals_oracle_render(po, wf, name, w, h):
    o = po.OScene(wf.Scene(name, w, h).arrays())
    sums = None
    total = [0, 0, 0]
    for sample in range(2):
        trace, sums, cnt = o.wavefront_pass(sample, 1234, 8, sums=sums)
        total = [a + b for a, b in zip(total, cnt)]
        assert trace[0][0] == w * h
        for b in range(len(trace) - 1):
            assert trace[b][5] == trace[b + 1][0]
        for rec in trace:
            active, ql, qm, qd, qs, nxt, _ = rec
            assert ql + qm + qd <= active and qs <= ql and nxt <= ql + qm + qd
        assert sum(r[0] for r in trace) == cnt[0]
    rsums, rcnt = o.render_sums(0, 2, 1234, 8)
    for a, b in zip(sums, rsums):
        assert_bitwise(a, b, "wavefront vs render")
    assert tuple(total) == rcnt


GOLDEN_W, GOLDEN_H = 64, 36   # tests/golden/make_golden.py
GOLDEN_PATHS = np.array([((k * 37) % GOLDEN_W, (k * 11) % GOLDEN_H, k % 7) for k in range(256)], np.int32)


@pytest.mark.parametrize("name", REF_SCENES)
def test_libm_trig_oracle_render_equals_reference_render_bitwise(wf, po, name):
    # 64x36, 4 spp, seed 1234, depth 8, written by the reference's own render_cpu. The oracle built
    # with the C library's binary32 sinf/cosf (the reference's trigonometry) gives every bit.
    want = np.load(os.path.join(GOLDEN, "ref_render.npz"))[name]
    o = po.OScene(wf.Scene(name, GOLDEN_W, GOLDEN_H).arrays(), libm_trig=True)
    rgb, var, _ = o.render(4, 1234, 8, threads=4)
    assert_bitwise(rgb, want, f"{name} libm-trig oracle vs reference render_cpu")


@pytest.mark.parametrize("name", REF_SCENES)
def test_libm_trig_oracle_single_paths_equal_reference_bitwise(wf, po, name):
    want = np.load(os.path.join(GOLDEN, "ref_paths.npz"))[name]
    o = po.OScene(wf.Scene(name, GOLDEN_W, GOLDEN_H).arrays(), libm_trig=True)
    L, _ = o.trace_paths(GOLDEN_PATHS, 1234, 8)
    assert_bitwise(L, want, f"{name} libm-trig oracle vs reference trace_path")


@pytest.mark.parametrize("name", REF_SCENES)
def test_oracle_and_product_differ_from_the_reference_only_through_trigonometry(wf, po, name):
    """The default oracle (binary64 sin/cos rounded once, which the product equals bit for bit)
    against the reference: the same oracle with the reference's binary32 sinf/cosf gives every bit
    (the tests above), so the pixels that differ here differ through trigonometry alone."""
    want = np.load(os.path.join(GOLDEN, "ref_render.npz"))[name]
    s = wf.Scene(name, GOLDEN_W, GOLDEN_H)
    o = po.OScene(s.arrays())
    rgb, var, _ = o.render(4, 1234, 8, threads=4)
    prod, pvar, _ = wf.render_cpu(s, wf.params(spp=4, threads=4))
    assert_bitwise(prod, rgb, f"{name} product CPU arm vs oracle")
    same = np.all(bits(rgb).reshape(-1, 3) == bits(want).reshape(-1, 3), axis=1)
    print(f"{name}: {same.mean():.4f} of pixels bitwise equal to the reference")
    assert same.mean() >= 0.98
    L, _ = o.trace_paths(GOLDEN_PATHS, 1234, 8)
    Lr = np.load(os.path.join(GOLDEN, "ref_paths.npz"))[name]
    assert np.all(bits(L) == bits(Lr), axis=1).mean() >= 0.97
    Lp, _ = wf.test_trace_paths(s, GOLDEN_PATHS)
    assert_bitwise(Lp, L, f"{name} product single paths vs oracle")


def test_cpu_arm_thread_count_does_not_change_the_image(wf):
    s = wf.Scene("divergence", 40, 30)
    a, _, _ = wf.render_cpu(s, wf.params(spp=3, threads=1))
    b, _, _ = wf.render_cpu(s, wf.params(spp=3, threads=7))
    assert_bitwise(a, b, "threads")


def test_cpu_progressive_equals_one_shot(wf):
    s = wf.Scene("cornell555", 32, 18)
    p = wf.params(spp=5, threads=4)
    one, _ = wf.render_cpu_sums(s, 0, 5, p)
    sums, _ = wf.render_cpu_sums(s, 0, 2, p)
    sums, _ = wf.render_cpu_sums(s, 2, 3, p, sums=sums)
    for a, b in zip(one, sums):
        assert_bitwise(a, b, "progressive")


def test_finalize_formula(wf, po):
    rs = np.random.default_rng(1)
    n, spp = 1000, 7
    rgb = rs.random(3 * n).astype(np.float32) * spp
    lum = rs.random(n).astype(np.float32) * spp
    lum2 = (lum * lum / spp + rs.random(n).astype(np.float32)).astype(np.float32)
    a, va = wf.finalize(rgb, lum, lum2, spp)
    b, vb = po.finalize(rgb, lum, lum2, spp)
    assert_bitwise(a, b, "mean")
    assert_bitwise(va, vb, "var")
    assert_bitwise(a, rgb / np.float32(spp), "mean formula")
    want = np.maximum(0.0, (lum2.astype(np.float64) - lum.astype(np.float64) ** 2 / spp) / (spp - 1))
    np.testing.assert_allclose(va, want, rtol=1e-6)
    _, v1 = wf.finalize(rgb, lum, lum2, 1)
    assert not v1.any()


# ---------------------------------------------------------------------------------------------
# closed forms (tests/test_path_core.cpp:880-977); these pin the estimator, including the
# extensions (sun, sky gradient, metal fuzz), independently of any implementation
# ---------------------------------------------------------------------------------------------
def one_sphere(wf, mtype, albedo, emission=(0, 0, 0), env=(0, 0, 0), ior=1.5, fuzz=0.0, camera_inside=False,
               size=16, radius=1.0, extra_env=None):
    sph = np.zeros((1, 8), np.float32)
    sph[0, :4] = [0, 0, 0 if camera_inside else -4, 5.0 if camera_inside else radius]
    mat = np.zeros((1, 12), np.float32)
    mat[0, 0] = np.int32(mtype).view(np.float32)
    mat[0, 1], mat[0, 2] = ior, fuzz
    mat[0, 4:7] = albedo
    mat[0, 8:11] = emission
    e = np.zeros(20, np.float32)
    e[0:3] = env
    if extra_env:
        for k, v in extra_env.items():
            e[k] = v
    cam = np.array([0, 0, 0, 0, 0, -1, 0, 1, 0, 90 if camera_inside else 30], np.float32)
    return wf.Scene.from_arrays(sph, np.zeros((0, 3)), np.zeros((0, 4)), mat, e, cam, size, size)


def mean_se(rgb, var, spp):
    lum = 0.2126 * rgb[0::3] + 0.7152 * rgb[1::3] + 0.0722 * rgb[2::3]
    return lum.mean(), np.sqrt((var / spp).mean() / lum.size + 1e-30)


def test_furnace_converges_to_le_over_one_minus_rho(wf):
    s = one_sphere(wf, 0, (0.5, 0.5, 0.5), emission=(1, 1, 1), camera_inside=True, size=32)
    rgb, var, _ = wf.render_cpu(s, wf.params(spp=64, max_depth=64, threads=8))
    m, se = mean_se(rgb, var, 64)
    assert abs(m - 2.0) <= 4.5 * se + 1e-4 and se < 0.01


@pytest.mark.parametrize("k,expect", [(0, 1.0), (1, 1.5), (2, 1.75)])
def test_furnace_truncated_series(wf, k, expect):
    s = one_sphere(wf, 0, (0.5, 0.5, 0.5), emission=(1, 1, 1), camera_inside=True, size=32)
    rgb, var, _ = wf.render_cpu(s, wf.params(spp=64, max_depth=k, threads=8))
    m, se = mean_se(rgb, var, 64)
    assert abs(m - expect) <= 4.5 * se + 1e-5


def test_black_furnace_returns_exactly_the_emission(wf):
    s = one_sphere(wf, 0, (0.0, 0.0, 0.0), emission=(1, 1, 1), camera_inside=True)
    rgb, _, _ = wf.render_cpu(s, wf.params(spp=4, max_depth=16, threads=4))
    np.testing.assert_allclose(rgb, 1.0, atol=1e-6)


def all_paths(size, spp):
    return np.array([(x, y, k) for y in range(size) for x in range(size) for k in range(spp)], np.int32)


def test_convex_lambert_sphere_under_constant_env(wf):
    # Outside a convex object every scattered ray escapes: each path returns env (missed the
    # sphere) or rho * env (tests/test_path_core.cpp:935-950).
    s = one_sphere(wf, 0, (0.5, 0.4, 0.3), env=(1, 1, 1), size=16)
    L, _ = wf.test_trace_paths(s, all_paths(16, 4))
    hit = np.abs(L[:, 0] - 1.0) > 1e-6
    assert hit.sum() > 50
    np.testing.assert_allclose(L[hit], np.broadcast_to(np.float32([0.5, 0.4, 0.3]), L[hit].shape), atol=1e-6)
    L0, _ = wf.test_trace_paths(s, all_paths(16, 1), max_depth=0)   # depth 0: nothing scatters
    assert not L0[np.abs(L0[:, 0] - 1.0) > 1e-6].any()


def test_convex_mirror_and_fuzzed_metal(wf):
    s = one_sphere(wf, 1, (0.9, 0.8, 0.7), env=(1, 2, 3), size=16)
    L, _ = wf.test_trace_paths(s, all_paths(16, 4))
    hit = np.abs(L[:, 0] - 1.0) > 1e-6
    assert hit.sum() > 50
    np.testing.assert_allclose(L[hit], np.broadcast_to(np.float32([0.9, 1.6, 2.1]), L[hit].shape), atol=1e-5)
    # Fuzz: a path that hits either escapes (tint * env) or is absorbed below the surface (0);
    # grazing hits get absorbed, so fuzz > 0 absorbs some paths.
    f = one_sphere(wf, 1, (0.9, 0.8, 0.7), env=(1, 2, 3), size=16, fuzz=0.3)
    Lf, _ = wf.test_trace_paths(f, all_paths(16, 8))
    fh = np.abs(Lf[:, 0] - 1.0) > 1e-6
    on_tint = np.all(np.abs(Lf[fh] - np.float32([0.9, 1.6, 2.1])) < 1e-5, axis=1)
    absorbed = np.all(Lf[fh] == 0, axis=1)
    assert (on_tint | absorbed).all()
    assert absorbed.sum() > 0 and on_tint.sum() > absorbed.sum()


def test_glass_sphere_conserves_energy(wf):
    s = one_sphere(wf, 2, (1, 1, 1), env=(1, 1, 1), size=16, radius=1.5)
    rgb, var, _ = wf.render_cpu(s, wf.params(spp=64, max_depth=64, threads=8))
    m, se = mean_se(rgb, var, 64)
    assert abs(m - 1.0) <= 4.5 * se + 1e-4


def test_sun_on_a_lambert_sphere_gives_the_cosine_law(wf):
    # Sun only (no sky light, black env): every camera hit gets exactly rho * E * cos / pi from NEE;
    # BSDF-sampled rays escape into black.
    d = np.array([0.0, 0.6, 0.8], np.float32)
    s = one_sphere(wf, 0, (0.5, 0.5, 0.5), size=24, radius=1.0,
                   extra_env={11: np.int32(1).view(np.float32), 12: d[0], 13: d[1], 14: d[2], 16: 2.0, 17: 2.0,
                              18: 2.0})
    rgb, _, _ = wf.render_cpu(s, wf.params(spp=2, max_depth=1, threads=4))
    px = rgb.reshape(-1, 3)[:, 0]
    assert px.max() > 0.1
    assert px.max() <= 0.5 * 2.0 / np.pi * 1.0001   # cos <= 1


def test_sky_gradient_seen_directly(wf):
    # A camera ray that escapes returns exactly sky(d) = lo * (1 - t) + hi * t, t = 0.5 (d.y + 1).
    ex = {3: np.int32(1).view(np.float32), 4: 1.0, 5: 1.0, 6: 1.0, 8: 0.5, 9: 0.7, 10: 1.0}
    s = one_sphere(wf, 0, (0.5, 0.5, 0.5), size=8, radius=1e-3, extra_env=ex)
    rgb, _, _ = wf.render_cpu(s, wf.params(spp=1, threads=1))
    px = rgb.reshape(-1, 3)
    assert ((px[:, 0] >= 0.5 - 1e-6) & (px[:, 0] <= 1.0 + 1e-6)).all()
    assert np.allclose(px[:, 2], 1.0)


def test_sky_light_nee_is_unbiased_for_a_convex_lambert_sphere(wf):
    # Constant sky used as a light: E[L] = rho * env on the sphere (MIS of NEE and BSDF sampling).
    # The sphere (radius 3 at distance 4, vfov 30) covers every pixel.
    ex = {7: np.int32(1).view(np.float32)}
    s = one_sphere(wf, 0, (0.5, 0.5, 0.5), env=(1, 1, 1), size=24, radius=3.0, extra_env=ex)
    rgb, var, _ = wf.render_cpu(s, wf.params(spp=64, max_depth=4, threads=8))
    m, se = mean_se(rgb, var, 64)
    assert abs(m - 0.5) <= 4.5 * se + 1e-4 and se < 0.005


# ---------------------------------------------------------------------------------------------
# heightfield parameters (6-octave fBm, amplitude 0.25, lacunarity 2, gain 0.5)
# ---------------------------------------------------------------------------------------------
def _lowbias32(x):
    x = np.asarray(x, np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def _grad_noise(x, z, seed):
    """2D gradient noise: a unit gradient per lattice corner from hashing the corner and the
    seed, dotted with the offset to the corner, blended by the quintic fade; scaled by sqrt(2)."""
    fx0, fz0 = np.floor(x), np.floor(z)
    ix, iz = fx0.astype(np.int64), fz0.astype(np.int64)
    fx, fz = x - fx0, z - fz0
    g = []
    for dx, dz in ((0, 0), (1, 0), (0, 1), (1, 1)):
        cx = ((ix + dx) & 0xFFFFFFFF) * 0x8DA6B343 & 0xFFFFFFFF
        cz = ((iz + dz) & 0xFFFFFFFF) * 0xD8163841 & 0xFFFFFFFF
        h = _lowbias32(_lowbias32(cx ^ cz) ^ np.uint64(seed & 0xFFFFFFFF))
        a = (h >> 8).astype(np.float64) * (1.0 / 16777216.0) * 6.283185307179586
        g.append(np.cos(a) * (fx - dx) + np.sin(a) * (fz - dz))
    fade = lambda t: t * t * t * (t * (t * 6.0 - 15.0) + 10.0)
    u, v = fade(fx), fade(fz)
    a = g[0] + (g[1] - g[0]) * u
    b = g[2] + (g[3] - g[2]) * u
    return (a + (b - a) * v) * 1.4142135623730951


def _fbm(x, z, octaves=6, amp=0.25, lacunarity=2.0, gain=0.5, f0=0.4, seed=7):
    h = np.zeros_like(x)
    f = f0
    for o in range(octaves):
        h += amp * _grad_noise(x * f, z * f, (seed + o * 0x9E3779B9) & 0xFFFFFFFF)
        amp *= gain
        f *= lacunarity
    return h


def test_heightfield_heights_are_the_stated_fbm(wf, po):
    """Vertex heights equal 6-octave fBm (amplitude 0.25, lacunarity 2, gain 0.5, base frequency
    0.4, seed 7) of gradient noise, evaluated here in numpy on a sample of vertices; each parameter
    changed by a little moves the heights, so the comparison pins all of them. The corners of
    [-5, 5]^2 sit on the lattice of every octave, where gradient noise vanishes."""
    for v in (0, 1, 12345, 0xFFFFFFFF):
        assert int(_lowbias32(v)) == po.hash_u32(v)
    s = wf.Scene("heightfield", 64, 36)
    vert = s.arrays()["vertices"].reshape(1024, 1024, 3)
    rng = np.random.default_rng(5)
    i, j = rng.integers(0, 1024, 20000), rng.integers(0, 1024, 20000)
    x = -5.0 + 10.0 * i / 1023.0
    z = -5.0 + 10.0 * j / 1023.0
    assert np.array_equal(vert[i, j, 0], x.astype(np.float32)) and np.array_equal(vert[i, j, 2], z.astype(np.float32))
    y = vert[i, j, 1]
    want = _fbm(x, z).astype(np.float32)
    assert np.array_equal(y, want), int((y != want).sum())
    for k in ((0, 0), (0, 1023), (1023, 0), (1023, 1023)):
        assert vert[k[0], k[1], 1] == 0.0
    for change in ({"octaves": 5}, {"octaves": 7}, {"amp": 0.26}, {"lacunarity": 2.01}, {"gain": 0.51},
                   {"f0": 0.41}, {"seed": 8}):
        other = _fbm(x, z, **change).astype(np.float32)
        assert (other != y).mean() > 0.5, change
