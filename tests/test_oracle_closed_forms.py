"""Closed forms that pin the oracle's estimator extensions (sun, sky gradient, sky next-event
estimation, metal fuzz), which have no counterpart in the reference and so no golden vectors.

Each check is computed independently of the oracle's code: camera rays come from the camera rig
and the counter RNG (both pinned bit for bit by the reference's own golden vectors), geometry is
intersected analytically in binary64, and the expected radiance is the textbook formula. A
plausible mistake fails one of them: a dropped cosine or 1/pi, a sun direction of the wrong sign,
a gradient on d.x instead of d.y, a missing or doubled MIS weight, a wrong fuzz scale or a fuzz
vector drawn inside the ball instead of on the sphere.
Reference test followed for the approach: tests/test_path_core.cpp:880-977 (closed forms of the
whole-path radiance).
"""
import numpy as np
import pytest

# ---------------------------------------------------------------------------------------------
# scene arrays (the C ABI layout of include/wfamd.h, wf_scene_desc) and exact camera rays
# ---------------------------------------------------------------------------------------------
LAMBERT, METAL = 0, 1


def scene_arrays(spheres, materials, env, camera, w, h):
    sph = np.zeros((len(spheres), 8), np.float32)
    for i, (c, r, m) in enumerate(spheres):
        sph[i, :3], sph[i, 3] = c, r
        sph[i, 4] = np.int32(m).view(np.float32)
    mat = np.zeros((len(materials), 12), np.float32)
    for i, (t, albedo, fuzz) in enumerate(materials):
        mat[i, 0] = np.int32(t).view(np.float32)
        mat[i, 1], mat[i, 2] = 1.5, fuzz
        mat[i, 4:7] = albedo
    return dict(spheres=sph, vertices=np.zeros((0, 3), np.float32), triangles=np.zeros((0, 4), np.int32),
                materials=mat, env=np.asarray(env, np.float32), camera=np.asarray(camera, np.float32),
                width=w, height=h)


def env20(env=(0, 0, 0), sky_mode=0, lo=(0, 0, 0), hi=(0, 0, 0), sky_light=0, sun_light=0, sun_dir=(0, 1, 0),
          sun_E=(0, 0, 0)):
    e = np.zeros(20, np.float32)
    e[0:3] = env
    e[3] = np.int32(sky_mode).view(np.float32)
    e[4:7] = lo
    e[7] = np.int32(sky_light).view(np.float32)
    e[8:11] = hi
    e[11] = np.int32(sun_light).view(np.float32)
    e[12:15] = sun_dir
    e[16:19] = sun_E
    return e


def lowbias32(x):
    """hash_u32 (include/wf/rng.hpp:8-13) on uint32 arrays; checked against the pinned oracle below."""
    x = np.asarray(x, np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def rng_dim(k0, k1, depth, dim):
    ctr = np.uint32(depth * 16 + dim)
    h = lowbias32(k0 ^ lowbias32(ctr * np.uint32(0x9E3779B9) + k1))
    return (h >> np.uint32(8)).astype(np.float64) * 2.0 ** -24


def camera_rays(po, o, w, h, spp, seed=1234):
    """Origins and unit directions (binary64) of every camera ray of an image, sample-major."""
    rig = o.camera_rig().astype(np.float64).reshape(4, 3)
    origin, ll, hor, ver = rig
    py, px = np.mgrid[0:h, 0:w]
    px, py = px.ravel(), py.ravel()
    keys = np.array([[po.path_key(seed, int(y * w + x), s) for x, y in zip(px, py)] for s in range(spp)],
                    np.uint32)                              # spp x pixels x 2
    k0, k1 = keys[..., 0], keys[..., 1]
    # The vectorised RNG must give the pinned oracle's bits.
    for i in range(0, k0.size, max(1, k0.size // 7)):
        a, b = k0.ravel()[i], k1.ravel()[i]
        assert np.float32(rng_dim(a, b, 0, 0)) == np.float32(po.rng(int(a), int(b), 0, 0))
    jx, jy = rng_dim(k0, k1, 0, 0), rng_dim(k0, k1, 0, 1)
    u = (px[None, :] + jx) / w
    v = ((h - 1 - py)[None, :] + jy) / h
    target = ll + u[..., None] * hor + v[..., None] * ver
    d = target - origin
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return origin, d                                        # d: spp x pixels x 3


def hit_sphere(o, d, c, r):
    """Nearest positive root (binary64) of |o + t d - c| = r, NaN for a miss; and |disc| margin."""
    oc = o - np.asarray(c, np.float64)
    b = (d * oc).sum(-1)
    disc = b * b - ((oc * oc).sum(-1) - r * r)
    t = np.where(disc >= 0, -b - np.sqrt(np.maximum(disc, 0)), np.nan)
    return np.where(t > 1e-4, t, np.nan), np.abs(disc)


# ---------------------------------------------------------------------------------------------
# sun (directional delta light): L = rho * E * max(0, n.l) / pi at every camera hit
# ---------------------------------------------------------------------------------------------
SUN_DIR = np.array([0.3, 0.6, 0.74], np.float64)
SUN_DIR /= np.linalg.norm(SUN_DIR)
ALBEDO = np.array([0.5, 0.4, 0.3])
SUN_E = np.array([2.0, 1.5, 1.0])


def check_sun_cosine_law(po, lib):
    w = h = 32
    cam = [0, 0, 0, 0, 0, -1, 0, 1, 0, 30]
    a = scene_arrays([((0, 0, -4), 1.0, 0)], [(LAMBERT, ALBEDO, 0.0)],
                     env20(sun_light=1, sun_dir=SUN_DIR, sun_E=SUN_E), cam, w, h)
    o = po.OScene(a, lib=lib)
    (rgb, _, _), cnt = o.render_sums(0, 1, 1234, 1)
    origin, d = camera_rays(po, o, w, h, 1)
    t, margin = hit_sphere(origin, d[0], (0, 0, -4), 1.0)
    hit = ~np.isnan(t)
    p = origin + t[:, None] * d[0]
    n = (p - np.array([0, 0, -4.0])) / 1.0
    cos = np.clip(n @ SUN_DIR, 0.0, None)
    want = np.where(hit[:, None], ALBEDO * SUN_E * cos[:, None] / np.pi, 0.0)
    got = rgb.reshape(-1, 3).astype(np.float64)
    clear = margin > 1e-4                      # not on the silhouette, where binary32 and 64 disagree
    assert hit.sum() > 100 and (cos[hit] > 0.5).sum() > 50
    peak = (ALBEDO * SUN_E).max() / np.pi
    err = np.abs(got - want)[clear]
    assert err.max() <= 1e-6 * peak, err.max() / peak
    # One shadow ray per lit hit (the sun is the only light), none for misses or unlit hits.
    assert cnt[1] == int((hit & (n @ SUN_DIR > 0)).sum()) or not clear.all()


@pytest.mark.parametrize("libm_trig", [False, True])
def test_sun_gives_the_cosine_law_per_pixel(po, libm_trig):
    check_sun_cosine_law(po, po.oracle(libm_trig))


def test_sun_is_blocked_by_an_occluder(po):
    check_sun_occluder(po, po.oracle())


def check_sun_occluder(po, lib):
    """A second sphere between the first one and the sun: hits in its shadow get exactly 0."""
    w = h = 32
    cam = [0, 0, 0, 0, 0, -1, 0, 1, 0, 30]
    blocker = np.array([0, 0, -4.0]) + 2.2 * SUN_DIR + np.array([0.3, 0.0, 0.0])
    a = scene_arrays([((0, 0, -4), 1.0, 0), (tuple(blocker), 0.6, 0)], [(LAMBERT, ALBEDO, 0.0)],
                     env20(sun_light=1, sun_dir=SUN_DIR, sun_E=SUN_E), cam, w, h)
    o = po.OScene(a, lib=lib)
    (rgb, _, _), _ = o.render_sums(0, 1, 1234, 1)
    origin, d = camera_rays(po, o, w, h, 1)
    t, margin = hit_sphere(origin, d[0], (0, 0, -4), 1.0)
    tb, _ = hit_sphere(origin, d[0], blocker, 0.6)
    first = ~np.isnan(t) & (np.isnan(tb) | (tb > t))
    p = origin + t[:, None] * d[0]
    n = p - np.array([0, 0, -4.0])
    org = p + n * 1e-4
    ts, smargin = hit_sphere(org, np.broadcast_to(SUN_DIR, org.shape), blocker, 0.6)
    lit = first & (n @ SUN_DIR > 0)
    shadowed = lit & ~np.isnan(ts)
    clear = (margin > 1e-4) & (smargin > 1e-4)
    got = rgb.reshape(-1, 3)
    assert (shadowed & clear).sum() > 20 and (lit & ~shadowed & clear).sum() > 20
    assert not got[shadowed & clear].any()
    cos = n[lit & ~shadowed & clear] @ SUN_DIR
    np.testing.assert_allclose(got[lit & ~shadowed & clear], ALBEDO * SUN_E * cos[:, None] / np.pi,
                               rtol=0, atol=1e-6 * (ALBEDO * SUN_E).max() / np.pi)


# ---------------------------------------------------------------------------------------------
# sky gradient seen directly: L = lo (1 - t) + hi t, t = (d.y + 1) / 2
# ---------------------------------------------------------------------------------------------
LO = np.array([1.0, 0.9, 0.8])
HI = np.array([0.2, 0.4, 1.0])


@pytest.mark.parametrize("look", [(0.0, 0.6, -1.0), (1.0, -0.4, 0.3)])
def test_sky_gradient_seen_directly_per_pixel(po, look):
    check_sky_direct(po, po.oracle(), look)


def check_sky_direct(po, lib, look):
    w, h = 40, 24
    cam = [0, 0, 0, look[0], look[1], look[2], 0, 1, 0, 90]
    # The only sphere is behind the camera: every camera ray escapes at depth 0 (specular, no MIS).
    a = scene_arrays([((-look[0] * 50, -look[1] * 50, -look[2] * 50), 0.5, 0)], [(LAMBERT, ALBEDO, 0.0)],
                     env20(sky_mode=1, lo=LO, hi=HI), cam, w, h)
    o = po.OScene(a, lib=lib)
    (rgb, _, _), cnt = o.render_sums(0, 2, 1234, 8)
    _, d = camera_rays(po, o, w, h, 2)
    t = 0.5 * (d[..., 1] + 1.0)
    want = (LO * (1 - t[..., None]) + HI * t[..., None]).sum(0)      # sum of the two samples
    assert np.ptp(d[..., 1]) > 0.5 and np.ptp(d[..., 0]) > 0.5
    np.testing.assert_allclose(rgb.reshape(-1, 3), want, rtol=0, atol=2e-6)
    assert cnt[0] == 2 * w * h and cnt[1] == 0


# ---------------------------------------------------------------------------------------------
# sky gradient as a light (next-event estimation + BSDF sampling with MIS) on a convex Lambert
# sphere: every scattered ray escapes, so E[L] = rho/pi * int_H (a + b w_y) cos dw
#                                             = rho (a + 2/3 b n_y),  a = (lo+hi)/2, b = (hi-lo)/2
# ---------------------------------------------------------------------------------------------
LUM = np.array([0.2126, 0.7152, 0.0722])


@pytest.mark.parametrize("max_depth", [1, 8])
def test_gradient_sky_nee_is_unbiased_on_a_convex_lambert_sphere(po, max_depth):
    check_sky_nee(po, po.oracle(), max_depth)


def check_sky_nee(po, lib, max_depth):
    w = h = 32
    spp = 192
    cam = [0, 0, 0, 0, 0, -1, 0, 1, 0, 40]
    c, r = np.array([0, 0, -4.0]), 1.0            # wholly in view: n_y spans [-1, 1]
    a = scene_arrays([(tuple(c), r, 0)], [(LAMBERT, ALBEDO, 0.0)],
                     env20(sky_mode=1, lo=LO, hi=HI, sky_light=1), cam, w, h)
    o = po.OScene(a, lib=lib)
    (rgb, lum, lum2), cnt = o.render_sums(0, spp, 1234, max_depth, threads=8)
    mean, var = po.finalize(rgb, lum, lum2, spp)
    origin, d = camera_rays(po, o, w, h, spp)
    t, _ = hit_sphere(origin, d, c, r)
    hit = ~np.isnan(t)
    nrm = (origin + np.nan_to_num(t)[..., None] * d - c) / r
    A, B = (LO + HI) / 2, (HI - LO) / 2
    sky = LO * (1 - 0.5 * (d[..., 1:2] + 1)) + HI * 0.5 * (d[..., 1:2] + 1)   # a camera ray that escapes

    def expect(n_axis):
        on = ALBEDO * (A + (2.0 / 3.0) * B * n_axis[..., None])
        return np.where(hit[..., None], on, sky).mean(0) @ LUM        # per pixel, luminance

    want_l, wrong_l = expect(nrm[..., 1]), expect(nrm[..., 0])
    got_l = mean.reshape(-1, 3) @ LUM
    se = np.sqrt(var / spp)
    # 4x4 pixel blocks: mean luminance within 4 standard errors of the closed form.
    blk = lambda v: v.reshape(h // 4, 4, w // 4, 4).mean(axis=(1, 3))
    blk_se = np.sqrt((se.reshape(h // 4, 4, w // 4, 4) ** 2).sum(axis=(1, 3))) / 16 + 1e-7
    z = (blk(got_l) - blk(want_l)) / blk_se
    assert np.abs(z).max() < 4.0, np.abs(z).max()
    z_all = (got_l.mean() - want_l.mean()) / (np.sqrt((se ** 2).sum()) / se.size)
    assert abs(z_all) < 4.0, z_all
    # Sensitivity: the same formula on n_x (a gradient on the wrong axis) is far outside.
    assert (np.abs(blk(got_l) - blk(wrong_l)) / blk_se).max() > 10
    # The sky is sampled uniformly over the sphere: about half of the samples face the surface and
    # trace a shadow ray.
    assert hit.mean() > 0.3 and 0.45 * hit.sum() < cnt[1] < 0.55 * hit.sum()


# ---------------------------------------------------------------------------------------------
# metal fuzz: w = normalize(r + f u), r the mirror direction, u uniform on the unit sphere;
# absorbed when (r + f u).n <= 0
# ---------------------------------------------------------------------------------------------
def stated_fuzz(d, n, f, m, seed):
    rs = np.random.default_rng(seed)
    u = rs.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = d - 2 * (d @ n) * n
    v = r + f * u
    alive = v @ n > 0
    return r, v[alive] / np.linalg.norm(v[alive], axis=1, keepdims=True), alive.mean()


@pytest.mark.parametrize("d,fuzz", [((0.0, -0.6, -0.8), 0.3), ((0.0, -0.6, -0.8), 0.05),
                                    ((0.98, -0.199, 0.0), 0.3)])
def test_fuzz_direction_follows_the_stated_distribution(po, d, fuzz):
    check_fuzz(po, po.oracle(), d, fuzz)


def check_fuzz(po, lib, d, fuzz):
    from scipy import stats
    d = np.asarray(d, np.float64)
    d /= np.linalg.norm(d)
    nrm = np.array([0.0, 1.0, 0.0])
    m = 40000
    keys = [po.path_key(77, i, 3) for i in range(m)]
    out = [po.metal_dir(d.astype(np.float32), nrm.astype(np.float32), fuzz, k0, k1, 2, lib=lib) for k0, k1 in keys]
    alive = np.array([v is not None for v in out])
    wv = np.array([v for v in out if v is not None], np.float64)
    r, wref, alive_ref = stated_fuzz(d, nrm, fuzz, 4 * m, 5)
    # Unit length, inside the cone of half-angle asin(f) about the mirror direction.
    np.testing.assert_allclose(np.linalg.norm(wv, axis=1), 1.0, atol=2e-6)
    cosang = wv @ r
    assert cosang.min() >= np.sqrt(1 - fuzz * fuzz) - 2e-6
    # Same distribution of the angle from r and of the azimuth about r as the stated one.
    t1 = np.cross(r, [0.0, 0.0, 1.0])
    if np.linalg.norm(t1) < 1e-3:
        t1 = np.cross(r, [1.0, 0.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(r, t1)
    az = lambda v: np.arctan2(v @ t2, v @ t1)
    assert stats.ks_2samp(cosang, wref @ r).pvalue > 1e-3
    assert stats.ks_2samp(az(wv), az(wref)).pvalue > 1e-3
    # Mean direction and absorbed fraction within 4 standard errors of the stated distribution.
    se = np.sqrt(wv.var(axis=0) / len(wv) + wref.var(axis=0) / len(wref)) + 1e-12
    assert np.all(np.abs(wv.mean(0) - wref.mean(0)) < 4 * se)
    p = alive_ref
    assert abs(alive.mean() - p) < 4 * np.sqrt(p * (1 - p) / m + p * (1 - p) / (4 * m)) + 1e-9
    if d[0] > 0.9:
        assert alive.mean() < 0.95          # grazing: some fuzzed directions point below the surface


def test_zero_fuzz_is_the_exact_mirror(po):
    rs = np.random.default_rng(3)
    for _ in range(200):
        n = rs.normal(size=3)
        n /= np.linalg.norm(n)
        d = rs.normal(size=3)
        d /= np.linalg.norm(d)
        if d @ n > 0:
            d = -d
        wv = po.metal_dir(d.astype(np.float32), n.astype(np.float32), 0.0, 1, 2)
        r = d - 2 * (d @ n) * n
        np.testing.assert_allclose(wv, r, atol=2e-6)
