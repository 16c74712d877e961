"""Scene validation and the failure cases of the reference's wavefront tests, through the C ABI.

Followed: src/scene.cpp:10-56 (validate_scene: each rejected input and its message) and
tests/catch/wavefront_failure.cu:61-260 (a scene the renderer cannot take is refused with an error
and the process stays usable; a queue capacity of exactly width x height never overflows; an
unlit scene with no lights renders the megakernel's image with no shadow rays; the wavefront's
device memory at 1280x720 lies between its path pool and the 3.5 GB budget). The reference's
"unfinalized scene" cannot be built here: every scene handle is validated and finalized when it is
created, so the same inputs are refused at creation instead of at renderer construction.
"""
import numpy as np
import pytest

from conftest import assert_equals_oracle


def base_arrays():
    """The reference's hand-built unlit scene (wavefront_failure.cu:61-90): two non-emissive
    Lambert spheres (the ground is a radius-100 sphere) under a constant environment of 0.5."""
    sph = np.zeros((2, 8), np.float32)
    sph[0, :4] = [0, 0, -3, 1]
    sph[1, :4] = [0, -101, -3, 100]
    mat = np.zeros((1, 12), np.float32)
    mat[0, 1] = 1.0
    mat[0, 4:7] = 0.5
    env = np.zeros(20, np.float32)
    env[0:3] = 0.5
    cam = np.array([0, 0, 0, 0, 0, -3, 0, 1, 0, 50], np.float32)
    return dict(spheres=sph, vertices=np.zeros((0, 3), np.float32), triangles=np.zeros((0, 4), np.int32),
                materials=mat, env=env, camera=cam)


def build(wf, a, w=16, h=9):
    return wf.Scene.from_arrays(a["spheres"], a["vertices"], a["triangles"], a["materials"], a["env"], a["camera"],
                                w, h)


def tri_arrays():
    a = base_arrays()
    a["vertices"] = np.array([[-1, 0, -2], [1, 0, -2], [0, 1, -2]], np.float32)
    a["triangles"] = np.array([[0, 1, 2, 0]], np.int32)
    return a


def bits(i):
    return np.float32(np.int32(i).view(np.float32))


CASES = [
    ("no primitives", lambda a: a.update(spheres=np.zeros((0, 8), np.float32))),
    ("no materials", lambda a: a.update(materials=np.zeros((0, 12), np.float32))),
    ("sphere material index out of range", lambda a: a["spheres"].__setitem__((0, 4), np.int32(1).view(np.float32))),
    ("sphere material index out of range", lambda a: a["spheres"].__setitem__((1, 4), np.int32(-1).view(np.float32))),
    ("sphere has invalid radius", lambda a: a["spheres"].__setitem__((0, 3), 0.0)),
    ("sphere has invalid radius", lambda a: a["spheres"].__setitem__((0, 3), -1.0)),
    ("sphere has invalid radius", lambda a: a["spheres"].__setitem__((0, 3), np.nan)),
    ("sphere has invalid radius", lambda a: a["spheres"].__setitem__((0, 3), np.inf)),
    ("sphere centre is non-finite", lambda a: a["spheres"].__setitem__((0, 1), np.inf)),
    ("sphere centre is non-finite", lambda a: a["spheres"].__setitem__((1, 2), np.nan)),
]
TRI_CASES = [
    ("triangle material index out of range", lambda a: a["triangles"].__setitem__((0, 3), 1)),
    ("triangle material index out of range", lambda a: a["triangles"].__setitem__((0, 3), -2)),
    ("triangle vertex index out of range", lambda a: a["triangles"].__setitem__((0, 2), 3)),
    ("triangle vertex index out of range", lambda a: a["triangles"].__setitem__((0, 0), -1)),
    ("vertex is non-finite", lambda a: a["vertices"].__setitem__((1, 0), np.nan)),
    ("vertex is non-finite", lambda a: a["vertices"].__setitem__((2, 1), -np.inf)),
]


@pytest.mark.parametrize("why,mutate", CASES + TRI_CASES)
def test_invalid_scene_is_refused_with_its_reason(wf, why, mutate):
    a = tri_arrays() if (why, mutate) in TRI_CASES else base_arrays()
    build(wf, a)                                            # the unmodified scene is accepted
    mutate(a)
    with pytest.raises(wf.InvalidArgument, match=why):
        build(wf, a)
    s = build(wf, base_arrays())                            # the process is unharmed
    rgb, var, st = wf.render_cpu(s, wf.params(spp=1, threads=2))
    assert rgb.sum() > 0 and np.isfinite(rgb).all() and st.shadow_rays == 0


def test_unlit_scene_on_the_cpu_has_no_lights_and_no_shadow_rays(wf, oracle_ref):
    s = build(wf, base_arrays(), 32, 18)
    assert s.info().n_lights == 0
    (rgb, lum, lum2), st = wf.render_cpu_sums(s, 0, 4, wf.params(spp=4, threads=4))
    assert_equals_oracle((rgb, lum, lum2), (st.extension_rays, st.shadow_rays, st.invalid_hits), oracle_ref(s, 4),
                         "unlit CPU arm")
    assert st.shadow_rays == 0 and rgb.sum() > 0


@pytest.mark.gpu
def test_unlit_scene_every_gpu_arm_equals_the_oracle(wf, oracle_ref):
    s = build(wf, base_arrays(), 32, 18)
    want = oracle_ref(s, 4, 7)
    for mode in ("megakernel", "wavefront", "persistent"):
        r = wf.Renderer(s, mode)
        r.reset()
        st = r.run_passes(0, 4, wf.params(spp=4, seed=7))
        sums = r.download_sums()
        r.close()
        assert_equals_oracle(sums, (st.extension_rays, st.shadow_rays, st.invalid_hits), want, f"unlit {mode}")
        assert st.shadow_rays == 0 and st.queue_overflows == 0 and sums[0].sum() > 0


@pytest.mark.gpu
def test_queue_capacity_of_exactly_the_path_count_never_overflows(wf, oracle_ref):
    s = wf.Scene("divergence", 64, 36)
    want = oracle_ref(s, 4, 7)
    r = wf.Renderer(s, "wavefront")
    for cap in (0, 64 * 36):
        r.reset()
        st = r.run_passes(0, 4, wf.params(spp=4, seed=7, queue_capacity=cap))
        assert st.queue_overflows == 0 and st.invalid_hits == 0, cap
        assert_equals_oracle(r.download_sums(), (st.extension_rays, st.shadow_rays, st.invalid_hits), want,
                             f"capacity {cap}")
    r.close()


@pytest.mark.gpu
def test_wavefront_memory_at_720p_between_its_pool_and_the_budget(wf):
    s = wf.Scene("divergence", 1280, 720)
    r = wf.Renderer(s, "wavefront")
    rgb, var, st = r.render(wf.params(spp=1))
    dev = r.device_bytes()
    r.close()
    assert np.isfinite(rgb).all() and (var >= 0).all() and st.queue_overflows == 0 and st.invalid_hits == 0
    # 921,600 paths x 48 B (origin, direction, throughput, radiance) is the least a pool can take.
    assert 44_000_000 <= dev < 1_000_000_000
    assert 0 < st.peak_vram_bytes < 1_000_000_000 < 3_500_000_000
