"""ctypes binding of libwfamd.so (include/wfamd.h) for Python callers: tests, bench.py and the
build/smoke entry points. Plumbing only -- the engine is the native library. Loading fails loudly when the
library has not been built; there is no fallback path.

Mirrors the reference API: Scene/make_scene (src/scenes.cpp:29), GpuRenderer
(include/wf/renderer.hpp:44-67), render_cpu (include/wf/renderer.hpp:39) and the queue
primitives (src/gpu/queues.cuh:29-54).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libwfamd.so")

MODE_CPU, MODE_MEGAKERNEL, MODE_WAVEFRONT, MODE_PERSISTENT = 0, 1, 2, 3
MODES = {"cpu": MODE_CPU, "megakernel": MODE_MEGAKERNEL, "wavefront": MODE_WAVEFRONT, "persistent": MODE_PERSISTENT}
WF_OK, WF_ERR_INVALID, WF_ERR_RUNTIME, WF_ERR_BUDGET = 0, 1, 2, 3


class WfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class InvalidArgument(WfError, ValueError):
    pass


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_int), ("seed", C.c_uint32), ("max_depth", C.c_int), ("threads", C.c_int),
                ("count_rays", C.c_int), ("stage_timing", C.c_int), ("queue_capacity", C.c_int),
                ("trace_bounces", C.c_int), ("block_size", C.c_int), ("pool_paths", C.c_int),
                ("scan_tile", C.c_int), ("lds_nodes", C.c_int), ("lds_kib", C.c_int), ("refill", C.c_int),
                ("leaf_batch", C.c_int), ("dir_bins", C.c_int), ("shade_block", C.c_int), ("graph_iters", C.c_int)]


class RenderStats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("total_ms", C.c_double), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("queue_high_water", C.c_int64 * 5), ("queue_overflows", C.c_int64),
                ("invalid_hits", C.c_int64), ("peak_vram_bytes", C.c_uint64), ("stage_ms", C.c_double * 8),
                ("iterations", C.c_int64), ("stage_items", C.c_int64 * 8), ("stage_launches", C.c_int64 * 8),
                ("stage_bytes", C.c_int64 * 8), ("stage_threads", C.c_int64 * 8)]

    STAGES = ("extend", "classify", "shade_lambert", "shade_metal", "shade_dielectric", "shadow", "compact", "other")

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        for k in ("queue_high_water", "stage_ms", "stage_items", "stage_launches", "stage_bytes", "stage_threads"):
            d[k] = list(getattr(self, k))
        return d

    def stages(self):
        return {n: {"ms": self.stage_ms[i], "items": self.stage_items[i], "launches": self.stage_launches[i],
                    "bytes": self.stage_bytes[i], "threads": self.stage_threads[i]}
                for i, n in enumerate(self.STAGES) if self.stage_launches[i] > 0}


class SceneInfo(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("n_spheres", C.c_int), ("n_triangles", C.c_int),
                ("n_vertices", C.c_int), ("n_materials", C.c_int), ("n_lights", C.c_int), ("n_nodes", C.c_int),
                ("bvh_leaves", C.c_int), ("bvh_internal", C.c_int), ("bvh_max_depth", C.c_int),
                ("bvh_sah_cost", C.c_double), ("hash", C.c_uint64)]


class SceneDesc(C.Structure):
    _fields_ = [("spheres", C.POINTER(C.c_float)), ("n_spheres", C.c_int), ("vertices", C.POINTER(C.c_float)),
                ("n_vertices", C.c_int), ("triangles", C.POINTER(C.c_int)), ("n_triangles", C.c_int),
                ("materials", C.POINTER(C.c_float)), ("n_materials", C.c_int), ("env", C.POINTER(C.c_float)),
                ("camera", C.POINTER(C.c_float)), ("width", C.c_int), ("height", C.c_int)]


class DeviceFacts(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int),
                ("runtime_version", C.c_int), ("driver_version", C.c_int), ("total_memory_bytes", C.c_uint64),
                ("free_memory_bytes", C.c_uint64)]


class BounceTrace(C.Structure):
    _fields_ = [("active", C.c_uint32), ("lambert", C.c_uint32), ("metal", C.c_uint32),
                ("dielectric", C.c_uint32), ("shadow", C.c_uint32), ("next_active", C.c_uint32),
                ("list_hash", C.c_uint64)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("spp", C.c_int), ("seed", C.c_uint32), ("max_depth", C.c_int),
                ("rgb", C.POINTER(C.c_float)), ("lum_var", C.POINTER(C.c_float)), ("width", C.c_int),
                ("height", C.c_int), ("yaw_deg", C.c_float), ("orbit", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("invalid_hits", C.c_int64), ("queue_overflows", C.c_int64), ("iterations", C.c_int64),
                ("graph_launches", C.c_int64), ("block_size", C.c_int), ("pool_paths", C.c_int),
                ("lds_kib", C.c_int), ("refill", C.c_int), ("leaf_batch", C.c_int),
                ("dir_bins", C.c_int), ("shade_block", C.c_int),
                ("device_bytes", C.c_uint64),
                ("stage_ms", C.c_double * 8), ("stage_items", C.c_int64 * 8), ("stage_launches", C.c_int64 * 8),
                ("stage_bytes", C.c_int64 * 8), ("stage_threads", C.c_int64 * 8)]

    def stages(self):
        return {n: {"ms": self.stage_ms[i], "items": self.stage_items[i], "launches": self.stage_launches[i],
                    "bytes": self.stage_bytes[i], "threads": self.stage_threads[i]}
                for i, n in enumerate(RenderStats.STAGES) if self.stage_launches[i] > 0}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
                ("rejected", C.c_int64), ("in_flight", C.c_int64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("bytes_in_use", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("busy_ms", C.c_double)]


_lib = None

# Every symbol declared in include/wfamd.h.
EXPORTS = [
    "wf_last_error", "wf_version", "wf_scene_create", "wf_scene_from_arrays", "wf_scene_orbit", "wf_scene_view",
    "wf_scene_destroy", "wf_scene_info_get", "wf_scene_export", "wf_scene_bvh", "wf_scene_camera_rig",
    "wf_renderer_create", "wf_renderer_destroy", "wf_renderer_render", "wf_renderer_reset",
    "wf_renderer_run_passes", "wf_renderer_download", "wf_renderer_download_mean", "wf_renderer_download_sums",
    "wf_renderer_device_bytes", "wf_renderer_device_name", "wf_renderer_trace", "wf_render_cpu",
    "wf_render_cpu_sums", "wf_finalize", "wf_query_device", "wf_scan_workspace_create",
    "wf_scan_workspace_destroy", "wf_scan_exclusive_u32", "wf_compact_u32", "wf_compact_u32_noalloc",
    "wf_test_queue_push",
    "wf_batch_render", "wf_autotune", "wf_autotune_budget", "wf_server_create", "wf_server_destroy", "wf_server_submit",
    "wf_server_step", "wf_server_poll", "wf_server_stats_get", "wf_test_sincos", "wf_test_sincos_grid_hash",
    "wf_test_intersect", "wf_test_slab", "wf_test_divu32", "wf_test_lds_plan", "wf_test_trace_paths", "wf_test_shade",
    "wf_copy_probe",
    "wf_cached_memory_bytes", "wf_release_cached_memory",
]


def lib():
    """Loads libwfamd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C "
                           f"graphics-book-tutorials_amd")
    L = C.CDLL(LIB_PATH)
    P, I, U64, F = C.c_void_p, C.c_int, C.c_uint64, C.POINTER(C.c_float)
    sig = {
        "wf_last_error": (C.c_char_p, []), "wf_version": (C.c_char_p, []),
        "wf_scene_create": (I, [C.c_char_p, I, I, C.POINTER(P)]),
        "wf_scene_from_arrays": (I, [C.POINTER(SceneDesc), C.POINTER(P)]),
        "wf_scene_orbit": (I, [P, C.c_float]), "wf_scene_destroy": (None, [P]),
        "wf_scene_view": (I, [P, I, I, C.c_float, I, C.POINTER(P)]),
        "wf_scene_info_get": (I, [P, C.POINTER(SceneInfo)]),
        "wf_scene_export": (I, [P, F, F, C.POINTER(C.c_int), F, F, F]),
        "wf_scene_bvh": (I, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "wf_scene_camera_rig": (I, [P, F]),
        "wf_renderer_create": (I, [P, I, C.POINTER(P)]), "wf_renderer_destroy": (None, [P]),
        "wf_renderer_render": (I, [P, C.POINTER(RenderParams), C.POINTER(RenderStats), F, F]),
        "wf_renderer_reset": (I, [P]),
        "wf_renderer_run_passes": (I, [P, I, I, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
        "wf_renderer_download": (I, [P, I, F, F]), "wf_renderer_download_mean": (I, [P, I, F]),
        "wf_renderer_download_sums": (I, [P, F, F, F]), "wf_renderer_device_bytes": (U64, [P]),
        "wf_renderer_device_name": (C.c_char_p, [P]),
        "wf_renderer_trace": (I, [P, C.POINTER(BounceTrace), I]),
        "wf_render_cpu": (I, [P, C.POINTER(RenderParams), C.POINTER(RenderStats), F, F]),
        "wf_render_cpu_sums": (I, [P, I, I, C.POINTER(RenderParams), C.POINTER(RenderStats), F, F, F]),
        "wf_finalize": (I, [F, F, F, I, I, F, F]),
        "wf_query_device": (I, [C.POINTER(DeviceFacts)]),
        "wf_scan_workspace_create": (I, [I, C.POINTER(P)]), "wf_scan_workspace_destroy": (None, [P]),
        "wf_scan_exclusive_u32": (I, [P, P, I, P, P, P, I]),
        "wf_compact_u32": (I, [P, P, P, P, P, I, P, P, I]),
        "wf_compact_u32_noalloc": (I, [P, P, P, P, I, P, I]),
        "wf_test_queue_push": (I, [P, I, P, C.c_uint32, P, P, P]),
        "wf_batch_render": (I, [C.POINTER(Job), I, C.POINTER(RenderParams), C.POINTER(BatchStats)]),
        "wf_autotune": (I, [C.POINTER(Job), I, I, C.POINTER(RenderParams)]),
        "wf_autotune_budget": (I, [C.POINTER(Job), I, I, U64, C.POINTER(RenderParams)]),
        "wf_server_create": (I, [U64, C.POINTER(RenderParams), C.POINTER(P)]),
        "wf_server_destroy": (None, [P]),
        "wf_server_submit": (I, [P, C.POINTER(Job), C.POINTER(C.c_int64)]),
        "wf_server_step": (I, [P, I]), "wf_server_poll": (I, [P, C.c_int64]),
        "wf_server_stats_get": (I, [P, C.POINTER(ServerStats)]),
        "wf_test_sincos": (I, [F, I, F, F, I]),
        "wf_test_sincos_grid_hash": (I, [C.c_uint32, C.c_uint32, I, C.POINTER(C.c_uint64)]),
        "wf_test_intersect": (I, [P, F, I, I, F, C.POINTER(C.c_int), I]),
        "wf_test_slab": (I, [F, I, C.POINTER(C.c_int), I]),
        "wf_test_divu32": (I, [C.POINTER(C.c_uint32), I, C.c_uint32, C.POINTER(C.c_uint32), I]),
        "wf_test_lds_plan": (I, [C.POINTER(C.c_int), C.POINTER(C.c_int), I, C.c_int64, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]),
        "wf_test_trace_paths": (I, [P, C.POINTER(C.c_int), I, C.c_uint32, I, F, C.POINTER(C.c_uint32), I]),
        "wf_test_shade": (I, [P, C.POINTER(C.c_uint32), I, I, C.POINTER(C.c_uint32)]),
        "wf_cached_memory_bytes": (U64, []), "wf_release_cached_memory": (I, [C.POINTER(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != WF_OK:
        msg = lib().wf_last_error().decode()
        if rc == WF_ERR_INVALID:
            raise InvalidArgument(rc, msg)
        raise WfError(rc, msg)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


def params(spp=64, seed=1234, max_depth=8, threads=0, stage_timing=False, queue_capacity=0, trace_bounces=False,
           block_size=0, pool_paths=0, scan_tile=0, lds_nodes=-1, lds_kib=0, refill=0, leaf_batch=0,
           dir_bins=0, shade_block=0, graph_iters=0):
    return RenderParams(spp, seed, max_depth, threads, 0, int(stage_timing), queue_capacity, int(trace_bounces),
                        block_size, pool_paths, scan_tile, lds_nodes, lds_kib, refill, leaf_batch, dir_bins,
                        shade_block, graph_iters)


class Scene:
    """A finalized scene (make_scene, src/scenes.cpp:29, or arrays)."""

    def __init__(self, name=None, width=1280, height=720, handle=None):
        self._h = C.c_void_p()
        if handle is not None:
            self._h = handle
        else:
            _check(lib().wf_scene_create(name.encode(), width, height, C.byref(self._h)))

    @classmethod
    def from_arrays(cls, spheres, vertices, triangles, materials, env, camera, width, height):
        sph = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1)
        ver = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1)
        tri = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1)
        mat = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1)
        env = np.ascontiguousarray(env, dtype=np.float32).reshape(-1)
        cam = np.ascontiguousarray(camera, dtype=np.float32).reshape(-1)
        d = SceneDesc(_fptr(sph), sph.size // 8, _fptr(ver), ver.size // 3, _iptr(tri), tri.size // 4, _fptr(mat),
                      mat.size // 12, _fptr(env), _fptr(cam), width, height)
        h = C.c_void_p()
        _check(lib().wf_scene_from_arrays(C.byref(d), C.byref(h)))
        return cls(handle=h)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.wf_scene_destroy(self._h)
            self._h = C.c_void_p()

    @property
    def handle(self):
        return self._h

    def info(self):
        i = SceneInfo()
        _check(lib().wf_scene_info_get(self._h, C.byref(i)))
        return i

    @property
    def width(self):
        return self.info().width

    @property
    def height(self):
        return self.info().height

    def orbit(self, yaw_deg):
        _check(lib().wf_scene_orbit(self._h, float(yaw_deg)))

    def view(self, width, height, yaw_deg=None):
        """A copy of this scene seen at width x height and, if yaw_deg is given, from the orbited
        camera (the standalone scene of a job view)."""
        h = C.c_void_p()
        _check(lib().wf_scene_view(self._h, int(width), int(height), float(yaw_deg or 0.0),
                                   0 if yaw_deg is None else 1, C.byref(h)))
        return Scene(handle=h)

    def arrays(self):
        i = self.info()
        sph = np.zeros((i.n_spheres, 8), np.float32)
        ver = np.zeros((i.n_vertices, 3), np.float32)
        tri = np.zeros((i.n_triangles, 4), np.int32)
        mat = np.zeros((i.n_materials, 12), np.float32)
        env = np.zeros(20, np.float32)
        cam = np.zeros(10, np.float32)
        _check(lib().wf_scene_export(self._h, _fptr(sph), _fptr(ver), _iptr(tri), _fptr(mat), _fptr(env), _fptr(cam)))
        return dict(spheres=sph, vertices=ver, triangles=tri, materials=mat, env=env, camera=cam,
                    width=i.width, height=i.height)

    def bvh(self):
        i = self.info()
        nodes = np.zeros((i.n_nodes, 12), np.int32)
        order = np.zeros(i.n_spheres + i.n_triangles, np.int32)
        _check(lib().wf_scene_bvh(self._h, _iptr(nodes), _iptr(order)))
        return nodes, order

    def camera_rig(self):
        r = np.zeros(12, np.float32)
        _check(lib().wf_scene_camera_rig(self._h, _fptr(r)))
        return r


class Renderer:
    """wf::GpuRenderer over the C ABI (include/wf/renderer.hpp:44-67)."""

    def __init__(self, scene, mode):
        self.scene = scene
        self.mode = MODES[mode] if isinstance(mode, str) else mode
        self._h = C.c_void_p()
        _check(lib().wf_renderer_create(scene.handle, self.mode, C.byref(self._h)))
        i = scene.info()
        self.width, self.height = i.width, i.height

    def close(self):
        if self._h and self._h.value:
            lib().wf_renderer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _lib is not None:
            self.close()

    def render(self, p):
        rgb = np.zeros(3 * self.width * self.height, np.float32)
        var = np.zeros(self.width * self.height, np.float32)
        st = RenderStats()
        _check(lib().wf_renderer_render(self._h, C.byref(p), C.byref(st), _fptr(rgb), _fptr(var)))
        return rgb, var, st

    def reset(self):
        _check(lib().wf_renderer_reset(self._h))

    def run_passes(self, first, count, p):
        st = RenderStats()
        _check(lib().wf_renderer_run_passes(self._h, first, count, C.byref(p), C.byref(st)))
        return st

    def download(self, spp):
        rgb = np.zeros(3 * self.width * self.height, np.float32)
        var = np.zeros(self.width * self.height, np.float32)
        _check(lib().wf_renderer_download(self._h, spp, _fptr(rgb), _fptr(var)))
        return rgb, var

    def download_mean(self, spp):
        rgb = np.zeros(3 * self.width * self.height, np.float32)
        _check(lib().wf_renderer_download_mean(self._h, spp, _fptr(rgb)))
        return rgb

    def download_sums(self):
        n = self.width * self.height
        rgb, lum, lum2 = np.zeros(3 * n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(lib().wf_renderer_download_sums(self._h, _fptr(rgb), _fptr(lum), _fptr(lum2)))
        return rgb, lum, lum2

    def device_bytes(self):
        return lib().wf_renderer_device_bytes(self._h)

    def device_name(self):
        return lib().wf_renderer_device_name(self._h).decode()

    def trace(self):
        n = lib().wf_renderer_trace(self._h, None, 0)
        arr = (BounceTrace * max(n, 1))()
        lib().wf_renderer_trace(self._h, arr, n)
        return [(t.active, t.lambert, t.metal, t.dielectric, t.shadow, t.next_active, t.list_hash) for t in arr[:n]]


def render_cpu(scene, p):
    i = scene.info()
    rgb = np.zeros(3 * i.width * i.height, np.float32)
    var = np.zeros(i.width * i.height, np.float32)
    st = RenderStats()
    _check(lib().wf_render_cpu(scene.handle, C.byref(p), C.byref(st), _fptr(rgb), _fptr(var)))
    return rgb, var, st


def render_cpu_sums(scene, first, count, p, sums=None):
    i = scene.info()
    n = i.width * i.height
    if sums is None:
        sums = (np.zeros(3 * n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))
    st = RenderStats()
    _check(lib().wf_render_cpu_sums(scene.handle, first, count, C.byref(p), C.byref(st), _fptr(sums[0]),
                                    _fptr(sums[1]), _fptr(sums[2])))
    return sums, st


def finalize(rgb_sum, lum_sum, lum2_sum, spp):
    n = lum_sum.size
    rgb, var = np.zeros(3 * n, np.float32), np.zeros(n, np.float32)
    _check(lib().wf_finalize(_fptr(rgb_sum), _fptr(lum_sum), _fptr(lum2_sum), n, spp, _fptr(rgb), _fptr(var)))
    return rgb, var


def query_device():
    f = DeviceFacts()
    _check(lib().wf_query_device(C.byref(f)))
    return dict(name=f.name.decode(), arch=f.arch.decode(), compute_units=f.compute_units,
                runtime_version=f.runtime_version, driver_version=f.driver_version,
                total_memory_bytes=f.total_memory_bytes, free_memory_bytes=f.free_memory_bytes)


def cached_memory_bytes():
    """Device bytes of ended persistent engines kept for reuse (wf_cached_memory_bytes)."""
    return int(lib().wf_cached_memory_bytes())


def release_cached_memory():
    """Frees the kept blocks; returns the bytes freed (wf_release_cached_memory)."""
    n = C.c_uint64()
    _check(lib().wf_release_cached_memory(C.byref(n)))
    return n.value


def copy_probe(nbytes, reps=10):
    """Average ms of a plain device copy of nbytes (reads and writes nbytes; counter calibration)."""
    ms = C.c_double()
    _check(lib().wf_copy_probe(C.c_uint64(nbytes), reps, C.byref(ms)))
    return ms.value


def test_sincos(x, device=False):
    x = np.ascontiguousarray(x, np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    _check(lib().wf_test_sincos(_fptr(x), x.size, _fptr(s), _fptr(c), int(device)))
    return s, c


def test_sincos_grid_hash(k0=0, k1=1 << 24, device=False):
    h = C.c_uint64()
    _check(lib().wf_test_sincos_grid_hash(k0, k1, int(device), C.byref(h)))
    return h.value


def test_slab(cases, device=False):
    """(reference slab test, signed slab test or -1) per 14-float case."""
    cases = np.ascontiguousarray(cases, np.float32).reshape(-1, 14)
    out = np.zeros((len(cases), 2), np.int32)
    _check(lib().wf_test_slab(_fptr(cases), len(cases), _iptr(out), int(device)))
    return out


def test_divu32(n, d, device=False):
    n = np.ascontiguousarray(n, np.uint32)
    q = np.zeros(len(n), np.uint32)
    _check(lib().wf_test_divu32(n.ctypes.data_as(C.POINTER(C.c_uint32)), len(n), int(d),
                                q.ctypes.data_as(C.POINTER(C.c_uint32)), int(device)))
    return q


def test_lds_plan(n_nodes, n_prims, budget_bytes):
    """(nodes, primitive records) each scene keeps in a persistent workgroup's LDS."""
    nn = np.ascontiguousarray(n_nodes, np.int32)
    np_ = np.ascontiguousarray(n_prims, np.int32)
    nc = np.zeros(len(nn), np.int32)
    pc = np.zeros(len(nn), np.int32)
    _check(lib().wf_test_lds_plan(nn.ctypes.data_as(C.POINTER(C.c_int)), np_.ctypes.data_as(C.POINTER(C.c_int)),
                                  len(nn), int(budget_bytes), nc.ctypes.data_as(C.POINTER(C.c_int)),
                                  pc.ctypes.data_as(C.POINTER(C.c_int))))
    return nc, pc


def test_intersect(scene, rays, anyhit=False, device=False):
    """device: False/0 host, True/1 one GPU thread per ray, 2 wave-scheduled traversal with the
    tree top in LDS, 3 the same with 8 nodes in LDS."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    t = np.zeros(len(rays), np.float32)
    p = np.zeros(len(rays), np.int32)
    _check(lib().wf_test_intersect(scene.handle, _fptr(rays), len(rays), int(anyhit), _fptr(t), _iptr(p), int(device)))
    return t, p


def test_trace_paths(scene, pxys, seed=1234, max_depth=8, device=False):
    q = np.ascontiguousarray(pxys, np.int32).reshape(-1, 3)
    L = np.zeros((len(q), 3), np.float32)
    rays = np.zeros((len(q), 2), np.uint32)
    _check(lib().wf_test_trace_paths(scene.handle, _iptr(q), len(q), seed, max_depth, _fptr(L),
                                     rays.ctypes.data_as(C.POINTER(C.c_uint32)), int(device)))
    return L, rays


SHADE_CASE_WORDS, SHADE_OUT_WORDS = 20, 34


def test_shade(scene, cases, form=0):
    """One classify + shade step per case (SHADE_CASE_WORDS uint32 words, floats as their bits).
    form: 0 the per-material functions on the host, 1 the one-pass shade_surface on the host, 2 and 3
    the same on the GPU. A case whose primitive word is -2 has its closest hit found on the host."""
    q = np.ascontiguousarray(cases, np.uint32).reshape(-1, SHADE_CASE_WORDS)
    out = np.zeros((len(q), SHADE_OUT_WORDS), np.uint32)
    u32 = C.POINTER(C.c_uint32)
    _check(lib().wf_test_shade(scene.handle, q.ctypes.data_as(u32), len(q), int(form), out.ctypes.data_as(u32)))
    return out


def make_job(scene, spp, seed=1234, max_depth=8, width=0, height=0, yaw_deg=None):
    """A Job and its output arrays. width/height 0: the scene's size; yaw_deg: orbit the camera."""
    i = scene.info()
    w, h = (width or i.width), (height or i.height)
    rgb = np.zeros(3 * w * h, np.float32)
    var = np.zeros(w * h, np.float32)
    job = Job(scene.handle, spp, seed, max_depth, _fptr(rgb), _fptr(var), int(width), int(height),
              float(yaw_deg or 0.0), 0 if yaw_deg is None else 1)
    return job, (rgb, var)


def make_jobs(specs):
    """specs: list of (scene, spp, seed, max_depth[, width, height[, yaw_deg]]). Returns (ctypes Job
    array, output arrays)."""
    arr = (Job * len(specs))()
    outs = []
    for k, spec in enumerate(specs):
        arr[k], out = make_job(*spec)
        outs.append(out)
    return arr, outs


def batch_render(specs, tuning=None):
    jobs, outs = make_jobs(specs)
    st = BatchStats()
    p = tuning if tuning is not None else params()
    _check(lib().wf_batch_render(jobs, len(specs), C.byref(p), C.byref(st)))
    return outs, st


def autotune(specs, probe_spp=2, state_budget=None):
    """Tuning for a batch (wf_autotune); state_budget: the most device bytes of path state its pool
    may take (wf_autotune_budget)."""
    jobs, _ = make_jobs(specs)
    p = params()
    if state_budget is None:
        _check(lib().wf_autotune(jobs, len(specs), probe_spp, C.byref(p)))
    else:
        _check(lib().wf_autotune_budget(jobs, len(specs), probe_spp, int(state_budget), C.byref(p)))
    return p


class Server:
    """Continuous-batching serving loop (wf_server_*)."""

    def __init__(self, budget_bytes, tuning=None):
        self._h = C.c_void_p()
        p = tuning if tuning is not None else params()
        _check(lib().wf_server_create(int(budget_bytes), C.byref(p), C.byref(self._h)))
        self._keep = {}

    def close(self):
        if self._h and self._h.value:
            lib().wf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _lib is not None:
            self.close()

    def submit(self, scene, spp, seed=1234, max_depth=8, width=0, height=0, yaw_deg=None):
        job, (rgb, var) = make_job(scene, spp, seed, max_depth, width, height, yaw_deg)
        rid = C.c_int64()
        _check(lib().wf_server_submit(self._h, C.byref(job), C.byref(rid)))
        self._keep[rid.value] = (scene, rgb, var)
        return rid.value

    def step(self, launches=1):
        _check(lib().wf_server_step(self._h, launches))

    def poll(self, rid):
        return lib().wf_server_poll(self._h, rid)

    def result(self, rid):
        _, rgb, var = self._keep[rid]
        return rgb, var

    def release(self, rid):
        self._keep.pop(rid, None)

    def stats(self):
        s = ServerStats()
        _check(lib().wf_server_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}
