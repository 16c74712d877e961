"""ctypes binding of the oracle (oracle/_build/liboracle.so) and, where it has been built in this
container, of the reference's own host code (oracle/_ref/libwfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker; the product never loads it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "_build", "liboracle.so")
# The same oracle built with the C library's binary32 sinf/cosf (the reference's host trigonometry).
ORACLE_LIBM_PATH = os.path.join(HERE, "_build", "liboracle_libm.so")
REF_PATH = os.path.join(HERE, "_ref", "libwfref.so")

_o = None
_olibm = None
_r = None
F = C.POINTER(C.c_float)
I = C.POINTER(C.c_int)


def _f(a):
    return a.ctypes.data_as(F) if a is not None else None


def _i(a):
    return a.ctypes.data_as(I) if a is not None else None


def oracle(libm_trig=False):
    """The oracle library; libm_trig=True: the build with the C library's binary32 sinf/cosf."""
    global _o, _olibm
    if libm_trig:
        if _olibm is None:
            _olibm = _load(ORACLE_LIBM_PATH)
        return _olibm
    if _o is None:
        _o = _load(ORACLE_PATH)
    return _o


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (make -C oracle)")
    L = C.CDLL(path)
    P = C.c_void_p
    L.or_scene_new.restype = P
    L.or_scene_new.argtypes = [F, C.c_int, F, C.c_int, I, C.c_int, F, C.c_int, F, F, C.c_int, C.c_int]
    L.or_scene_free.argtypes = [P]
    L.or_scene_n_nodes.argtypes = [P]
    L.or_scene_n_lights.argtypes = [P]
    L.or_scene_bvh.argtypes = [P, I, I]
    L.or_camera_rig.argtypes = [P, F]
    L.or_scene_set_view.argtypes = [P, F, C.c_int, C.c_int]
    L.or_scene_hash.restype = C.c_uint64
    L.or_scene_hash.argtypes = [P]
    L.or_hash_u32.restype = C.c_uint32
    L.or_hash_u32.argtypes = [C.c_uint32]
    L.or_path_key.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    L.or_rng.restype = C.c_float
    L.or_rng.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    L.or_sincos.argtypes = [F, C.c_int, F, F]
    L.or_intersect.argtypes = [P, F, C.c_int, C.c_int, F, I]
    L.or_intersect_brute.argtypes = [P, F, C.c_int, F, I]
    L.or_render.argtypes = [P, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, F, F, F,
                            C.POINTER(C.c_uint64)]
    L.or_wavefront_pass.restype = C.c_int
    L.or_wavefront_pass.argtypes = [P, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint64), C.c_int, F, F, F, C.POINTER(C.c_uint64)]
    L.or_finalize.argtypes = [F, F, F, C.c_int, C.c_int, F, F]
    L.or_compare.argtypes = [F, F, C.c_int, F, F, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.or_trace_paths.argtypes = [P, I, C.c_int, C.c_uint32, C.c_int, F, C.POINTER(C.c_uint64)]
    L.or_metal_dir.restype = C.c_int
    L.or_metal_dir.argtypes = [F, F, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32, F]
    return L


class OScene:
    """Oracle scene built from C-ABI-layout arrays (wfamd.Scene.arrays())."""

    def __init__(self, arrays, libm_trig=False, lib=None):
        a = arrays
        self.width, self.height = a["width"], a["height"]
        self.L = lib if lib is not None else oracle(libm_trig)
        self._keep = [np.ascontiguousarray(a[k]) for k in ("spheres", "vertices", "triangles", "materials", "env",
                                                           "camera")]
        s, v, t, m, e, c = self._keep
        self.h = self.L.or_scene_new(_f(s), len(s), _f(v), len(v), _i(t), len(t), _f(m), len(m), _f(e), _f(c),
                                     self.width, self.height)
        if not self.h:
            raise ValueError("oracle rejected the scene")
        self.n_prims = len(s) + len(t)

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "L", None) is not None:
            self.L.or_scene_free(self.h)
            self.h = None

    def bvh(self):
        n = self.L.or_scene_n_nodes(self.h)
        nodes = np.zeros((n, 12), np.int32)
        order = np.zeros(self.n_prims, np.int32)
        self.L.or_scene_bvh(self.h, _i(nodes), _i(order))
        return nodes, order

    def set_view(self, camera10, width, height):
        """Same geometry and BVH, another camera and image size."""
        cam = np.ascontiguousarray(camera10, np.float32)
        self.L.or_scene_set_view(self.h, _f(cam), int(width), int(height))
        self.width, self.height = int(width), int(height)

    def camera_rig(self):
        r = np.zeros(12, np.float32)
        self.L.or_camera_rig(self.h, _f(r))
        return r

    def hash(self):
        return self.L.or_scene_hash(self.h)

    def n_lights(self):
        return self.L.or_scene_n_lights(self.h)

    def intersect(self, rays, anyhit=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        t = np.zeros(len(rays), np.float32)
        p = np.zeros(len(rays), np.int32)
        self.L.or_intersect(self.h, _f(rays), len(rays), int(anyhit), _f(t), _i(p))
        return t, p

    def intersect_brute(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        t = np.zeros(len(rays), np.float32)
        p = np.zeros(len(rays), np.int32)
        self.L.or_intersect_brute(self.h, _f(rays), len(rays), _f(t), _i(p))
        return t, p

    def render_sums(self, first, count, seed=1234, max_depth=8, threads=1, sums=None):
        n = self.width * self.height
        if sums is None:
            sums = (np.zeros(3 * n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))
        cnt = (C.c_uint64 * 3)()
        self.L.or_render(self.h, first, count, seed, max_depth, threads, _f(sums[0]), _f(sums[1]), _f(sums[2]),
                         cnt)
        return sums, (cnt[0], cnt[1], cnt[2])

    def trace_paths(self, pxys, seed=1234, max_depth=8):
        """Radiance (n x 3) of single paths given as (px, py, sample) rows, and the ray counts."""
        q = np.ascontiguousarray(pxys, np.int32).reshape(-1, 3)
        L = np.zeros((len(q), 3), np.float32)
        cnt = (C.c_uint64 * 3)()
        self.L.or_trace_paths(self.h, _i(q), len(q), seed, max_depth, _f(L), cnt)
        return L, (cnt[0], cnt[1], cnt[2])

    def render(self, spp, seed=1234, max_depth=8, threads=1):
        sums, cnt = self.render_sums(0, spp, seed, max_depth, threads)
        rgb, var = finalize(*sums, spp)
        return rgb, var, cnt

    def wavefront_pass(self, sample, seed=1234, max_depth=8, queue_capacity=0, sums=None):
        n = self.width * self.height
        if sums is None:
            sums = (np.zeros(3 * n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))
        maxr = 128
        rec = np.zeros((maxr, 6), np.uint32)
        hsh = np.zeros(maxr, np.uint64)
        cnt = (C.c_uint64 * 3)()
        nb = self.L.or_wavefront_pass(self.h, sample, seed, max_depth, queue_capacity,
                                      rec.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      hsh.ctypes.data_as(C.POINTER(C.c_uint64)), maxr, _f(sums[0]), _f(sums[1]),
                                      _f(sums[2]), cnt)
        trace = [tuple(int(x) for x in rec[b]) + (int(hsh[b]),) for b in range(min(nb, maxr))]
        return trace, sums, (cnt[0], cnt[1], cnt[2])


def hash_u32(x):
    return oracle().or_hash_u32(x)


def path_key(seed, pixel, sample):
    k = (C.c_uint32 * 2)()
    oracle().or_path_key(seed, pixel, sample, k)
    return k[0], k[1]


def rng(k0, k1, depth, dim):
    return oracle().or_rng(k0, k1, depth, dim)


def load(path):
    """An oracle library built elsewhere (for instance a copy with a planted mistake)."""
    return _load(path)


def metal_dir(d, n, fuzz, k0, k1, depth=0, libm_trig=False, lib=None):
    """Scattered direction of a metal surface (or None if the path is absorbed)."""
    d = np.ascontiguousarray(d, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    L = lib if lib is not None else oracle(libm_trig)
    alive = L.or_metal_dir(_f(d), _f(n), float(fuzz), k0, k1, depth, _f(out))
    return out if alive else None


def sincos(x):
    x = np.ascontiguousarray(x, np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    oracle().or_sincos(_f(x), x.size, _f(s), _f(c))
    return s, c


def finalize(rgb_sum, lum_sum, lum2_sum, spp):
    n = lum_sum.size
    rgb, var = np.zeros(3 * n, np.float32), np.zeros(n, np.float32)
    oracle().or_finalize(_f(rgb_sum), _f(lum_sum), _f(lum2_sum), n, spp, _f(rgb), _f(var))
    return rgb, var


def compare(rgb_a, var_a, spp_a, rgb_b, var_b, spp_b):
    out = (C.c_double * 4)()
    n = var_a.size
    oracle().or_compare(_f(np.ascontiguousarray(rgb_a, np.float32)), _f(np.ascontiguousarray(var_a, np.float32)),
                        spp_a, _f(np.ascontiguousarray(rgb_b, np.float32)),
                        _f(np.ascontiguousarray(var_b, np.float32)), spp_b, n, out)
    return dict(max_abs_rel=out[0], mean_rel=out[1], frac_within_3se=out[2], nonfinite=int(out[3]))


# ---------------------------------------------------------------------------------------------
# reference's own host code (built in this container only)
# ---------------------------------------------------------------------------------------------
def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _r
    if _r is None:
        L = C.CDLL(REF_PATH)
        P = C.c_void_p
        L.ref_make_scene.restype = P
        L.ref_make_scene.argtypes = [C.c_char_p, C.c_int, C.c_int]
        L.ref_scene_from_arrays.restype = P
        L.ref_scene_from_arrays.argtypes = [F, C.c_int, F, C.c_int, I, C.c_int, F, C.c_int, F, F, C.c_int, C.c_int]
        L.ref_scene_free.argtypes = [P]
        L.ref_scene_counts.argtypes = [P, I]
        L.ref_scene_export.argtypes = [P, F, F, I, F, F, F]
        L.ref_scene_bvh.argtypes = [P, I, I]
        L.ref_scene_hash.restype = C.c_uint64
        L.ref_scene_hash.argtypes = [P]
        L.ref_camera_rig.argtypes = [P, F]
        L.ref_hash_u32.restype = C.c_uint32
        L.ref_hash_u32.argtypes = [C.c_uint32]
        L.ref_path_key.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
        L.ref_rng.restype = C.c_float
        L.ref_rng.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.ref_intersect.argtypes = [P, F, C.c_int, C.c_int, F, I]
        L.ref_trace_path.argtypes = [P, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_int, F]
        L.ref_render_cpu.argtypes = [P, C.c_int, C.c_uint32, C.c_int, C.c_int, F, F]
        _r = L
    return _r


class RefScene:
    def __init__(self, name=None, width=0, height=0, arrays=None):
        if arrays is not None:
            a = arrays
            self._keep = [np.ascontiguousarray(a[k]) for k in ("spheres", "vertices", "triangles", "materials",
                                                               "env", "camera")]
            s, v, t, m, e, c = self._keep
            self.h = ref().ref_scene_from_arrays(_f(s), len(s), _f(v), len(v), _i(t), len(t), _f(m), len(m), _f(e),
                                                 _f(c), a["width"], a["height"])
            self.width, self.height = a["width"], a["height"]
        else:
            self.h = ref().ref_make_scene(name.encode(), width, height)
            self.width, self.height = width, height
        if not self.h:
            raise ValueError("reference rejected the scene")

    def __del__(self):
        if getattr(self, "h", None) and _r is not None:
            _r.ref_scene_free(self.h)

    def counts(self):
        c = np.zeros(6, np.int32)
        ref().ref_scene_counts(self.h, _i(c))
        return c

    def arrays(self):
        ns, nv, nt, nm, _, _ = self.counts()
        sph, ver = np.zeros((ns, 8), np.float32), np.zeros((nv, 3), np.float32)
        tri, mat = np.zeros((nt, 4), np.int32), np.zeros((nm, 12), np.float32)
        env, cam = np.zeros(20, np.float32), np.zeros(10, np.float32)
        ref().ref_scene_export(self.h, _f(sph), _f(ver), _i(tri), _f(mat), _f(env), _f(cam))
        return dict(spheres=sph, vertices=ver, triangles=tri, materials=mat, env=env, camera=cam,
                    width=self.width, height=self.height)

    def bvh(self):
        c = self.counts()
        nodes = np.zeros((c[4], 12), np.int32)
        order = np.zeros(c[0] + c[2], np.int32)
        ref().ref_scene_bvh(self.h, _i(nodes), _i(order))
        return nodes, order

    def hash(self):
        return ref().ref_scene_hash(self.h)

    def camera_rig(self):
        r = np.zeros(12, np.float32)
        ref().ref_camera_rig(self.h, _f(r))
        return r

    def intersect(self, rays, anyhit=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        t = np.zeros(len(rays), np.float32)
        p = np.zeros(len(rays), np.int32)
        ref().ref_intersect(self.h, _f(rays), len(rays), int(anyhit), _f(t), _i(p))
        return t, p

    def trace_path(self, px, py, sample, seed, max_depth):
        L = np.zeros(3, np.float32)
        ref().ref_trace_path(self.h, px, py, sample, seed, max_depth, _f(L))
        return L

    def render_cpu(self, spp, seed=1234, max_depth=8, threads=1):
        n = self.width * self.height
        rgb, var = np.zeros(3 * n, np.float32), np.zeros(n, np.float32)
        ref().ref_render_cpu(self.h, spp, seed, max_depth, threads, _f(rgb), _f(var))
        return rgb, var
