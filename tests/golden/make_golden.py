"""Writes the golden fixtures in tests/golden/ from the reference's own host code, compiled in this
container by `make -C oracle ref` into oracle/_ref/libwfref.so (the reference sources never leave
/root/reference). Run from the repository root: python tests/golden/make_golden.py

Fixtures (data only):
  rng.json        hash_u32, make_path_key and rng bit patterns (include/wf/rng.hpp)
  scenes.json     reference scenes at 1280x720: scene hash, node count, FNV-1a of the node array
                  and of prim_order (src/scenes.cpp, src/bvh_build.cpp)
  ref_render.npz  render_cpu mean images of the reference scenes at 64x36, 4 spp, seed 1234, depth 8
  ref_paths.npz   trace_path radiance of 256 single paths per reference scene (64x36 image)
  intersect.npz   closest-hit (t, prim) of 1024 seeded rays per reference scene; and closest- and
                  any-hit results of 2048 axis-aligned rays (two exact zeros in d, of either sign,
                  so 1/d has infinite components) from origins on BVH node planes, the cases where
                  the slab test meets 0 * inf
"""
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SCENES = ["cornell", "divergence", "trianglehv", "furnace"]
RW, RH = 64, 36   # golden render and single-path image size
# (px, py, sample) of the single-path fixtures.
PATHS = [((k * 37) % RW, (k * 11) % RH, k % 7) for k in range(256)]


def fnv64(data: bytes) -> int:
    h = 14695981039346656037
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def seeded_rays(n, seed, center, radius):
    rs = np.random.default_rng(seed)
    o = center + rs.uniform(-radius, radius, size=(n, 3))
    d = rs.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = 3.402823466e38
    return rays


def axis_rays(nodes, n, seed):
    """Rays along an axis (the other two components +0 or -0) from origins on node box planes
    (sometimes nudged by 1e-3), tmin 1e-4, tmax FLT_MAX. nodes: int32 rows with lo at columns 0-2
    and hi at 4-6 (float bits)."""
    lo = np.ascontiguousarray(nodes[:, 0:3]).view(np.float32)
    hi = np.ascontiguousarray(nodes[:, 4:7]).view(np.float32)
    rs = np.random.default_rng(seed)
    k = rs.integers(0, len(nodes), n)
    o = np.where(rs.random((n, 3)) < 0.5, lo[k], hi[k]) + \
        rs.choice(np.array([0.0, 0.0, 1e-3, -1e-3], np.float32), (n, 3))
    d = np.zeros((n, 3), np.float32)
    d[np.arange(n), rs.integers(0, 3, n)] = rs.choice([-1.0, 1.0], n)
    d[(d == 0) & (rs.random((n, 3)) < 0.5)] = -0.0
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = 3.402823466e38
    return rays


def main():
    if not po.ref_available():
        sys.exit("oracle/_ref/libwfref.so is not built (make -C oracle ref)")
    R = po.ref()
    rng = {"hash_u32": {}, "path_key": [], "rng": []}
    for x in [0, 1, 2, 3, 255, 256, 65535, 0x12345678, 0x9E3779B9, 0xFFFFFFFF, 1234]:
        rng["hash_u32"][str(x)] = int(R.ref_hash_u32(x))
    for seed, pixel, sample in [(1234, 0, 0), (1234, 1, 0), (1234, 0, 1), (1234, 921599, 63), (0, 0, 0),
                                (7, 5000, 17), (0xFFFFFFFF, 123456, 255)]:
        kk = (C.c_uint32 * 2)()
        R.ref_path_key(seed, pixel, sample, kk)
        rng["path_key"].append([seed, pixel, sample, int(kk[0]), int(kk[1])])
        for depth in [0, 1, 3, 7]:
            for dim in [0, 1, 2, 3, 4, 5, 6, 7, 8, 15]:
                v = R.ref_rng(kk[0], kk[1], depth, dim)
                rng["rng"].append([int(kk[0]), int(kk[1]), depth, dim, int(np.float32(v).view(np.uint32))])
    with open(os.path.join(OUT, "rng.json"), "w") as f:
        json.dump(rng, f, indent=1)

    scenes = {}
    renders, paths, inter = {}, {}, {}
    for name in SCENES:
        s = po.RefScene(name, 1280, 720)
        nodes, order = s.bvh()
        scenes[name] = {
            "hash": "%016x" % s.hash(),
            "nodes": int(len(nodes)),
            "nodes_fnv": "%016x" % fnv64(np.ascontiguousarray(nodes[:, :9]).tobytes()),
            "prim_order_fnv": "%016x" % fnv64(np.ascontiguousarray(order).tobytes()),
            "camera_rig_bits": [int(v) for v in s.camera_rig().view(np.uint32)],
        }
        small = po.RefScene(name, RW, RH)
        rgb, var = small.render_cpu(4, 1234, 8, 1)
        renders[name] = rgb
        pl = []
        for px, py, smp in PATHS:
            pl.append(small.trace_path(int(px), int(py), int(smp), 1234, 8))
        paths[name] = np.array(pl, np.float32)
        a = s.arrays()
        pts = a["vertices"] if len(a["vertices"]) else a["spheres"][:, :3]
        center = pts.mean(axis=0)
        radius = float(np.abs(pts - center).max()) + 1.0
        rays = seeded_rays(1024, 99, center, radius)
        t, p = s.intersect(rays, False)
        inter[name + "_rays"] = rays
        inter[name + "_t"] = t
        inter[name + "_prim"] = p
        ar = axis_rays(nodes, 2048, 17)
        t, p = s.intersect(ar, False)
        ta, pa = s.intersect(ar, True)
        inter[name + "_axis_rays"] = ar
        inter[name + "_axis_t"] = t
        inter[name + "_axis_prim"] = p
        inter[name + "_axis_any_t"] = ta
        inter[name + "_axis_any_prim"] = pa
    with open(os.path.join(OUT, "scenes.json"), "w") as f:
        json.dump(scenes, f, indent=1)
    np.savez_compressed(os.path.join(OUT, "ref_render.npz"), **renders)
    np.savez_compressed(os.path.join(OUT, "ref_paths.npz"), **paths)
    np.savez_compressed(os.path.join(OUT, "intersect.npz"), **inter)
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
