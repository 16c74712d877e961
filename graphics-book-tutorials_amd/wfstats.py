"""wfstats: the measurements behind the analysis figures, from real renders (the reference's
tools/wfstats.cpp, restated over this engine's C ABI), written as one JSON document.

  python graphics-book-tutorials_amd/wfstats.py OUT.json [section ...] [--quick]

Sections (all by default):
  convergence  RMSE of the megakernel (and of the CPU arm up to 64 spp) against a 16384-spp render
               of the Cornell box, 1..1024 spp, with the predicted error from the per-pixel
               variance and a log-log least-squares fit (tools/wfstats.cpp:49-78)
  furnace      furnace radiance at depth caps 1..16 on the CPU arm, the megakernel, the plain and
               the persistent wavefront, against the geometric series Le * sum rho^k (:81-100)
  variance     concentration of the per-pixel luminance variance: Lorenz curve, Gini coefficient,
               share of the top 1%, kurtosis (:103-121)
  survival     extension rays at depth caps 0..16 (:124-137)
  divergence   material mix of the primary hits seen by 32-pixel (16x2) tiles as in the reference,
               and by the 64-pixel (8x8) tiles that a gfx950 wave shades (:140-166)
  scaling      render time against pixel count for every GPU arm, with least-squares fits (:169-193)
  bvh          tree statistics: node and leaf counts, depth, SAH cost (root-area normalised),
               leaf depth and size histograms, child-to-parent area by depth (:196-225)
  rng          counter-based RNG quality: avalanche of the path key, autocorrelation, a
               Kolmogorov-Smirnov distance and 16x16 pair chi-square tests (:228-262)
--quick shrinks every render (for a smoke run); the full sizes are the reference's.
Every number comes from running this engine (GPU arms and the CPU arm) or its host-side code.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import wfamd  # noqa: E402

REF_SCENES = ("cornell", "divergence", "trianglehv")
_LUMA = np.array([0.2126, 0.7152, 0.0722])


def _lum(rgb):
    return np.asarray(rgb, np.float64).reshape(-1, 3) @ _LUMA


def ols(x, y):
    """y = a + b x by least squares; r2 and the standard error of b."""
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    mx, my = x.mean(), y.mean()
    sxx, sxy, syy = ((x - mx) ** 2).sum(), ((x - mx) * (y - my)).sum(), ((y - my) ** 2).sum()
    b = sxy / sxx
    a = my - b * mx
    r2 = sxy * sxy / (sxx * syy) if syy > 0 else 1.0
    rss = ((y - a - b * x) ** 2).sum()
    se = math.sqrt(rss / (len(x) - 2) / sxx) if len(x) > 2 else 0.0
    return {"a": float(a), "b": float(b), "r2": float(r2), "se_b": float(se)}


def _params(spp, seed, depth=8):
    return wfamd.params(spp=spp, seed=seed, max_depth=depth)


def _render(r, spp, seed, depth=8):
    rgb, var, st = r.render(_params(spp, seed, depth))
    return rgb, var, st


# ---- 1. convergence -----------------------------------------------------------------------------
def convergence(q):
    w, h = (40, 24) if q else (160, 90)
    ref_spp, top, reps, cpu_top = (256, 16, 2, 4) if q else (16384, 1024, 6, 64)
    sc = wfamd.Scene("cornell", w, h)
    mk = wfamd.Renderer(sc, "megakernel")
    ref_rgb, ref_var, _ = _render(mk, ref_spp, 99991)
    ref_l = _lum(ref_rgb)
    n = w * h
    floor_mse = float(ref_var.astype(np.float64).sum()) / n / ref_spp

    def err(rgb, var, spp):
        mse = float(((_lum(rgb) - ref_l) ** 2).mean())
        return math.sqrt(max(mse - floor_mse, 0.0)), math.sqrt(float(var.astype(np.float64).sum()) / n / spp)

    spps, rm, rs, pm, cm, cs = [], [], [], [], [], []
    spp = 1
    while spp <= top:
        e, pr = [], []
        for rep in range(reps):
            rgb, var, _ = _render(mk, spp, 100 + rep)
            a, b = err(rgb, var, spp)
            e.append(a)
            pr.append(b)
        spps.append(spp)
        rm.append(float(np.mean(e)))
        rs.append(float(np.std(e, ddof=1)) if len(e) > 1 else 0.0)
        pm.append(float(np.mean(pr)))
        if spp <= cpu_top:
            ce = []
            for rep in range(max(2, reps // 2)):
                rgb, var, _ = wfamd.render_cpu(sc, wfamd.params(spp=spp, seed=200 + rep))
                ce.append(err(rgb, var, spp)[0])
            cm.append(float(np.mean(ce)))
            cs.append(float(np.std(ce, ddof=1)))
        spp *= 2
    mk.close()
    return {"scene": "cornell", "width": w, "height": h, "reference_spp": ref_spp,
            "reference_floor_rmse": math.sqrt(floor_mse), "spp": spps, "rmse_mean": rm, "rmse_sd": rs,
            "predicted": pm, "cpu_rmse_mean": cm, "cpu_rmse_sd": cs,
            "loglog_fit": ols(np.log2(spps), np.log2(rm))}


# ---- 2. furnace ---------------------------------------------------------------------------------
def furnace(q):
    w = 16 if q else 64
    spp = 16 if q else 64
    sc = wfamd.Scene("furnace", w, w)
    mat = sc.arrays()["materials"][0]
    rho, le = float(mat[4]), float(mat[8])
    arms = {m: wfamd.Renderer(sc, m) for m in ("megakernel", "wavefront", "persistent")}
    out = {"albedo": rho, "emission": le, "limit": le / (1.0 - rho), "depth": []}
    for k in ("cpu", "megakernel", "wavefront", "persistent"):
        out[k], out[k + "_se"] = [], []
    out["series_to_depth_minus_1"], out["series_to_depth"] = [], []

    def summ(rgb, var, spp_):
        return float(_lum(rgb).mean()), math.sqrt(float(var.astype(np.float64).sum()) / spp_) / (w * w)

    for d in range(1, 17):
        rgb, var, _ = wfamd.render_cpu(sc, wfamd.params(spp=spp, seed=7, max_depth=d))
        m, se = summ(rgb, var, spp)
        out["cpu"].append(m)
        out["cpu_se"].append(se)
        for name, r in arms.items():
            rgb, var, _ = _render(r, spp, 7, d)
            m, se = summ(rgb, var, spp)
            out[name].append(m)
            out[name + "_se"].append(se)
        s0 = sum(rho ** k for k in range(d))
        out["depth"].append(d)
        out["series_to_depth_minus_1"].append(le * s0)
        out["series_to_depth"].append(le * (s0 + rho ** d))
    for r in arms.values():
        r.close()
    return out


# ---- 3. variance concentration ------------------------------------------------------------------
def lorenz_stats(v):
    v = np.sort(np.asarray(v, np.float64))
    n = len(v)
    tot = v.sum()
    cum = np.concatenate([[0.0], np.cumsum(v)])
    lor = [float(cum[int(n * k / 100.0)] / tot) if tot > 0 else 0.0 for k in range(101)]
    gini = 1.0 - 2.0 * sum((lor[k] + lor[k + 1]) / 2 * 0.01 for k in range(100))
    top1 = float(v[n - n // 100:].sum() / tot) if tot > 0 else 0.0
    mu = tot / n
    d = v - mu
    m2, m4 = float((d ** 2).mean()), float((d ** 4).mean())
    return {"lorenz": lor, "gini": gini, "top1pct_share": top1, "mean_var": mu, "median_var": float(v[n // 2]),
            "max_over_mean": float(v[-1] / mu) if mu > 0 else 0.0, "kurtosis": m4 / (m2 * m2) if m2 > 0 else 0.0}


def variance(q):
    w, h, spp = (80, 45, 16) if q else (320, 180, 64)
    out = {}
    for name in REF_SCENES:
        r = wfamd.Renderer(wfamd.Scene(name, w, h), "megakernel")
        _, var, _ = _render(r, spp, 5)
        r.close()
        out[name] = lorenz_stats(var)
    return out


# ---- 4. survival --------------------------------------------------------------------------------
def survival(q):
    w, h, spp = (80, 45, 2) if q else (320, 180, 8)
    out = {}
    for name in REF_SCENES:
        r = wfamd.Renderer(wfamd.Scene(name, w, h), "megakernel")
        e = [int(_render(r, spp, 3, d)[2].extension_rays) for d in range(17)]
        r.close()
        out[name] = {"paths": w * h * spp, "extension_rays_by_cap": e}
    return out


# ---- 5. divergence ------------------------------------------------------------------------------
def _hash_u32(x):
    x = np.asarray(x, np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def path_key(seed, pixel, sample):
    """The engine's path key (wf_core.h make_path_key) for arrays of pixel/sample indices."""
    m = 0xFFFFFFFF
    seed, pixel, sample = (np.asarray(v, np.uint64) & m for v in (seed, pixel, sample))
    a = _hash_u32(seed ^ 0x9E3779B9)
    k0 = _hash_u32(_hash_u32((a + pixel * 0xC2B2AE35) & m) ^ ((sample * 0x27D4EB2F + 0x165667B1) & m))
    b = _hash_u32((seed + 0x85EBCA6B) & m)
    k1 = _hash_u32((_hash_u32(b ^ ((pixel * 0x9E3779B1) & m)) + ((sample * 0x85EBCA77 + 0xC2B2AE3D) & m)) & m)
    return k0, k1


def rng(k0, k1, depth, dim):
    """The engine's draw (wf_core.h rng): uniform on the 2^-24 grid."""
    m = 0xFFFFFFFF
    ctr = (np.asarray(depth, np.uint64) * 16 + dim) & m
    h = _hash_u32(np.asarray(k0, np.uint64) ^ _hash_u32((ctr * 0x9E3779B9 + np.asarray(k1, np.uint64)) & m))
    return (h >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)


def primary_categories(scene, seed=1234):
    """Material type of each pixel's first hit (-1: miss) for the jittered camera rays of sample 0."""
    i = scene.info()
    w, h = i.width, i.height
    rig = scene.camera_rig().astype(np.float32)
    org, ll, hor, ver = rig[0:3], rig[3:6], rig[6:9], rig[9:12]
    py, px = np.divmod(np.arange(w * h, dtype=np.int64), w)
    k0, k1 = path_key(seed, py * w + px, 0)
    jx, jy = rng(k0, k1, 0, 0), rng(k0, k1, 0, 1)
    u = (px.astype(np.float32) + jx) / np.float32(w)
    v = ((h - 1 - py).astype(np.float32) + jy) / np.float32(h)
    d = ll + hor * u[:, None] + ver * v[:, None] - org
    d = d / np.sqrt((d * d).sum(axis=1, dtype=np.float32))[:, None]
    rays = np.zeros((w * h, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    _, prim = wfamd.test_intersect(scene, rays)
    a = scene.arrays()
    mats = a["materials"][:, 0].view(np.int32)
    ns = len(a["spheres"])
    pm = np.concatenate([a["spheres"][:, 4].view(np.int32), a["triangles"][:, 3]])
    cat = np.where(prim >= 0, mats[pm[np.maximum(prim, 0)]], -1)
    return cat.reshape(h, w), ns


def tile_stats(cat, tw, th):
    h, w = cat.shape
    kinds, counts = np.unique(cat, return_counts=True)
    p = counts / cat.size
    lanes = tw * th
    distinct = [len(np.unique(cat[y:y + th, x:x + tw])) for y in range(0, h - th + 1, th)
                for x in range(0, w - tw + 1, tw)]
    return {"categories": kinds.tolist(), "probability": p.tolist(),
            "entropy_bits": float(-(p * np.log2(p)).sum()),
            "expected_distinct_iid": float((1 - (1 - p) ** lanes).sum()),
            "mean_distinct_per_tile": float(np.mean(distinct)), "tile": [tw, th]}


def divergence(q):
    w, h = (80, 45) if q else (320, 180)
    out = {}
    for name in REF_SCENES + ("sphere_grid",):
        cat, _ = primary_categories(wfamd.Scene(name, w, h))
        out[name] = {"warp32": tile_stats(cat, 16, 2), "wave64": tile_stats(cat, 8, 8)}
        out[name]["mean_distinct_per_warp"] = out[name]["warp32"]["mean_distinct_per_tile"]
    return out


# ---- 6. scaling ---------------------------------------------------------------------------------
def scaling(q):
    widths = (128, 256, 512) if q else (128, 192, 256, 384, 512, 640, 768, 1024, 1280, 1600, 1920)
    out = {}
    for name in ("cornell", "divergence"):
        px, times = [], {m: [] for m in ("megakernel", "wavefront", "persistent")}
        for w in widths:
            h = (w * 9 // 16) & ~1
            sc = wfamd.Scene(name, w, h)
            for m in times:
                r = wfamd.Renderer(sc, m)
                _render(r, 4, 11)
                ms = [_render(r, 4, 11)[2].render_ms for _ in range(3 if q else 5)]
                r.close()
                times[m].append(float(np.median(ms)))
            px.append(w * h / 1e6)
        out[name] = {"megapixels": px, "spp": 4, **{f"{m}_ms": t for m, t in times.items()},
                     **{f"{m}_fit": ols(px, t) for m, t in times.items()}}
    return out


# ---- 7. BVH -------------------------------------------------------------------------------------
def _area(lo, hi):
    d = hi - lo
    return np.float32(2.0) * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def bvh_tree_stats(nodes):
    """nodes: the reference layout (wf_scene_bvh: lo xyz bits, skip, hi xyz bits, prim_start,
    prim_count, ...); the left child of node N is N + 1 and the right child is its skip target."""
    lo = nodes[:, 0:3].view(np.float32)
    hi = nodes[:, 4:7].view(np.float32)
    skip, cnt = nodes[:, 3], nodes[:, 8]
    root = float(_area(lo[0], hi[0]))
    depth_hist, size_hist = {}, [0, 0, 0, 0]
    ratio_sum, ratio_cnt = {}, {}
    sah, max_depth, internal, leaves = 0.0, 0, 0, 0
    stack = [(0, 0)]
    while stack:
        i, dpt = stack.pop()
        a = float(_area(lo[i], hi[i])) / root
        max_depth = max(max_depth, dpt)
        if cnt[i] > 0:
            leaves += 1
            sah += a * int(cnt[i])
            depth_hist[min(dpt, 63)] = depth_hist.get(min(dpt, 63), 0) + 1
            size_hist[min(int(cnt[i]), 4) - 1] += 1
        else:
            internal += 1
            sah += a
            pa = max(float(_area(lo[i], hi[i])), 1e-20)
            for c in (i + 1, int(skip[i + 1])):
                ratio_sum[dpt] = ratio_sum.get(dpt, 0.0) + float(_area(lo[c], hi[c])) / pa
                ratio_cnt[dpt] = ratio_cnt.get(dpt, 0) + 1
                stack.append((c, dpt + 1))
    return {"nodes": len(nodes), "internal": internal, "leaves": leaves, "max_depth": max_depth, "sah_cost": sah,
            "leaf_depth_hist": [depth_hist.get(d, 0) for d in range(max_depth + 1)], "leaf_size_hist": size_hist,
            "child_to_parent_area_by_depth": [ratio_sum[d] / ratio_cnt[d] if ratio_cnt.get(d) else 0.0
                                              for d in range(max_depth)]}


def bvh(q):
    out = {}
    for name in REF_SCENES + (() if q else ("cornell555", "sphere_grid", "heightfield")):
        s = wfamd.Scene(name, 1280, 720)
        nodes, _ = s.bvh()
        i = s.info()
        st = bvh_tree_stats(nodes)
        st["primitives"] = i.n_spheres + i.n_triangles
        st["sah_per_primitive"] = st["sah_cost"] / st["primitives"]
        out[name] = st
    return out


# ---- 8. RNG -------------------------------------------------------------------------------------
def rng_stats(q):
    trials = 2000 if q else 20000
    state = np.random.default_rng(88172645463325252)
    inp = state.integers(0, 1 << 32, size=(trials, 3), dtype=np.uint64)

    def key64(s, p, smp):
        k0, k1 = path_key(s, p, smp)
        return (k1 << np.uint64(32)) | k0

    base = key64(inp[:, 0], inp[:, 1], inp[:, 2])
    flip = np.zeros((96, 64))
    for bit in range(96):
        m = inp.copy()
        m[:, bit // 32] ^= np.uint64(1 << (bit % 32))
        d = key64(m[:, 0], m[:, 1], m[:, 2]) ^ base
        flip[bit] = ((d[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).mean(axis=0)
    n = 1 << (16 if q else 20)
    idx = np.arange(n, dtype=np.uint64)
    k0, k1 = path_key(1234, idx, 0)
    u = rng(k0, k1, 0, 5).astype(np.float64)
    mu, var = u.mean(), u.var()
    acf = [float(((u[:-lag] - mu) * (u[lag:] - mu)).mean() / var) for lag in range(1, 65)]
    s = np.sort(u)
    i = np.arange(0, n, 16)
    ks = float(max(np.abs((i + 1.0) / n - s[i]).max(), np.abs(i / n - s[i]).max()))
    B = 16
    chi2 = []
    for seed in range(1, 33):
        a0, a1 = path_key(seed, idx, 0)
        v = rng(a0, a1, 0, 5)
        cell = np.minimum(B - 1, (v[:-1] * B).astype(np.int64)) * B + np.minimum(B - 1, (v[1:] * B).astype(np.int64))
        c = np.bincount(cell, minlength=B * B).astype(np.float64)
        e = (n - 1) / (B * B)
        chi2.append(float(((c - e) ** 2 / e).sum()))
    return {"trials": trials, "avalanche": flip.ravel().tolist(), "avalanche_mean": float(flip.mean()),
            "avalanche_max_dev": float(np.abs(flip - 0.5).max()), "avalanche_se": 0.5 / math.sqrt(trials),
            "stream_samples": n, "acf": acf, "acf_band95": 1.96 / math.sqrt(n), "ks_d": ks,
            "ks_critical_5pct": 1.358 / math.sqrt(n), "pair_chi2": chi2[0], "pair_chi2_by_seed": chi2,
            "pair_dof": B * B - 1}


SECTIONS = {"convergence": convergence, "furnace": furnace, "variance": variance, "survival": survival,
            "divergence": divergence, "scaling": scaling, "bvh": bvh, "rng": rng_stats}


def run(sections, quick=False, log=sys.stderr):
    doc = {"schema_version": 1}
    t0 = time.perf_counter()
    for name in sections:
        doc[name] = SECTIONS[name](quick)
        print(f"{name} done ({time.perf_counter() - t0:.1f} s)", file=log, flush=True)
    try:
        doc["device"] = wfamd.query_device()["name"]
    except Exception:   # a CPU-only run (bvh, rng) on a machine without a GPU
        doc["device"] = "cpu"
    return doc


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("out")
    ap.add_argument("sections", nargs="*")
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args(argv)
    unknown = [s for s in a.sections if s not in SECTIONS]
    if unknown:
        print(f"Error: unknown section(s) {unknown}; sections: {' '.join(SECTIONS)}", file=sys.stderr)
        return 2
    wfamd.lib()
    doc = run(a.sections or list(SECTIONS), a.quick)
    text = json.dumps(doc)
    with open(a.out, "w") as f:
        f.write(text + "\n")
    print(f"wrote {a.out} ({len(text) + 1} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
