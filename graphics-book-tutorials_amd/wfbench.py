"""Benchmark driver for the five benchmark configurations.

  python graphics-book-tutorials_amd/wfbench.py --config N [--reps R] [--warmup W] [--out FILE] [--quick]

Configs 0-2 follow the reference's wfbench protocol (src/bench.cpp:14-92,95-196): warm-up renders,
`reps` timed renders per condition, linear-interpolation statistics (finalize_condition), quality
checks and the verdict rule (apply_verdict), written as the reference's bench document, schema
version 1 (src/bench.cpp:198-330): conditions megakernel and wavefront, the same keys. Changes:
  - every quality check is measured (the reference writes fixed values, src/bench.cpp:181-189):
    megakernel vs wavefront maximum relative difference, and for the CPU arm the fraction of
    pixels within 3 combined standard errors of each GPU arm on the same samples;
  - the persistent batched wavefront is a third condition, reported under the key "mi355x" (with
    its speedup over the megakernel, the bitwise check and Mrays/s of every condition), so the
    schema-v1 part stays valid for the reference's validator (src/bench.cpp:360-441).
  config 0  Cornell box 320x180, 16 spp: also times the CPU arm (the config names it).
  config 1  sphere_grid 1280x720, 64 spp (material-divergent).
  config 2  heightfield 1280x720, 64 spp (2.1 M triangles).
Config 3 renders 64 jobs of random scene, orbit angle, resolution and spp in one persistent
launch (autotuned) and compares with rendering the same views one by one on the megakernel; every
job's image must equal its standalone render bit for bit.
Config 4 is the serving loop: Poisson arrivals at a fixed rate (60 requests/s), requests of
random scene, orbit angle, resolution and spp under a device-memory budget; it reports Mrays/s and
p50/p99 latency (arrival to result) and checks that requests give the same bits alone and in the mix.
Every random choice comes from a seeded generator, and every pinned value is in PROTOCOL.
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

# Persistent-engine tuning fields (wf_render_params) the autotuner chooses.
TUNING_KEYS = ("block_size", "pool_paths", "lds_kib", "refill", "leaf_batch", "dir_bins", "shade_block")

# Pinned protocol (the reference pins its values in bench/protocol.json).
PROTOCOL = {
    "schema_version": 1,
    "speedup_target": 1.30,            # include/wf/bench.hpp:12 (wavefront vs megakernel)
    "vram_budget_bytes": 3.5e9,        # include/wf/bench.hpp:13
    "se_pixel_fraction": 0.999,        # stricter than include/wf/bench.hpp:14 (0.995)
    "bitwise_rel_tolerance": 1e-5,     # include/wf/bench.hpp:16
    "seed": 1234,
    "max_depth": 8,
    "configs": {
        "0": {"scene": "cornell555", "width": 320, "height": 180, "spp": 16, "reps": 20, "warmup": 2, "cpu_spp": 16},
        "1": {"scene": "sphere_grid", "width": 1280, "height": 720, "spp": 64, "reps": 20, "warmup": 2, "cpu_spp": 2},
        "2": {"scene": "heightfield", "width": 1280, "height": 720, "spp": 64, "reps": 20, "warmup": 2, "cpu_spp": 1},
        "3": {"jobs": 64, "job_seed": 2025, "scenes": ["cornell555", "sphere_grid", "heightfield"],
              "width": [128, 640], "height": [128, 360], "spp": [1, 4, 16, 64], "reps": 5, "warmup": 1},
        "4": {"requests": 48, "request_seed": 4242, "scenes": ["cornell555", "sphere_grid", "heightfield"],
              "width": [256, 1920], "height": [256, 1080], "spp": [4, 256], "rate_per_s": 60.0,
              "budget_bytes": 8 << 30, "check_smallest": 6},
    },
    "quick": {"width": 96, "height": 54, "spp": 4, "reps": 3, "warmup": 1, "jobs": 8, "requests": 8},
}


def percentile(v, p):
    """Linear interpolation between order statistics at index (n-1)*p (src/bench.cpp:14-37)."""
    s = sorted(v)
    x = (len(s) - 1) * p
    lo = int(x)
    if lo + 1 >= len(s):
        return s[-1]
    return s[lo] * (1.0 - (x - lo)) + s[lo + 1] * (x - lo)


def finalize_condition(c, width, height, spp):
    """min, median, p10, p90, max, spread and samples/s of c["run_ms"] (src/bench.cpp:14-37)."""
    if not c["run_ms"]:
        raise ValueError("finalize_condition: empty run_ms")
    s = sorted(c["run_ms"])
    c["min"], c["max"] = s[0], s[-1]
    c["median"] = percentile(s, 0.5)
    c["p10"] = percentile(s, 0.10)
    c["p90"] = percentile(s, 0.90)
    c["spread_rel"] = (c["p90"] - c["p10"]) / c["median"]
    c["samples_per_second"] = width * height * spp / (c["median"] / 1000.0)
    return c


def apply_verdict(r):
    """The reference's verdict rule (src/bench.cpp:39-92): supported iff the wavefront is at least
    speedup_target faster than the megakernel, every condition is within the VRAM budget, every
    quality check passes and no condition overflowed, hit invalid geometry or produced non-finite
    pixels."""
    r["verdict"] = "hypothesis_weakened"
    reasons = []
    mode = {c["mode"]: c for c in r["conditions"]}
    if "megakernel" in mode and "wavefront" in mode:
        r["speedup"] = mode["megakernel"]["median"] / mode["wavefront"]["median"]
        r["speedup_meets_target"] = r["speedup"] >= PROTOCOL["speedup_target"]
        if not r["speedup_meets_target"]:
            reasons.append("speedup below 1.30")
    else:
        r["speedup"] = 0.0
        r["speedup_meets_target"] = False
        reasons.append("missing condition")
    r["vram_within_budget"] = all(int(c["peak_vram_bytes"]) <= PROTOCOL["vram_budget_bytes"]
                                  for c in r["conditions"])
    if not r["vram_within_budget"]:
        reasons.append("peak VRAM exceeds 3.5 GB")
    reasons += ["quality check failed: " + q["name"] for q in r["quality"] if not q["pass"]]
    for key, why in (("queue_overflows", "queue overflow"), ("invalid_hits", "invalid hits"),
                     ("nonfinite_pixels", "non-finite pixels")):
        if any(c[key] > 0 for c in r["conditions"]):
            reasons.append(why)
    r["reasons"] = reasons
    if r["speedup_meets_target"] and r["vram_within_budget"] and not reasons:
        r["verdict"] = "hypothesis_supported"
    return r


BENCH_KEYS = ("scene_name", "scene_hash", "width", "height", "spp", "max_depth", "reps", "warmup", "seed",
              "device_name", "driver_version", "cuda_runtime_version", "conditions", "quality", "speedup",
              "speedup_meets_target", "vram_within_budget", "verdict", "reasons")
CONDITION_KEYS = ("mode", "run_ms", "min", "median", "p10", "p90", "max", "spread_rel", "samples_per_second",
                  "rays_per_second", "peak_vram_bytes", "queue_high_water", "queue_overflows", "invalid_hits",
                  "nonfinite_pixels")


def validate_bench_doc(doc):
    """Checks a bench document (a parsed JSON object) against schema version 1 and the verdict rule
    (src/bench.cpp:369-441). Returns "" when valid, else the name of the first offending key or
    check ("verdict" when the verdict does not follow from the document's own fields)."""
    def num(v):
        return isinstance(v, (int, float)) and not isinstance(v, bool)
    if not isinstance(doc, dict):
        return "root"
    if not num(doc.get("schema_version")) or int(doc["schema_version"]) != PROTOCOL["schema_version"]:
        return "schema_version"
    for k in BENCH_KEYS:
        if k not in doc:
            return k
    h = doc["scene_hash"]
    if not isinstance(h, str) or len(h) != 16 or any(c not in "0123456789abcdef" for c in h):
        return "scene_hash"
    if not num(doc["speedup"]):
        return "speedup"
    for k in ("speedup_meets_target", "vram_within_budget"):
        if not isinstance(doc[k], bool):
            return k
    conds = doc["conditions"]
    if not isinstance(conds, list) or len(conds) != 2:
        return "conditions"
    if not isinstance(doc["quality"], list):
        return "quality"
    if not isinstance(doc["reasons"], list):
        return "reasons"
    if conds[0].get("mode") != "megakernel" or conds[1].get("mode") != "wavefront":
        return "mode"
    for c in conds:
        for k in CONDITION_KEYS:
            if k not in c:
                return k
        if not isinstance(c["run_ms"], list) or not c["run_ms"] or not all(num(e) for e in c["run_ms"]):
            return "run_ms"
        if not (c["min"] <= c["p10"] <= c["median"] <= c["p90"] <= c["max"]):
            return "stats"
    for q in doc["quality"]:
        for k in ("name", "pass", "value", "threshold", "detail"):
            if k not in q:
                return k
        if not isinstance(q["pass"], bool):
            return "pass"
    qual_ok = all(q["pass"] for q in doc["quality"])
    valid_ok = all(c["queue_overflows"] == 0 and c["invalid_hits"] == 0 and c["nonfinite_pixels"] == 0 for c in conds)
    speedup_ok = doc["speedup"] >= PROTOCOL["speedup_target"]
    all_ok = speedup_ok and qual_ok and doc["vram_within_budget"] and valid_ok and not doc["reasons"]
    if doc["verdict"] == "hypothesis_supported":
        return "" if all_ok else "verdict"
    if doc["verdict"] == "hypothesis_weakened":
        return "verdict" if all_ok else ""
    return "verdict"


def nonfinite_pixels(rgb):
    return int((~np.isfinite(rgb.reshape(-1, 3))).any(axis=1).sum())


def image_diff(rgb_a, var_a, spp_a, rgb_b, var_b, spp_b):
    """max relative channel difference and fraction of pixels whose luminance differs by at most 3
    combined standard errors (compare_results, src/film.cpp:45-107)."""
    a, b = rgb_a.reshape(-1, 3).astype(np.float64), rgb_b.reshape(-1, 3).astype(np.float64)
    m = np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-6)
    rel = float(np.max(np.abs(a - b) / m)) if a.size else 0.0
    w = np.array([0.2126, 0.7152, 0.0722])
    la, lb = a @ w, b @ w
    se = np.sqrt(var_a.astype(np.float64) / spp_a + var_b.astype(np.float64) / spp_b)
    within = float(np.mean(np.abs(la - lb) <= 3.0 * se)) if la.size else 1.0
    same = float(np.mean((rgb_a.reshape(-1, 3).view(np.uint32) == rgb_b.reshape(-1, 3).view(np.uint32)).all(axis=1)))
    return rel, within, same


def run_condition(scene, mode, params, reps, warmup, w, h):
    r = wfamd.Renderer(scene, mode)
    for _ in range(warmup):
        r.render(params)
    c = {"mode": mode, "run_ms": []}
    first = None
    rays = 0
    for k in range(reps):
        rgb, var, st = r.render(params)
        c["run_ms"].append(st.render_ms)
        if k == 0:
            first = (rgb, var)
            rays = (st.extension_rays, st.shadow_rays)
            c["peak_vram_bytes"] = str(st.peak_vram_bytes)
            c["queue_high_water"] = [int(x) for x in st.queue_high_water]
            c["queue_overflows"] = int(st.queue_overflows)
            c["invalid_hits"] = int(st.invalid_hits)
            c["nonfinite_pixels"] = nonfinite_pixels(rgb)
    r.close()
    finalize_condition(c, w, h, params.spp)
    # rays_per_second: extension rays per second, as the reference (src/bench.cpp:139-145).
    c["rays_per_second"] = rays[0] / (c["median"] / 1000.0)
    extra = {"mrays_per_s": (rays[0] + rays[1]) / (c["median"] * 1e3), "extension_rays": rays[0],
             "shadow_rays": rays[1], "p99_ms": percentile(c["run_ms"], 0.99)}
    return c, first, extra


def stage_profile(scene, mode, spp, seed, depth, tuned=None):
    """One stage-timed render (HIP events around every stage, synchronising): per stage ms,
    items, launches, algorithmic device-memory bytes and the GB/s they imply (the reference's
    stage_timing run, src/bench.cpp:166-172)."""
    kw = {}
    if tuned is not None:
        kw = {k: getattr(tuned, k) for k in TUNING_KEYS}
    r = wfamd.Renderer(scene, mode)
    _, _, st = r.render(wfamd.params(spp=spp, seed=seed, max_depth=depth, stage_timing=True, **kw))
    r.close()
    out = {}
    for k, v in st.stages().items():
        v = dict(v)
        v["gbs"] = v["bytes"] / (v["ms"] / 1e3) / 1e9 if v["ms"] > 0 and v["bytes"] else None
        out[k] = v
    return out


def bench_scene(cfg_id, cfg, quick):
    w, h, spp, reps, warmup = cfg["width"], cfg["height"], cfg["spp"], cfg["reps"], cfg["warmup"]
    if quick:
        q = PROTOCOL["quick"]
        w, h, spp, reps, warmup = q["width"], q["height"], q["spp"], q["reps"], q["warmup"]
    seed, depth = PROTOCOL["seed"], PROTOCOL["max_depth"]
    scene = wfamd.Scene(cfg["scene"], w, h)
    dev = wfamd.query_device()
    p = wfamd.params(spp=spp, seed=seed, max_depth=depth)
    mk, mk_img, mk_x = run_condition(scene, "megakernel", p, reps, warmup, w, h)
    wv, wv_img, wv_x = run_condition(scene, "wavefront", p, reps, warmup, w, h)
    tuned = wfamd.autotune([(scene, spp, seed, depth)], probe_spp=min(spp, 16))
    pp = wfamd.params(spp=spp, seed=seed, max_depth=depth, **{k: getattr(tuned, k) for k in TUNING_KEYS})
    pe, pe_img, pe_x = run_condition(scene, "persistent", pp, reps, warmup, w, h)
    quality = []
    rel, _, _ = image_diff(mk_img[0], mk_img[1], spp, wv_img[0], wv_img[1], spp)
    quality.append({"name": "megakernel_vs_wavefront_bitwise", "pass": rel <= PROTOCOL["bitwise_rel_tolerance"],
                    "value": rel, "threshold": PROTOCOL["bitwise_rel_tolerance"],
                    "detail": "max relative channel difference"})
    # CPU arm against each GPU arm on the same samples 0..cpu_spp-1.
    cpu_spp = min(cfg["cpu_spp"], spp)
    cpu_x = None
    if cpu_spp > 0:
        t0 = time.perf_counter()
        (cr, cl, cl2), cst = wfamd.render_cpu_sums(scene, 0, cpu_spp, wfamd.params(spp=cpu_spp, seed=seed,
                                                                                 max_depth=depth))
        cpu_s = time.perf_counter() - t0
        crgb, cvar = wfamd.finalize(cr, cl, cl2, cpu_spp)
        cpu_x = {"render_s": cpu_s, "spp": cpu_spp, "threads": os.cpu_count(),
                 "mrays_per_s": (cst.extension_rays + cst.shadow_rays) / cpu_s / 1e6}
        for mode in ("megakernel", "wavefront"):
            r = wfamd.Renderer(scene, mode)
            grgb, gvar, _ = r.render(wfamd.params(spp=cpu_spp, seed=seed, max_depth=depth))
            r.close()
            _, within, same = image_diff(crgb, cvar, cpu_spp, grgb, gvar, cpu_spp)
            quality.append({"name": f"cpu_vs_{mode}", "pass": within >= PROTOCOL["se_pixel_fraction"],
                            "value": within, "threshold": PROTOCOL["se_pixel_fraction"],
                            "detail": f"{cpu_spp} spp, pixels within 3 SE; bitwise-equal pixels {same:.6f}"})
    prel, _, psame = image_diff(mk_img[0], mk_img[1], spp, pe_img[0], pe_img[1], spp)
    stages = {"wavefront": stage_profile(scene, "wavefront", spp, seed, depth),
              "persistent": stage_profile(scene, "persistent", spp, seed, depth, tuned)}
    doc = {"schema_version": PROTOCOL["schema_version"], "scene_name": cfg["scene"],
           "scene_hash": f"{scene.info().hash:016x}",
           "width": w, "height": h, "spp": spp, "max_depth": depth, "reps": reps, "warmup": warmup, "seed": seed,
           "device_name": dev["name"], "driver_version": str(dev["driver_version"]),
           "cuda_runtime_version": str(dev["runtime_version"]), "conditions": [mk, wv], "quality": quality}
    apply_verdict(doc)
    doc["mi355x"] = {
        "config": int(cfg_id), "persistent": pe,
        "persistent_speedup_vs_megakernel": mk["median"] / pe["median"],
        "persistent_equals_megakernel_bitwise": psame == 1.0 and prel == 0.0,
        "mrays_per_s": {"megakernel": mk_x, "wavefront": wv_x, "persistent": pe_x},
        "tuning": {k: getattr(tuned, k) for k in TUNING_KEYS},
        "cpu_arm": cpu_x,
        "stages": stages,
    }
    return doc


def log_uniform(rng, lo, hi):
    return float(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def make_views(rng, n, cfg, quick):
    """(scene name, width, height, yaw, spp) per job; seeded."""
    out = []
    for _ in range(n):
        name = cfg["scenes"][int(rng.integers(len(cfg["scenes"])))]
        yaw = float(rng.uniform(0.0, 360.0))
        w = int(round(log_uniform(rng, *cfg["width"])))
        h = int(round(log_uniform(rng, *cfg["height"])))
        if isinstance(cfg["spp"], list) and len(cfg["spp"]) == 4:
            spp = int(cfg["spp"][int(rng.integers(4))])
        else:
            spp = int(round(log_uniform(rng, *cfg["spp"])))
        if quick:
            w, h, spp = max(8, w // 8), max(8, h // 8), min(spp, 4)
        out.append((name, w, h, yaw, spp))
    return out


def bench_batch(cfg, quick):
    n = PROTOCOL["quick"]["jobs"] if quick else cfg["jobs"]
    rng = np.random.default_rng(cfg["job_seed"])
    views = make_views(rng, n, cfg, quick)
    base = {s: wfamd.Scene(s, 64, 36) for s in cfg["scenes"]}
    specs = [(base[s], spp, PROTOCOL["seed"] + k, PROTOCOL["max_depth"], w, h, yaw)
             for k, (s, w, h, yaw, spp) in enumerate(views)]
    t0 = time.perf_counter()
    tuned = wfamd.autotune(specs, probe_spp=4)
    tune_s = time.perf_counter() - t0
    runs, outs, st = [], None, None
    for k in range(cfg["warmup"] + cfg["reps"]):
        t0 = time.perf_counter()
        outs, st = wfamd.batch_render(specs, tuned)
        wall = (time.perf_counter() - t0) * 1e3
        if k >= cfg["warmup"]:
            runs.append((st.render_ms, wall))
    rays = st.extension_rays + st.shadow_rays
    med = percentile([r[0] for r in runs], 0.5)
    # The same views one by one on the megakernel (each view's own scene, as the reference builds
    # one renderer per camera), and every batched image against its standalone render.
    seq_ms, mk_rays, mismatches = 0.0, 0, []
    for k, ((s, w, h, yaw, spp), spec, (rgb, var)) in enumerate(zip(views, specs, outs)):
        v = base[s].view(w, h, yaw)
        r = wfamd.Renderer(v, "megakernel")
        p = wfamd.params(spp=spp, seed=spec[2], max_depth=PROTOCOL["max_depth"])
        r.render(p)
        want, wvar, mst = r.render(p)
        r.close()
        seq_ms += mst.render_ms
        mk_rays += mst.extension_rays + mst.shadow_rays
        if not (np.array_equal(rgb.view(np.uint32), want.view(np.uint32)) and
                np.array_equal(var.view(np.uint32), wvar.view(np.uint32))):
            mismatches.append(k)
    pixels = sum(w * h for _, w, h, _, _ in views)
    paths = sum(w * h * spp for _, w, h, _, spp in views)
    return {"config": 3, "jobs": n, "pixels": pixels, "paths": paths, "rays": rays,
            "batch_render_ms": {"median": med, "runs": [r[0] for r in runs], "wall_ms": [r[1] for r in runs]},
            "mrays_per_s": rays / (med * 1e3),
            "megakernel_one_by_one_ms": seq_ms, "megakernel_mrays_per_s": mk_rays / (seq_ms * 1e3),
            "speedup_vs_megakernel_one_by_one": seq_ms / med,
            "rays_equal": rays == mk_rays, "jobs_bitwise_equal_standalone": not mismatches, "mismatched_jobs": mismatches,
            "queue_overflows": st.queue_overflows, "invalid_hits": st.invalid_hits, "iterations": st.iterations,
            "autotune_s": tune_s,
            "tuning": {k: getattr(tuned, k) for k in TUNING_KEYS},
            "views": [{"scene": s, "width": w, "height": h, "yaw": yaw, "spp": spp} for s, w, h, yaw, spp in views]}


def serve(views, base, budget, arrivals, tuning):
    """Runs the serving loop over the requests; arrivals (seconds from start) or None = all at
    once. Returns per-request latency (ms), makespan (s), rays and the server."""
    srv = wfamd.Server(budget_bytes=budget, tuning=tuning)
    n = len(views)
    ids, t_arr, t_done = [None] * n, [0.0] * n, [None] * n
    nxt = 0
    peak = 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while nxt < n and (arrivals is None or arrivals[nxt] <= now):
            s, w, h, yaw, spp = views[nxt]
            ids[nxt] = srv.submit(base[s], spp, PROTOCOL["seed"] + nxt, PROTOCOL["max_depth"], w, h, yaw)
            t_arr[nxt] = now if arrivals is None else arrivals[nxt]
            nxt += 1
        srv.step(1)
        now = time.perf_counter() - t0
        for k in range(nxt):
            if t_done[k] is None and srv.poll(ids[k]) == 1:
                t_done[k] = now
        peak = max(peak, srv.stats()["in_flight"])
        if nxt == n and all(t is not None for t in t_done):
            break
        if nxt < n and arrivals is not None and srv.stats()["in_flight"] == 0:
            time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.002)))
    makespan = time.perf_counter() - t0
    st = srv.stats()
    lat = [(t_done[k] - t_arr[k]) * 1e3 for k in range(n)]
    return lat, makespan, st, srv, ids, peak


def bench_serving(cfg, quick):
    n = PROTOCOL["quick"]["requests"] if quick else cfg["requests"]
    rng = np.random.default_rng(cfg["request_seed"])
    views = make_views(rng, n, cfg, quick)
    base = {s: wfamd.Scene(s, 64, 36) for s in cfg["scenes"]}
    tuning = wfamd.params()
    # Capacity: every request submitted at once (closed loop).
    lat0, span0, st0, srv0, _, _ = serve(views, base, cfg["budget_bytes"], None, tuning)
    srv0.close()
    capacity = n / span0
    # A fixed arrival rate (PROTOCOL), or a fraction of the capacity just measured.
    rate = cfg["rate_per_s"] if "rate_per_s" in cfg else cfg["load"] * capacity
    if quick:
        rate = 0.75 * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n)).tolist()
    lat, span, st, srv, ids, peak = serve(views, base, cfg["budget_bytes"], arrivals, tuning)
    # Alone vs in the mix: the smallest requests rendered standalone on the megakernel.
    order = sorted(range(n), key=lambda k: views[k][1] * views[k][2] * views[k][4])[:cfg["check_smallest"]]
    mismatches = []
    for k in order:
        s, w, h, yaw, spp = views[k]
        rgb, var = srv.result(ids[k])
        r = wfamd.Renderer(base[s].view(w, h, yaw), "megakernel")
        want, wvar, _ = r.render(wfamd.params(spp=spp, seed=PROTOCOL["seed"] + k, max_depth=PROTOCOL["max_depth"]))
        r.close()
        if not (np.array_equal(rgb.view(np.uint32), want.view(np.uint32)) and
                np.array_equal(var.view(np.uint32), wvar.view(np.uint32))):
            mismatches.append(k)
    srv.close()
    rays = st["extension_rays"] + st["shadow_rays"]
    return {"config": 4, "requests": n, "budget_bytes": cfg["budget_bytes"],
            "capacity_requests_per_s": capacity, "closed_loop_makespan_s": span0,
            "closed_loop_mrays_per_s": (st0["extension_rays"] + st0["shadow_rays"]) / span0 / 1e6,
            "arrival_rate_per_s": rate, "load": rate / capacity, "makespan_s": span,
            "mrays_per_s": rays / span / 1e6,
            "latency_ms": {"p50": percentile(lat, 0.5), "p99": percentile(lat, 0.99), "max": max(lat),
                           "min": min(lat)},
            "completed": st["completed"], "rejected": st["rejected"], "peak_in_flight": peak,
            "checked_alone_vs_mix": order, "alone_equals_mix_bitwise": not mismatches, "mismatched": mismatches,
            "paths": sum(w * h * spp for _, w, h, _, spp in views),
            "views": [{"scene": s, "width": w, "height": h, "yaw": yaw, "spp": spp} for s, w, h, yaw, spp in views]}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--config", type=int, required=True, choices=[0, 1, 2, 3, 4])
    ap.add_argument("--reps", type=int, default=0, help="timed repetitions (0 = protocol)")
    ap.add_argument("--warmup", type=int, default=-1)
    ap.add_argument("--quick", action="store_true", help="small sizes (smoke run)")
    ap.add_argument("--out", default="")
    ap.add_argument("--stage-profile", choices=["wavefront", "persistent"], default="",
                    help="only one stage-timed render of the configuration's scene with this arm (configs 0-2): "
                         "per stage ms, items, launches and algorithmic device-memory bytes")
    ap.add_argument("--spp", type=int, default=0, help="samples per pixel of --stage-profile (0 = protocol)")
    args = ap.parse_args()
    cfg = dict(PROTOCOL["configs"][str(args.config)])
    if args.reps > 0:
        cfg["reps"] = args.reps
    if args.warmup >= 0:
        cfg["warmup"] = args.warmup
    wfamd.lib()
    if args.stage_profile:
        if args.config > 2:
            ap.error("--stage-profile takes configs 0-2")
        spp = args.spp or cfg["spp"]
        scene = wfamd.Scene(cfg["scene"], cfg["width"], cfg["height"])
        st = stage_profile(scene, args.stage_profile, spp, PROTOCOL["seed"], PROTOCOL["max_depth"])
        print(json.dumps({"config": args.config, "scene": cfg["scene"], "width": cfg["width"],
                          "height": cfg["height"], "spp": spp, "arm": args.stage_profile, "stages": st}), flush=True)
        return
    if args.config <= 2:
        doc = bench_scene(args.config, cfg, args.quick)
        summary = {"config": args.config, "verdict": doc["verdict"], "reasons": doc["reasons"],
                   "megakernel_ms": doc["conditions"][0]["median"], "wavefront_ms": doc["conditions"][1]["median"],
                   "persistent_ms": doc["mi355x"]["persistent"]["median"],
                   "persistent_speedup": doc["mi355x"]["persistent_speedup_vs_megakernel"],
                   "persistent_bitwise": doc["mi355x"]["persistent_equals_megakernel_bitwise"],
                   "mrays_per_s": {k: v["mrays_per_s"] for k, v in doc["mi355x"]["mrays_per_s"].items()}}
    elif args.config == 3:
        doc = bench_batch(cfg, args.quick)
        summary = {k: doc[k] for k in ("config", "jobs", "mrays_per_s", "megakernel_mrays_per_s",
                                       "speedup_vs_megakernel_one_by_one", "jobs_bitwise_equal_standalone")}
    else:
        doc = bench_serving(cfg, args.quick)
        summary = {k: doc[k] for k in ("config", "requests", "mrays_per_s", "latency_ms", "arrival_rate_per_s",
                                       "capacity_requests_per_s", "peak_in_flight", "alone_equals_mix_bitwise")}
    if args.out:
        with open(args.out, "w") as f:
            json.dump(doc, f)
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
