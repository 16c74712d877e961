"""Benchmark: Mrays/s of the persistent batched wavefront on the material-divergent scene.

Workload: sphere_grid (1024 spheres on a jittered 32x32 grid, Lambert / metal (fuzz) / dielectric,
sky-gradient emitter used for next-event estimation) at 1280x720, 64 spp, depth 8, seed 1234, FP32,
synthetic procedural input. One step = one full render of the workload (reset + all 64 passes),
scene and film resident in HBM; the download of the image is outside the timed region.

Prints one JSON line with:
  value        (extension + shadow rays of all K steps) / wall time of the K steps, in Mrays/s
  p50/p99      per-step render latency (HIP events around each step)
  roofline     dominant kernel of the flagship arm: algorithmic bytes per launch / average launch
               duration (stage-timed render with HIP events on the engine's stream), against 8 TB/s
  cpu_baseline the oracle (a C restatement of the same estimator) on the host cores, bounded sample
  quality      persistent == megakernel bitwise, zero overflows / invalid hits / non-finite pixels
  pools        the same render with the path pool capped at 2^24 and 2^22 slots (3.1 and 0.8 GB of
               path state instead of the work-sized 10.9 GB): Mrays/s and p50 at each
  configs      the other benchmark workloads, each also on its own clock: heightfield 1280x720x64
               (Mrays/s, p50/p99, speedup over the megakernel), the 64-job batch in one launch
               (Mrays/s, speedup over the megakernel rendering the jobs one at a time) and the
               serving loop (capacity, and Mrays/s with p50/p99 latency at a fixed Poisson rate)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "graphics-book-tutorials_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "Mrays/s at 1280x720x64spp depth 8 per scene; p50/p99 render latency"
TUNING_KEYS = ("block_size", "pool_paths", "lds_kib", "refill", "leaf_batch", "dir_bins", "shade_block")
HBM_PEAK_GBS = 8000.0
# Host threads of the CPU baseline: the CPUs of this process's affinity mask, at most the host share
# of one GPU on a shared multi-GPU host (16).
CPU_SHARE = 16
# Capped pool sizes timed beside the work-sized pool, and renders per timing.
POOL_CAPS = (1 << 24, 1 << 22)
POOL_STEPS = 10
POOL_SLOT_BYTES = 184   # two 64 B path-state buffers, the 8 B hit and a 48 B shadow-queue entry per slot
HF_STEPS = 10


def percentile(v, p):
    v = sorted(v)
    if len(v) == 1:
        return v[0]
    x = p * (len(v) - 1)
    i = int(x)
    f = x - i
    return v[i] + (v[min(i + 1, len(v) - 1)] - v[i]) * f


def cpu_description():
    """The host CPU as /proc/cpuinfo describes it: model name, family/model/stepping, clock, logical
    CPUs and sockets (a virtual machine may report a generic model name; the numbers identify it)."""
    d = {"model": "", "family": None, "model_number": None, "stepping": None, "mhz": None, "logical_cpus": 0,
         "sockets": 0, "flags_avx512": False}
    phys = set()
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                d["logical_cpus"] += 1
            elif k == "model name" and not d["model"]:
                d["model"] = v
            elif k == "cpu family" and d["family"] is None:
                d["family"] = int(v)
            elif k == "model" and d["model_number"] is None:
                d["model_number"] = int(v)
            elif k == "stepping" and d["stepping"] is None:
                d["stepping"] = int(v) if v.isdigit() else v
            elif k == "cpu MHz" and d["mhz"] is None:
                d["mhz"] = float(v)
            elif k == "physical id":
                phys.add(v)
            elif k == "flags" and not d["flags_avx512"]:
                d["flags_avx512"] = "avx512f" in v.split()
    except OSError:
        pass
    d["sockets"] = len(phys)
    return d


def cpu_baseline(scene, scene_name, w, h, spp_sample, gpu_sums, runs=3):
    """The oracle (a plain-C restatement of the same estimator) rendering the first spp_sample
    samples of the workload on the host cores: median wall time of `runs` renders. Its image is
    also the CPU side of the quality gate: the GPU's render of the same samples must have at least
    99.9% of its pixels within 3 combined standard errors of it (reported with the fraction of
    bitwise-equal pixels)."""
    import numpy as np
    import pyoracle
    visible = len(os.sched_getaffinity(0))
    threads = max(1, min(CPU_SHARE, visible))
    o = pyoracle.OScene(scene.arrays())
    times, sums, cnt = [], None, None
    for _ in range(runs):
        t0 = time.perf_counter()
        sums, cnt = o.render_sums(0, spp_sample, 1234, 8, threads=threads)
        times.append(time.perf_counter() - t0)
    dt = percentile(times, 0.5)
    rgb_c, var_c = pyoracle.finalize(*sums, spp_sample)
    rgb_g, var_g = pyoracle.finalize(*gpu_sums, spp_sample)
    cmp = pyoracle.compare(rgb_g, var_g, spp_sample, rgb_c, var_c, spp_sample)
    same = (rgb_g.view(np.uint32).reshape(-1, 3) == rgb_c.view(np.uint32).reshape(-1, 3)).all(axis=1)
    cpu = cpu_description()
    base = {"value": (cnt[0] + cnt[1]) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{scene_name} {w}x{h}, samples 0..{spp_sample - 1} of 64, depth 8, seed 1234: {cnt[0]} "
                      f"extension + {cnt[1]} shadow rays, median of {runs} renders {dt:.3f} s "
                      f"({', '.join(f'{t:.3f}' for t in times)}); {threads} threads of the {visible} CPUs in "
                      f"this process's affinity mask (at most {CPU_SHARE}: the host share of one GPU); "
                      f"{cpu['model']}",
            "host_cpus_visible": visible, "cpu": cpu}
    gate = {"frac_within_3se": cmp["frac_within_3se"], "threshold": 0.999,
            "pass": cmp["frac_within_3se"] >= 0.999 and cmp["nonfinite"] == 0,
            "bitwise_pixel_fraction": float(same.mean()), "samples": spp_sample}
    return base, gate


def other_configs(wfamd):
    """The other benchmark workloads, each timed on its own (wfbench.py holds their protocol):
    heightfield 1280x720x64 on the persistent engine (autotuned) beside the megakernel; the 64-job
    batch in one launch beside the megakernel rendering the same views one at a time, every job
    bitwise equal to its standalone render; the serving loop's capacity (every request at once) and
    its latency and throughput at a fixed Poisson arrival rate."""
    import numpy as np
    import wfbench
    out = {}
    t_all = time.perf_counter()
    # heightfield: 2.1 M triangles, BVH built on the host (outside the timed region).
    cfg = wfbench.PROTOCOL["configs"]["2"]
    w, h, spp = cfg["width"], cfg["height"], cfg["spp"]
    seed, depth = wfbench.PROTOCOL["seed"], wfbench.PROTOCOL["max_depth"]
    scene = wfamd.Scene(cfg["scene"], w, h)
    tuned = wfamd.autotune([(scene, spp, seed, depth)], probe_spp=16)
    pt = wfamd.params(spp=spp, seed=seed, max_depth=depth, **{k: getattr(tuned, k) for k in TUNING_KEYS})
    r = wfamd.Renderer(scene, "persistent")
    for _ in range(2):
        r.reset()
        r.run_passes(0, spp, pt)
    lat, rays = [], 0
    t0 = time.perf_counter()
    for _ in range(HF_STEPS):
        r.reset()
        s = r.run_passes(0, spp, pt)
        lat.append(s.render_ms)
        rays += s.extension_rays + s.shadow_rays
    el = time.perf_counter() - t0
    sums_p = r.download_sums()
    r.close()
    mk = wfamd.Renderer(scene, "megakernel")
    mk.reset()
    mk.run_passes(0, spp, pt)
    mk_lat = []
    for _ in range(3):
        mk.reset()
        ms = mk.run_passes(0, spp, pt)
        mk_lat.append(ms.render_ms)
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(mk.download_sums(), sums_p))
    mk.close()
    mk_mrays = (ms.extension_rays + ms.shadow_rays) / (percentile(mk_lat, 0.5) * 1e3)
    out["2"] = {"workload": "heightfield 1280x720 64spp depth 8 (2,093,058 triangles)",
                "mrays_per_s": rays / el / 1e6, "ms_per_step": el / HF_STEPS * 1e3, "steps": HF_STEPS,
                "p50_ms": percentile(lat, 0.5), "p99_ms": percentile(lat, 0.99),
                "megakernel": {"mrays_per_s": mk_mrays, "p50_ms": percentile(mk_lat, 0.5)},
                "speedup_vs_megakernel": percentile(mk_lat, 0.5) / percentile(lat, 0.5),
                "persistent_equals_megakernel_bitwise": bool(same),
                "queue_overflows": int(s.queue_overflows), "invalid_hits": int(s.invalid_hits),
                "tuning": {k: getattr(tuned, k) for k in TUNING_KEYS}}
    print(f"[bench] heightfield: {out['2']['ms_per_step']:.2f} ms ({out['2']['mrays_per_s']:.0f} Mrays/s), "
          f"megakernel {out['2']['megakernel']['p50_ms']:.1f} ms, bitwise={same}", file=sys.stderr, flush=True)
    # 64 jobs in one launch.
    d = wfbench.bench_batch(dict(wfbench.PROTOCOL["configs"]["3"]), False)
    out["3"] = {"workload": "64 jobs (scene of configs 0-2, orbit angle, 128x128..640x360 log-uniform, "
                            "spp in {1,4,16,64}, depth 8, seed 2025) in one persistent launch",
                **{k: d[k] for k in ("jobs", "pixels", "paths", "rays", "mrays_per_s", "megakernel_mrays_per_s",
                                     "megakernel_one_by_one_ms", "speedup_vs_megakernel_one_by_one",
                                     "jobs_bitwise_equal_standalone", "rays_equal", "queue_overflows",
                                     "invalid_hits", "iterations", "tuning", "autotune_s")},
                "batch_ms_p50": d["batch_render_ms"]["median"], "batch_ms_runs": d["batch_render_ms"]["runs"]}
    print(f"[bench] batch: {out['3']['batch_ms_p50']:.2f} ms ({out['3']['mrays_per_s']:.0f} Mrays/s), "
          f"{out['3']['speedup_vs_megakernel_one_by_one']:.1f}x the megakernel one by one, "
          f"bitwise={out['3']['jobs_bitwise_equal_standalone']}", file=sys.stderr, flush=True)
    # Serving loop.
    d = wfbench.bench_serving(dict(wfbench.PROTOCOL["configs"]["4"]), False)
    out["4"] = {"workload": "serving: 48 requests (scene of configs 0-2, orbit angle, 256x256..1920x1080 and "
                            "4..256 spp log-uniform, depth 8, seed 4242), Poisson arrivals, 8 GiB budget",
                **{k: d[k] for k in ("requests", "capacity_requests_per_s", "closed_loop_mrays_per_s",
                                     "arrival_rate_per_s", "mrays_per_s", "latency_ms", "completed", "rejected",
                                     "peak_in_flight", "alone_equals_mix_bitwise", "checked_alone_vs_mix")}}
    print(f"[bench] serving: capacity {out['4']['capacity_requests_per_s']:.1f} req/s; at "
          f"{out['4']['arrival_rate_per_s']:.0f} req/s: {out['4']['mrays_per_s']:.0f} Mrays/s, p50 "
          f"{out['4']['latency_ms']['p50']:.1f} ms, p99 {out['4']['latency_ms']['p99']:.1f} ms",
          file=sys.stderr, flush=True)
    out["wall_s"] = time.perf_counter() - t_all
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scene", default="sphere_grid")
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-spp", type=int, default=16, help="samples per pixel of the CPU baseline sample")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-pools", action="store_true", help="skip the renders at capped pool sizes")
    ap.add_argument("--no-configs", action="store_true", help="skip the heightfield, batch and serving workloads")
    ap.add_argument("--pool-paths", type=int, default=0, help="persistent path slots (0: sized to the work)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain-copy calibration kernel over BYTES (traffic-counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (across GPUs the work would shard by jobs and pixel tiles, "
              f"one process per GPU); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import wfamd

    wfamd.lib()
    if args.copy_probe:
        ms = wfamd.copy_probe(args.copy_probe, reps=args.steps)
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "reps": args.steps, "ms_per_copy": ms,
                          "gbs": 2 * args.copy_probe / (ms / 1e3) / 1e9}), flush=True)
        return
    dev = wfamd.query_device()
    print(f"[bench] device {dev['name']} {dev['arch']} {dev['compute_units']} CUs", file=sys.stderr, flush=True)
    scene = wfamd.Scene(args.scene, args.width, args.height)
    spp, depth, seed = args.spp, 8, 1234

    tuning = wfamd.params(spp=spp, seed=seed, max_depth=depth, pool_paths=args.pool_paths)
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuned = wfamd.autotune([(scene, spp, seed, depth)], probe_spp=16)
        for k in TUNING_KEYS:
            if k != "pool_paths" or not args.pool_paths:
                setattr(tuning, k, getattr(tuned, k))
        print(f"[bench] autotune: {', '.join(f'{k} {getattr(tuned, k)}' for k in TUNING_KEYS)} "
              f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    r = wfamd.Renderer(scene, "persistent")
    for _ in range(args.warmup):
        r.reset()
        r.run_passes(0, spp, tuning)
    lat, ext, sh = [], 0, 0
    last = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.reset()
        st = r.run_passes(0, spp, tuning)
        lat.append(st.render_ms)
        ext += st.extension_rays
        sh += st.shadow_rays
        last = st
    elapsed = time.perf_counter() - t0
    print(f"[bench] persistent: {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    sums_p = r.download_sums()
    rgb_p, var_p = wfamd.finalize(*sums_p, spp)
    # The boundary hands the image back to the host: time that download (mean rgb + luminance
    # variance, 16 B per pixel over PCIe) beside the render for the PCIe-inclusive rate.
    t1 = time.perf_counter()
    for _ in range(5):
        r.download(spp)
    download_ms = (time.perf_counter() - t1) / 5 * 1e3

    # Megakernel on the same workload: throughput beside the flagship, and the bitwise gate.
    mk = wfamd.Renderer(scene, "megakernel")
    mk.reset()
    mk.run_passes(0, spp, tuning)
    mk_lat = []
    for _ in range(max(2, min(5, args.steps))):
        mk.reset()
        mst = mk.run_passes(0, spp, tuning)
        mk_lat.append(mst.render_ms)
    sums_m = mk.download_sums()
    mk.close()
    bitwise = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(sums_p, sums_m))
    rays_per_render = last.extension_rays + last.shadow_rays
    mk_mrays = (mst.extension_rays + mst.shadow_rays) / (percentile(mk_lat, 0.5) * 1e3)
    print(f"[bench] megakernel median {percentile(mk_lat, 0.5):.2f} ms ({mk_mrays:.0f} Mrays/s); bitwise={bitwise}",
          file=sys.stderr, flush=True)

    # Stage-timed render (eager launches, HIP events on the engine's stream around each stage):
    # per-kernel average launch durations, items and algorithmic device-memory bytes per launch
    # (computed by the engine from each iteration's counts; layouts in wf_persistent.hip).
    roof = None
    prof = wfamd.params(spp=spp, seed=seed, max_depth=depth, stage_timing=True,
                        **{k: getattr(tuning, k) for k in TUNING_KEYS})
    r.reset()
    pst = r.run_passes(0, spp, prof)
    names = {"extend": "extend", "shade_lambert": "shade", "shadow": "shadow"}
    stages = {names[k]: v for k, v in pst.stages().items() if k in names}
    print(f"[bench] stages: {json.dumps(stages)}", file=sys.stderr, flush=True)
    if stages:
        dom = max(stages, key=lambda k: stages[k]["ms"])
        d = stages[dom]
        launches = max(d["launches"], 1)
        avg_s = d["ms"] / 1e3 / launches
        achieved = d["bytes"] / launches / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "note": ("extend and shadow are BVH traversal: bound by the issue of their node visits and "
                         "leaf tests (VALU issue, lane utilisation), not by HBM; traffic null: the persistent "
                         "render is not run under rocprofv3 (a profiler crash in that path has no shown "
                         "cause), so no FETCH_SIZE/WRITE_SIZE pass covers these kernels"),
                "avg_launch_us": avg_s * 1e6, "bytes_per_launch": d["bytes"] / launches,
                "items_per_launch": d["items"] / launches,
                "stages": {k: {"ms": v["ms"], "launches": v["launches"], "items": v["items"],
                               "gbs": v["bytes"] / (v["ms"] / 1e3) / 1e9 if v["ms"] > 0 else None,
                               # occupancy: waves of the persistent grid resident on each SIMD
                               # (4 SIMDs per CU; the device allows at most 8 per SIMD)
                               "waves_per_simd": v["threads"] / 64 / (4 * dev["compute_units"])}
                           for k, v in stages.items()}}
    r.close()

    # The same render with the path pool capped below the work (fewer slots: more iterations).
    pools = None
    if not args.no_pools:
        pools = {}
        for cap in POOL_CAPS:
            pt = wfamd.params(spp=spp, seed=seed, max_depth=depth,
                              **{k: getattr(tuning, k) for k in TUNING_KEYS if k != "pool_paths"}, pool_paths=cap)
            rp = wfamd.Renderer(scene, "persistent")
            for _ in range(2):
                rp.reset()
                rp.run_passes(0, spp, pt)
            pl, prays = [], 0
            tp = time.perf_counter()
            for _ in range(POOL_STEPS):
                rp.reset()
                s = rp.run_passes(0, spp, pt)
                pl.append(s.render_ms)
                prays += s.extension_rays + s.shadow_rays
            pel = time.perf_counter() - tp
            same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(rp.download_sums(), sums_p))
            pools[str(cap)] = {"mrays_per_s": prays / pel / 1e6, "ms_per_step": pel / POOL_STEPS * 1e3,
                               "p50_ms": percentile(pl, 0.5), "p99_ms": percentile(pl, 0.99),
                               "iterations": int(s.iterations), "path_state_bytes": cap * POOL_SLOT_BYTES,
                               "equals_work_sized_pool_bitwise": bool(same)}
            rp.close()
            print(f"[bench] pool {cap}: {pools[str(cap)]['ms_per_step']:.2f} ms per render "
                  f"({pools[str(cap)]['mrays_per_s']:.0f} Mrays/s), {s.iterations} iterations, bitwise={same}",
                  file=sys.stderr, flush=True)

    configs = None
    if not args.no_configs:
        configs = other_configs(wfamd)

    cpu, cpu_gate = None, None
    if not args.no_cpu:
        # GPU side of the CPU quality gate: the same first cpu_spp samples on the persistent engine.
        r = wfamd.Renderer(scene, "persistent")
        r.reset()
        r.run_passes(0, args.cpu_spp, tuning)
        gpu_sample = r.download_sums()
        r.close()
        cpu, cpu_gate = cpu_baseline(scene, args.scene, args.width, args.height, args.cpu_spp, gpu_sample)
    value = (ext + sh) / elapsed / 1e6
    out = {
        "metric": METRIC, "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: procedural sphere_grid scene (1024 jittered spheres, seeded), no external data",
        "config": {"workload": f"{args.scene} {args.width}x{args.height} {spp}spp depth {depth} seed {seed} "
                               f"(material-divergent scene)", "arm": "persistent batched wavefront",
                   **{k: getattr(tuning, k) for k in TUNING_KEYS},
                   "pool_paths_used": int(tuning.pool_paths) if tuning.pool_paths > 0 else
                   min(1 << 26, (args.width * args.height * spp + 1023) // 1024 * 1024),
                   "parallelism": "single GPU"},
        "p50_ms": percentile(lat, 0.5), "p99_ms": percentile(lat, 0.99),
        "extension_mrays_per_s": ext / elapsed / 1e6, "rays_per_render": rays_per_render,
        "download_ms": download_ms,
        "mrays_per_s_with_download": rays_per_render / ((elapsed / args.steps * 1e3 + download_ms) * 1e3),
        "megakernel": {"mrays_per_s": mk_mrays, "p50_ms": percentile(mk_lat, 0.5)},
        "speedup_vs_megakernel": value / mk_mrays if mk_mrays else None,
        "quality": {"persistent_equals_megakernel_bitwise": bool(bitwise),
                    "queue_overflows": int(last.queue_overflows), "invalid_hits": int(last.invalid_hits),
                    "nonfinite_pixels": int((~np.isfinite(rgb_p.reshape(-1, 3))).any(axis=1).sum()),
                    "nonfinite_var": int((~np.isfinite(var_p)).sum()),
                    "cpu_vs_persistent_within_3se": cpu_gate},
        "roofline": roof, "cpu_baseline": cpu, "pools": pools, "configs": configs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
