"""The batched and serving benchmark configurations under the GPU parity bar, against the oracle.

Batch: the 64-job set of the benchmark protocol (seed 2025; scene uniform over cornell555 /
sphere_grid / heightfield, orbit yaw U[0, 360), width log-uniform [128, 640], height log-uniform
[128, 360], spp uniform over {1, 4, 16, 64}, depth 8) rendered in ONE autotuned persistent launch:
every job's image equals the oracle's render of the same view bit for bit, and the batch's ray
counts equal the oracle's.

Serving: requests of the serving configuration (scene uniform, yaw U[0, 360), resolution
log-uniform 256x256 .. 1920x1080, spp log-uniform [4, 256], depth 8) arriving as a Poisson process
under a device-memory budget that binds; the smallest requests equal the oracle bit for bit, every
request equals its standalone render bit for bit, and p50/p99 latency are printed.
Reference behaviour followed: src/gpu/gpu_renderer.cu:242-353 (a render is a function of the scene,
view, spp, seed and depth only).
"""
import time

import numpy as np
import pytest

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu


def oracle_image(po, oracle_ref, scene, spp, seed, depth=8):
    sums, cnt = oracle_ref(scene, spp, seed, depth)
    rgb, var = po.finalize(*sums, spp)
    return rgb, var, cnt


def test_batch_configuration_64_jobs_one_autotuned_launch_equal_oracle(wf, po, oracle_ref):
    import wfbench
    cfg = wfbench.PROTOCOL["configs"]["3"]
    rng = np.random.default_rng(cfg["job_seed"])
    views = wfbench.make_views(rng, cfg["jobs"], cfg, False)
    base = {s: wf.Scene(s, 64, 36) for s in cfg["scenes"]}
    seed0, depth = wfbench.PROTOCOL["seed"], wfbench.PROTOCOL["max_depth"]
    specs = [(base[s], spp, seed0 + k, depth, w, h, yaw) for k, (s, w, h, yaw, spp) in enumerate(views)]
    paths = sum(w * h * spp for _, w, h, _, spp in views)
    print(f"[batch] {len(specs)} jobs, {paths} paths: autotune", flush=True)
    t0 = time.perf_counter()
    tuned = wf.autotune(specs, probe_spp=4)
    print(f"[batch] tuned in {time.perf_counter() - t0:.1f} s: block {tuned.block_size} lds {tuned.lds_kib} KiB "
          f"refill {tuned.refill} leaf_batch {tuned.leaf_batch} shade_block {tuned.shade_block}", flush=True)
    outs, st = wf.batch_render(specs, tuned)
    print(f"[batch] render {st.render_ms:.2f} ms, {st.iterations} iterations, "
          f"{(st.extension_rays + st.shadow_rays) / (st.render_ms * 1e3):.0f} Mrays/s", flush=True)
    assert st.queue_overflows == 0 and st.invalid_hits == 0
    ext = sh = 0
    t0 = time.perf_counter()
    for k, ((s, w, h, yaw, spp), spec, (rgb, var)) in enumerate(zip(views, specs, outs)):
        orgb, ovar, ocnt = oracle_image(po, oracle_ref, base[s].view(w, h, yaw), spp, spec[2], depth)
        assert_bitwise(rgb, orgb, f"job {k} {s} {w}x{h} yaw {yaw:.1f} spp {spp} rgb vs oracle")
        assert_bitwise(var, ovar, f"job {k} var vs oracle")
        ext += ocnt[0]
        sh += ocnt[1]
    print(f"[batch] oracle: {ext + sh} rays in {time.perf_counter() - t0:.1f} s", flush=True)
    assert (st.extension_rays, st.shadow_rays) == (ext, sh)


def test_serving_configuration_poisson_arrivals_under_a_binding_budget(wf, po, oracle_ref):
    """A dozen requests of the serving configuration (seed 4242) arriving as a Poisson process
    at 100 requests/s under a 768 MiB budget (a path pool of 2^21 slots and 400 MB for films and
    sample buffers), which holds a few requests at a time: the others wait in the queue."""
    import wfbench
    cfg = dict(wfbench.PROTOCOL["configs"]["4"])
    rng = np.random.default_rng(cfg["request_seed"])
    n = 12
    views = wfbench.make_views(rng, n, cfg, False)
    base = {s: wf.Scene(s, 64, 36) for s in cfg["scenes"]}
    seed0 = wfbench.PROTOCOL["seed"]
    paths = [w * h * spp for _, w, h, _, spp in views]
    print(f"[serving] {n} requests, {sum(paths)} paths, largest {max(paths)}", flush=True)
    budget = 768 << 20
    # Arrivals at 100 requests/s: Poisson, seeded.
    arrivals = np.cumsum(rng.exponential(1.0 / 100.0, size=n))
    srv = wf.Server(budget_bytes=budget)
    ids, t_done = [None] * n, [None] * n
    nxt, peak, peak_bytes, waited = 0, 0, 0, 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            s, w, h, yaw, spp = views[nxt]
            ids[nxt] = srv.submit(base[s], spp, seed0 + nxt, 8, w, h, yaw)
            nxt += 1
        srv.step(1)
        now = time.perf_counter() - t0
        st = srv.stats()
        peak, peak_bytes = max(peak, st["in_flight"]), max(peak_bytes, st["bytes_in_use"])
        waited += 1 if st["submitted"] - st["admitted"] > 0 else 0
        for k in range(nxt):
            if t_done[k] is None:
                p = srv.poll(ids[k])
                assert p >= 0, f"request {k} failed"
                if p == 1:
                    t_done[k] = now
        if nxt == n and all(t is not None for t in t_done):
            break
        if nxt < n and st["in_flight"] == 0:
            time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.002)))
        assert now < 300, "serving did not finish"
    lat = [(t_done[k] - arrivals[k]) * 1e3 for k in range(n)]
    print(f"[serving] p50 {np.percentile(lat, 50):.1f} ms p99 {np.percentile(lat, 99):.1f} ms, "
          f"up to {peak} on the GPU, peak {peak_bytes / 2**30:.2f} GiB, {waited} steps with requests queued",
          flush=True)
    assert peak_bytes <= budget and peak < n and waited > 0
    # Every request equals its standalone render; the four smallest equal the oracle.
    smallest = set(sorted(range(n), key=lambda k: paths[k])[:4])
    for k in range(n):
        s, w, h, yaw, spp = views[k]
        rgb, var = srv.result(ids[k])
        v = base[s].view(w, h, yaw)
        r = wf.Renderer(v, "megakernel")
        want, wvar, _ = r.render(wf.params(spp=spp, seed=seed0 + k, max_depth=8))
        r.close()
        assert_bitwise(rgb, want, f"request {k} {s} {w}x{h} spp {spp} alone vs in the mix")
        assert_bitwise(var, wvar, f"request {k} var alone vs in the mix")
        if k in smallest:
            orgb, ovar, _ = oracle_image(po, oracle_ref, v, spp, seed0 + k)
            assert_bitwise(rgb, orgb, f"request {k} rgb vs oracle")
            assert_bitwise(var, ovar, f"request {k} var vs oracle")
    srv.close()


def test_autotuner_queue_capacity_within_a_state_budget(wf, oracle_ref):
    """The autotuner's queue-capacity search: under a path-state budget every candidate pool fits
    it, the chosen tuning renders the same bits as the megakernel (and the oracle), and without a
    budget the work-sized pool competes with 2^22 slots."""
    s = wf.Scene("sphere_grid", 320, 180)
    spp = 96                                   # 5,529,600 paths: more than 2^22
    slot = 184
    for budget in (None, 256 << 20, 2 << 30):
        print(f"[autotune] sphere_grid 320x180x{spp} state budget {budget}", flush=True)
        t = wf.autotune([(s, spp, 1234, 8)], probe_spp=4, state_budget=budget)
        pool = t.pool_paths if t.pool_paths > 0 else (320 * 180 * spp + 1023) // 1024 * 1024
        print(f"[autotune]   block {t.block_size} pool {t.pool_paths} ({pool * slot / 2**20:.0f} MiB) "
              f"lds {t.lds_kib} refill {t.refill} leaf_batch {t.leaf_batch} shade_block {t.shade_block}", flush=True)
        assert pool * slot <= (budget or 1 << 62)
        if budget == 256 << 20:
            assert t.pool_paths == 1 << 20             # the largest power of two within the budget
        else:
            assert t.pool_paths in (0, 1 << 22)        # work-sized, or the smaller candidate
        r = wf.Renderer(s, "persistent")
        r.reset()
        st = r.run_passes(0, spp, t)
        got = r.download_sums()
        r.close()
        from conftest import assert_equals_oracle
        assert_equals_oracle(got, (st.extension_rays, st.shadow_rays, st.invalid_hits), oracle_ref(s, spp),
                             f"autotuned (budget {budget})")
    with pytest.raises(wf.InvalidArgument):
        wf.autotune([(s, 4, 1234, 8)], probe_spp=2, state_budget=1000 * slot)
