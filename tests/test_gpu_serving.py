"""Serving loop and engine lifetimes on the GPU.

- A request's memory is bounded by its pixels: a 1920x1080 request at 256 spp (8.5 GB of samples)
  completes under an 8 GiB budget in sample-range chunks, bit for bit equal to the megakernel.
- Job slots are recycled while the server is busy: more requests than the 4096-slot job ring pass
  through one server with a steady latency.
- Nine distinct scenes in flight (the engine's scene table has eight slots): the ninth waits for a
  free slot and every result is exact; a scene handle destroyed after its request and a new one
  created in its place render their own geometry.
- Create/destroy cycles of renderers, batches and servers leave no device memory behind
  (tests/catch/wavefront_failure.cu:136-166, tests/test_gpu_queues.cu:193).
- The large blocks of an ended engine are kept for the next engine of the same size: a second
  renderer takes them without new device memory and renders the same bits; releasing the cache
  gives the memory back.
"""
import ctypes as C
import gc
import time

import numpy as np
import pytest

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu


def megakernel_image(wf, scene, spp, seed, depth=8):
    r = wf.Renderer(scene, "megakernel")
    rgb, var, st = r.render(wf.params(spp=spp, seed=seed, max_depth=depth))
    r.close()
    return rgb, var, st


def drain(srv, ids, max_steps=20000):
    for _ in range(max_steps):
        if all(srv.poll(i) != 0 for i in ids):
            return
        srv.step(1)
    raise AssertionError(f"requests did not complete in {max_steps} steps")


def test_1080p_256spp_request_under_an_8_gib_budget(wf):
    print("[serving] sphere_grid 1920x1080x256 under 8 GiB", flush=True)
    s = wf.Scene("sphere_grid", 1920, 1080)
    budget = 8 << 30
    srv = wf.Server(budget_bytes=budget)
    rid = srv.submit(s, 256, 77)
    peak = 0
    t0 = time.perf_counter()
    for _ in range(100000):
        srv.step(1)
        st = srv.stats()
        peak = max(peak, st["bytes_in_use"])
        if srv.poll(rid) != 0:
            break
    served_s = time.perf_counter() - t0
    assert srv.poll(rid) == 1
    rgb, var = srv.result(rid)
    srv.close()
    assert peak <= budget
    print(f"[serving] served in {served_s:.3f} s, peak {peak / 2**30:.2f} GiB of the budget", flush=True)
    want, wvar, _ = megakernel_image(wf, s, 256, 77)
    assert_bitwise(rgb, want, "1080p x 256 rgb")
    assert_bitwise(var, wvar, "1080p x 256 var")


def test_more_requests_than_job_slots_with_steady_latency(wf):
    """5000 small requests through one server, four at a time, so the job ring wraps; the p99
    latency of the last thousand stays within twice that of the first thousand."""
    s = wf.Scene("cornell555", 16, 16)
    srv = wf.Server(budget_bytes=2 << 30, tuning=wf.params(pool_paths=1 << 20))
    n, window = 5000, 1000
    lat, pending = [], []
    ids = []
    t_sub = {}
    k = 0
    t0 = time.perf_counter()
    steps_since = 0   # steps since the last completion: a request that stops progressing fails the test
    while k < n or pending:
        while k < n and len(pending) < 4:
            rid = srv.submit(s, 1, k % 97)
            t_sub[rid] = time.perf_counter()
            pending.append(rid)
            ids.append(rid)
            k += 1
        srv.step(1)
        now = time.perf_counter()
        still = []
        for rid in pending:
            st = srv.poll(rid)
            assert st >= 0, f"request {rid} failed"
            if st == 1:
                lat.append(now - t_sub[rid])
                srv.release(rid)
            else:
                still.append(rid)
        steps_since = 0 if len(still) < len(pending) else steps_since + 1
        assert steps_since < 2000, f"no request completed in 2000 steps ({len(lat)} of {n} done)"
        if len(still) < len(pending) and len(lat) % 1000 < len(pending) - len(still):
            print(f"[soak] {len(lat)} done, {time.perf_counter() - t0:.1f} s", flush=True)
        pending = still
    stats = srv.stats()
    srv.close()
    assert stats["completed"] == n and stats["in_flight"] == 0
    first = np.percentile(np.array(lat[:window]) * 1e3, 99)
    last = np.percentile(np.array(lat[-window:]) * 1e3, 99)
    print(f"[soak] p99 first {first:.2f} ms, last {last:.2f} ms", flush=True)
    assert last <= 2.0 * first + 1.0


def test_request_left_alone_completes_with_counted_final_shadow_rays(wf, po, oracle_ref):
    """A large Cornell box request beside a tiny one, two pool iterations per launch: the tiny one
    finishes first, and the large one continues alone (single-job kernels) while final shadow rays
    that its multi-job iterations queued (counted when traced) are still pending. It completes, and
    both equal the oracle."""
    big, tiny = wf.Scene("cornell555", 64, 36), wf.Scene("cornell555", 4, 4)
    srv = wf.Server(budget_bytes=2 << 30, tuning=wf.params(pool_paths=4096, graph_iters=2))
    a = srv.submit(big, 16, 31)
    b = srv.submit(tiny, 1, 32)
    saw_alone = False
    for _ in range(20000):
        srv.step(1)
        if srv.poll(b) == 1 and srv.poll(a) == 0:
            saw_alone = True
        if srv.poll(a) != 0 and srv.poll(b) != 0:
            break
    assert srv.poll(a) == 1 and srv.poll(b) == 1
    assert saw_alone
    for rid, sc, spp, seed in ((a, big, 16, 31), (b, tiny, 1, 32)):
        rgb, var = srv.result(rid)
        orgb, ovar = po.finalize(*oracle_ref(sc, spp, seed)[0], spp)
        assert_bitwise(rgb, orgb, "request left alone rgb vs oracle")
        assert_bitwise(var, ovar, "request left alone var vs oracle")
    srv.close()


def test_nine_distinct_scenes_in_flight_and_a_recreated_scene(wf, po, oracle_ref):
    names = ["cornell555", "sphere_grid", "divergence", "cornell", "trianglehv", "furnace", "sphere_grid",
             "cornell555", "divergence"]
    scenes = [wf.Scene(n, 24 + 4 * k, 16 + 2 * k) for k, n in enumerate(names)]
    srv = wf.Server(budget_bytes=8 << 30, tuning=wf.params(pool_paths=1 << 20))
    ids = [srv.submit(sc, 3, 500 + k) for k, sc in enumerate(scenes)]
    srv.step(1)
    assert srv.stats()["in_flight"] <= 8    # the ninth scene waits for a slot
    drain(srv, ids)
    for k, (sc, rid) in enumerate(zip(scenes, ids)):
        assert srv.poll(rid) == 1
        rgb, var = srv.result(rid)
        sums, _ = oracle_ref(sc, 3, 500 + k)
        orgb, ovar = po.finalize(*sums, 3)
        assert_bitwise(rgb, orgb, f"scene {k} {names[k]} rgb vs oracle")
        assert_bitwise(var, ovar, f"scene {k} var vs oracle")
    # A scene destroyed after its request; a different scene created afterwards (it may reuse the
    # handle's address) renders its own geometry.
    a = wf.Scene("sphere_grid", 40, 24)
    ra = srv.submit(a, 2, 9)
    drain(srv, [ra])
    srv.release(ra)
    del a
    gc.collect()
    b = wf.Scene("cornell555", 40, 24)
    rb = srv.submit(b, 2, 9)
    drain(srv, [rb])
    rgb, var = srv.result(rb)
    want, wvar, _ = megakernel_image(wf, b, 2, 9)
    assert_bitwise(rgb, want, "recreated scene rgb")
    assert_bitwise(var, wvar, "recreated scene var")
    srv.close()


def free_bytes(wf):
    gc.collect()
    from hipmem import sync
    sync()
    return wf.query_device()["free_memory_bytes"]


def test_large_blocks_of_ended_engines_are_reused_and_released(wf):
    print("[cache] sphere_grid 1280x720x16 persistent, twice", flush=True)
    wf.release_cached_memory()
    assert wf.cached_memory_bytes() == 0
    s = wf.Scene("sphere_grid", 1280, 720)
    p = wf.params(spp=16)
    base = free_bytes(wf)
    r = wf.Renderer(s, "persistent")
    rgb1, var1, _ = r.render(p)
    r.close()
    kept = wf.cached_memory_bytes()
    # the work-sized pool: 14,745,600 slots, its ten 236 MB arrays and two 118 MB ones
    assert kept >= 2_000_000_000, kept
    assert base - free_bytes(wf) >= kept - (64 << 20)
    before = free_bytes(wf)
    r = wf.Renderer(s, "persistent")
    rgb2, var2, _ = r.render(p)
    during = free_bytes(wf)
    r.close()
    print(f"[cache] kept {kept}, device memory taken by the second renderer {before - during}", flush=True)
    assert before - during < (256 << 20)          # its pool came from the cache
    assert_bitwise(rgb2, rgb1, "second render rgb")
    assert_bitwise(var2, var1, "second render var")
    assert wf.cached_memory_bytes() == kept
    freed = wf.release_cached_memory()
    assert freed == kept and wf.cached_memory_bytes() == 0
    assert free_bytes(wf) - before >= kept - (64 << 20)


def test_create_destroy_cycles_leave_no_device_memory(wf):
    s = wf.Scene("sphere_grid", 64, 36)
    # One cycle of each first: device libraries and pools that stay for the process are set up.
    def cycle():
        for mode in ("megakernel", "wavefront", "persistent"):
            r = wf.Renderer(s, mode)
            r.render(wf.params(spp=2))
            r.close()
        wf.batch_render([(s, 2, 1, 8), (s, 1, 2, 8, 32, 16, 45.0)], wf.params())
        srv = wf.Server(budget_bytes=2 << 30, tuning=wf.params(pool_paths=1 << 20))
        rid = srv.submit(s, 2, 3)
        drain(srv, [rid])
        srv.close()
        ws = C.c_void_p()
        wf._check(wf.lib().wf_scan_workspace_create(1 << 20, C.byref(ws)))
        wf.lib().wf_scan_workspace_destroy(ws)
    cycle()
    before = free_bytes(wf)
    for k in range(20):
        cycle()
    after = free_bytes(wf)
    print(f"[lifetime] free device memory before {before} after {after} (delta {before - after})", flush=True)
    assert before - after < (8 << 20)
