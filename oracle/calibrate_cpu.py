"""Times the oracle's CPU render against the reference's own render_cpu on the same cores.

TEST INFRASTRUCTURE ONLY (a timing calibration of the CPU baseline, run in the container where the
reference's host code is built into oracle/_ref by `make -C oracle ref`; it never runs on a GPU box).

  python oracle/calibrate_cpu.py [--width 320 --height 180 --spp 4 --runs 3 --threads N]

For each reference scene (cornell, divergence, trianglehv; src/scenes.cpp) both renderers trace the
same paths (the oracle's film equals render_cpu's bit for bit, up to the sin/cos of the C library:
tests/test_cpu_oracle.py), so their wall times compare the implementations, not the work. The
reference schedules static row bands per thread (src/cpu_renderer.cpp:78-88); the oracle hands out
rows dynamically. Wall clock around the render only, median of `runs`, at 1 thread and at N threads.
Prints one JSON line per scene and thread count.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pyoracle  # noqa: E402


def median(v):
    return float(np.median(np.asarray(v)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, default=320)
    ap.add_argument("--height", type=int, default=180)
    ap.add_argument("--spp", type=int, default=4)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    args = ap.parse_args()
    if not pyoracle.ref_available():
        print("oracle/_ref/libwfref.so is not built (make -C oracle ref)", file=sys.stderr)
        sys.exit(2)
    model = ""
    for line in open("/proc/cpuinfo"):
        if line.startswith("model name"):
            model = line.split(":", 1)[1].strip()
            break
    for name in ("cornell", "divergence", "trianglehv"):
        ref = pyoracle.RefScene(name, args.width, args.height)
        osc = pyoracle.OScene(ref.arrays(), libm_trig=True)
        for threads in sorted({1, args.threads}):
            t_ref, t_or, cnt = [], [], None
            for _ in range(args.runs):
                t0 = time.perf_counter()
                ref.render_cpu(args.spp, 1234, 8, threads)
                t_ref.append(time.perf_counter() - t0)
                t0 = time.perf_counter()
                _, cnt = osc.render_sums(0, args.spp, 1234, 8, threads=threads)
                t_or.append(time.perf_counter() - t0)
            rays = cnt[0] + cnt[1]
            print(json.dumps({
                "scene": name, "size": f"{args.width}x{args.height}x{args.spp}", "threads": threads,
                "cpu": model, "rays": int(rays),
                "reference_render_cpu_s": median(t_ref), "oracle_s": median(t_or),
                "reference_mrays_per_s": rays / median(t_ref) / 1e6, "oracle_mrays_per_s": rays / median(t_or) / 1e6,
                "oracle_over_reference_speed": median(t_ref) / median(t_or)}), flush=True)


if __name__ == "__main__":
    main()
