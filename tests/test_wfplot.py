"""wfplot: SVG charts of bench.py's line, wfbench documents and wfstats sections (the reference's
tools/svgplot.cpp and tools/svgstats.cpp), on the CPU with synthetic inputs of the real layouts.

Every chart must parse as SVG, carry a title and a description, and draw its data to scale: bar
heights in proportion to the values, points in the order of their values, log axes by decade."""
import json
import os
import sys
import xml.etree.ElementTree as ET

import pytest

from conftest import PKG

sys.path.insert(0, PKG)
import wfplot  # noqa: E402

NS = "{http://www.w3.org/2000/svg}"


def bench_line():
    st = {"extend": {"ms": 7.66, "launches": 9, "items": 138700153, "gbs": 478.1, "waves_per_simd": 8.0},
          "shade": {"ms": 4.87, "launches": 9, "items": 138700153, "gbs": 2750.4, "waves_per_simd": 6.0},
          "shadow": {"ms": 1.76, "launches": 9, "items": 23807819, "gbs": 783.4, "waves_per_simd": 8.0}}
    return {"metric": "Mrays/s", "value": 12138.6, "ms_per_step": 13.39,
            "config": {"workload": "sphere_grid", "pool_paths_used": 58982400},
            "megakernel": {"mrays_per_s": 2914.2, "p50_ms": 55.76},
            "roofline": {"bound": "hbm", "kernel": "extend", "peak": 8000.0, "stages": st},
            "pools": {"16777216": {"mrays_per_s": 11261.3, "ms_per_step": 14.43},
                      "4194304": {"mrays_per_s": 9098.7, "ms_per_step": 17.86}},
            "configs": {"2": {"mrays_per_s": 4961.0, "megakernel": {"mrays_per_s": 814.8}},
                        "3": {"mrays_per_s": 5995.3, "megakernel_mrays_per_s": 504.2},
                        "4": {"mrays_per_s": 4107.9, "arrival_rate_per_s": 60.0}}}


def stats_doc():
    spp = [1, 2, 4, 8, 16, 32, 64]
    rm = [0.1 * s ** -0.5 for s in spp]
    lorenz = [(i / 100) ** 6 for i in range(101)]
    return {
        "convergence": {"scene": "cornell", "width": 160, "height": 90, "reference_spp": 16384, "spp": spp,
                        "rmse_mean": rm, "cpu_rmse_mean": rm[:4],
                        "loglog_fit": {"a": -3.3219, "b": -0.5, "r2": 1.0}},
        "furnace": {"albedo": 0.5, "emission": 1.0, "limit": 2.0, "depth": [1, 2, 3],
                    "cpu": [1.5, 1.75, 1.875], "megakernel": [1.5, 1.75, 1.875],
                    "persistent": [1.5, 1.75, 1.875], "series_to_depth": [1.5, 1.75, 1.875]},
        "variance": {"cornell": {"lorenz": lorenz, "gini": 0.71}},
        "survival": {"cornell": {"paths": 100, "extension_rays_by_cap": [100, 170, 210, 218]}},
        "divergence": {"cornell": {"warp32": {"mean_distinct_per_tile": 1.1},
                                   "wave64": {"mean_distinct_per_tile": 1.05, "expected_distinct_iid": 2.0}}},
        "scaling": {"cornell": {"megapixels": [0.1, 0.5, 1.0], "spp": 4, "megakernel_ms": [0.3, 0.6, 1.0],
                                "persistent_ms": [0.5, 0.7, 0.9],
                                "persistent_fit": {"a": 0.45, "b": 0.45, "r2": 1.0}}},
        "bvh": {"cornell": {"nodes": 7, "sah_cost": 7.5, "leaf_depth_hist": [0, 1, 1, 2]}},
        "rng": {"acf": [0.001, -0.0005, 0.0002], "acf_band95": 0.0019, "stream_samples": 1 << 20, "ks_d": 0.0006,
                "pair_chi2": 232.6, "pair_dof": 255},
    }


def wfbench_doc(scene, mk, wv, pe):
    def cond(mode, med):
        return {"mode": mode, "median": med, "p10": med * 0.98, "p90": med * 1.03}
    return {"scene_name": scene, "width": 1280, "height": 720, "spp": 64, "speedup": mk / wv,
            "conditions": [cond("megakernel", mk), cond("wavefront", wv)],
            "mi355x": {"persistent": cond("persistent", pe), "persistent_speedup_vs_megakernel": mk / pe}}


def parse(path):
    root = ET.parse(path).getroot()
    assert root.tag == NS + "svg"
    assert root.find(NS + "title").text and root.find(NS + "desc").text
    return root


def bars(root):
    return [(r.find(NS + "title").text, float(r.get("height")), float(r.get("y"))) for r in root.iter(NS + "rect")
            if r.get("class") == "bar"]


def run(tmp_path, *args):
    out = tmp_path / "charts"
    assert wfplot.main([str(out), *map(str, args)]) == 0
    return out


def test_bench_charts(tmp_path):
    src = tmp_path / "bench.json"
    src.write_text("[bench] progress text\n" + json.dumps(bench_line()) + "\n")
    out = run(tmp_path, "--bench", src)
    assert sorted(os.listdir(out)) == ["bench-pools.svg", "bench-roofline.svg", "bench-stages.svg",
                                       "bench-throughput.svg"]
    b = bars(parse(out / "bench-throughput.svg"))
    assert len(b) == 9   # 6 persistent workloads, megakernel on 3
    h = {t.split(":")[0]: height for t, height, _ in b}
    assert h["sphere_grid\n(headline) persistent wavefront"] / h["sphere_grid\n(headline) megakernel"] == \
        pytest.approx(12138.6 / 2914.2, rel=1e-3)
    r = parse(out / "bench-roofline.svg")
    refs = [e for e in r.iter(NS + "line") if e.get("class") == "ref"]
    assert len(refs) == 1
    peak_y = float(refs[0].get("y1"))
    shade = [x for x in bars(r) if x[0].startswith("shade")][0]
    base = shade[2] + shade[1]
    assert (base - shade[2]) / (base - peak_y) == pytest.approx(2750.4 / 8000.0, rel=1e-3)
    pts = [e for e in parse(out / "bench-pools.svg").iter(NS + "circle")]
    xs = [float(p.get("cx")) for p in pts]
    assert len(pts) == 3 and xs == sorted(xs)
    # log axis: 2^22 -> 2^24 is as far as 2^24 -> 2^26 scaled by log(58982400 / 2^24) / log(4)
    import math
    assert (xs[2] - xs[1]) / (xs[1] - xs[0]) == pytest.approx(math.log(58982400 / 2 ** 24) / math.log(4), rel=1e-3)


def test_bench_line_without_optional_sections(tmp_path):
    line = bench_line()
    for k in ("pools", "configs", "roofline"):
        line.pop(k)
    src = tmp_path / "b.json"
    src.write_text(json.dumps(line))
    out = run(tmp_path, "--bench", src)
    assert os.listdir(out) == ["bench-throughput.svg"]
    assert len(bars(parse(out / "bench-throughput.svg"))) == 2


def test_wfbench_charts(tmp_path):
    paths = []
    for i, d in enumerate([wfbench_doc("cornell555", 5.2, 19.5, 1.24), wfbench_doc("sphere_grid", 57.8, 121.0, 13.9)]):
        p = tmp_path / f"d{i}.json"
        p.write_text(json.dumps(d))
        paths.append(p)
    out = run(tmp_path, "--wfbench", *paths)
    t = parse(out / "wfbench-timing.svg")
    assert len(bars(t)) == 6
    assert len([e for e in t.iter(NS + "line") if e.get("class") == "whisker"]) == 6
    s = bars(parse(out / "wfbench-speedup.svg"))
    h = {x[0].split(":")[0]: x[1] for x in s}
    assert h["sphere_grid\n1280x720x64 persistent wavefront"] / h["sphere_grid\n1280x720x64 plain wavefront"] == \
        pytest.approx((57.8 / 13.9) / (57.8 / 121.0), rel=1e-3)


def test_stats_charts(tmp_path):
    src = tmp_path / "s.json"
    src.write_text(json.dumps(stats_doc()))
    out = run(tmp_path, "--stats", src)
    names = sorted(os.listdir(out))
    assert names == ["stats-bvh.svg", "stats-convergence.svg", "stats-divergence.svg", "stats-furnace.svg",
                     "stats-rng.svg", "stats-scaling.svg", "stats-survival.svg", "stats-variance.svg"]
    for n in names:
        parse(out / n)
    conv = parse(out / "stats-convergence.svg")
    gpu = [e for e in conv.iter(NS + "circle") if e.find(NS + "title").text.startswith("GPU")]
    xs = [float(e.get("cx")) for e in gpu]
    ys = [float(e.get("cy")) for e in gpu]
    # log-log: equal steps per doubling, and RMSE ~ spp^-1/2 halves per quadrupling
    dx = [b - a for a, b in zip(xs, xs[1:])]
    dy = [b - a for a, b in zip(ys, ys[1:])]
    assert max(dx) - min(dx) < 0.2 and max(dy) - min(dy) < 0.2
    # the fit line (2^a spp^b, a = log2 0.1) lies on the points
    fit = [e for e in conv.iter(NS + "polyline") if e.find(NS + "title").text == "fit"][0]
    (fx0, fy0), (fx1, fy1) = [tuple(map(float, p.split(","))) for p in fit.get("points").split()]
    assert (fx0, fy0) == pytest.approx((xs[0], ys[0]), abs=0.2) and (fx1, fy1) == pytest.approx((xs[-1], ys[-1]), abs=0.2)
    dv = bars(parse(out / "stats-divergence.svg"))
    assert len(dv) == 3


def test_requires_an_input(tmp_path):
    with pytest.raises(SystemExit):
        wfplot.main([str(tmp_path / "x")])
