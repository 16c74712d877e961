"""wfplot: SVG charts of the measurements, the counterpart of the reference's chart tools
(tools/svgplot.cpp: timing with p10-p90 whiskers, speedup and stage breakdown from the bench
documents; tools/svgstats.cpp: one figure per wfstats section).

  python graphics-book-tutorials_amd/wfplot.py OUT_DIR [--bench LINE.json] [--wfbench DOC.json ...]
                                                       [--stats WFSTATS.json]

--bench    the JSON line `bench.py` prints: throughput of every workload (persistent engine beside
           the megakernel), the headline render's stage times and their algorithmic GB/s against the
           HBM peak, and render time against the path pool's size.
--wfbench  bench documents from `wfbench.py --json` (the reference's schema plus the `mi355x`
           object): median render time with p10-p90 whiskers per arm, and the speedups.
--stats    the document `wfstats.py` writes: convergence, furnace, variance concentration, survival,
           material divergence, scaling, BVH leaf depths and RNG autocorrelation.
Every chart is a self-contained SVG with a <title> and a <desc> naming its source; a line of the
bench input that is missing (an older document) leaves its chart out rather than failing.
"""
import argparse
import json
import math
import os
import sys

COLOURS = ("#2f6db5", "#d9622b", "#2a9d6f", "#b58a00", "#7b4fb3", "#c23b6a")
INK, INK2, GRID, BG = "#111111", "#555555", "#dddddd", "#ffffff"


def esc(s):
    return str(s).replace("&", "&amp;").replace("<", "&lt;").replace(">", "&gt;").replace('"', "&quot;")


def fmt(v):
    if v == 0:
        return "0"
    a = abs(v)
    if a < 1e6 and float(v).is_integer():
        return f"{int(v):,}"
    if a >= 1000:
        return f"{v:,.0f}"
    if a >= 10:
        return f"{v:.0f}"
    if a >= 1:
        return f"{v:.1f}"
    if a >= 0.01:
        return f"{v:.2f}"
    return f"{v:.1e}"


class Chart:
    """A plot area with linear or log10 axes; data coordinates map to SVG pixels."""

    def __init__(self, title, desc, w=720, h=450, left=78, right=24, top=86, bottom=62):
        self.w, self.h = w, h
        self.x0, self.x1, self.y0, self.y1 = left, w - right, h - bottom, top   # pixel box (y up)
        self.parts = []
        self.title, self.desc = title, desc
        self.xlog = self.ylog = False
        self.xr = (0.0, 1.0)
        self.yr = (0.0, 1.0)

    def axes(self, xr, yr, xlog=False, ylog=False, xlabel="", ylabel="", xticks=None, yticks=None):
        self.xlog, self.ylog = xlog, ylog
        self.xr = tuple(math.log10(v) for v in xr) if xlog else tuple(map(float, xr))
        self.yr = tuple(math.log10(v) for v in yr) if ylog else tuple(map(float, yr))
        if self.xr[1] == self.xr[0]:
            self.xr = (self.xr[0] - 1, self.xr[1] + 1)
        if self.yr[1] == self.yr[0]:
            self.yr = (self.yr[0] - 1, self.yr[1] + 1)
        for t in (yticks if yticks is not None else self._ticks(yr, ylog)):
            y = self.py(t)
            self.parts.append(f'<line x1="{self.x0}" y1="{y:.1f}" x2="{self.x1}" y2="{y:.1f}" stroke="{GRID}"/>')
            self.text(self.x0 - 6, y + 4, fmt(t), 11, "end", INK2)
        if xticks is not False:
            for t in (xticks if xticks is not None else self._ticks(xr, xlog)):
                x = self.px(t)
                self.parts.append(f'<line x1="{x:.1f}" y1="{self.y0}" x2="{x:.1f}" y2="{self.y0 + 4}" stroke="{INK2}"/>')
                self.text(x, self.y0 + 17, fmt(t), 11, "middle", INK2)
        self.parts.append(f'<line x1="{self.x0}" y1="{self.y0}" x2="{self.x1}" y2="{self.y0}" stroke="{INK2}"/>')
        if xlabel:
            self.text((self.x0 + self.x1) / 2, self.h - 18, xlabel, 12, "middle", INK2)
        if ylabel:
            cx, cy = 16, (self.y0 + self.y1) / 2
            self.parts.append(f'<text x="{cx}" y="{cy:.1f}" font-size="12" fill="{INK2}" text-anchor="middle" '
                              f'transform="rotate(-90 {cx} {cy:.1f})">{esc(ylabel)}</text>')

    @staticmethod
    def _ticks(r, log):
        lo, hi = r
        if log:
            a, b = math.floor(math.log10(lo) + 1e-9), math.ceil(math.log10(hi) - 1e-9)
            return [10.0 ** k for k in range(a, b + 1) if lo <= 10.0 ** k * (1 + 1e-9) and 10.0 ** k <= hi * (1 + 1e-9)]
        span = hi - lo
        step = 10 ** math.floor(math.log10(span / 5)) if span > 0 else 1
        for m in (1, 2, 5, 10):
            if span / (step * m) <= 6:
                step *= m
                break
        t = math.ceil(lo / step) * step
        out = []
        while t <= hi + step * 1e-9:
            out.append(round(t, 12))
            t += step
        return out

    def px(self, v):
        v = math.log10(v) if self.xlog else v
        return self.x0 + (v - self.xr[0]) / (self.xr[1] - self.xr[0]) * (self.x1 - self.x0)

    def py(self, v):
        v = math.log10(v) if self.ylog else v
        return self.y0 - (v - self.yr[0]) / (self.yr[1] - self.yr[0]) * (self.y0 - self.y1)

    def text(self, x, y, s, size=12, anchor="start", colour=INK, bold=False):
        weight = ' font-weight="600"' if bold else ""
        self.parts.append(f'<text x="{x:.1f}" y="{y:.1f}" font-size="{size}" fill="{colour}" text-anchor="{anchor}"'
                          f'{weight}>{esc(s)}</text>')

    def polyline(self, xs, ys, colour, width=2.0, dash=None, label=None):
        pts = " ".join(f"{self.px(x):.1f},{self.py(y):.1f}" for x, y in zip(xs, ys))
        d = f' stroke-dasharray="{dash}"' if dash else ""
        self.parts.append(f'<polyline class="series" fill="none" stroke="{colour}" stroke-width="{width}"{d} '
                          f'points="{pts}"><title>{esc(label or "")}</title></polyline>')

    def points(self, xs, ys, colour, r=3.0, label=None):
        for x, y in zip(xs, ys):
            self.parts.append(f'<circle class="point" cx="{self.px(x):.1f}" cy="{self.py(y):.1f}" r="{r}" '
                              f'fill="{colour}"><title>{esc(label or "")}: {fmt(x)}, {fmt(y)}</title></circle>')

    def whisker(self, x, lo, hi, colour=INK2):
        xp = self.px(x)
        a, b = self.py(lo), self.py(hi)
        self.parts.append(f'<line class="whisker" x1="{xp:.1f}" y1="{a:.1f}" x2="{xp:.1f}" y2="{b:.1f}" '
                          f'stroke="{colour}" stroke-width="1.5"/>')
        for y in (a, b):
            self.parts.append(f'<line x1="{xp - 5:.1f}" y1="{y:.1f}" x2="{xp + 5:.1f}" y2="{y:.1f}" stroke="{colour}"/>')

    def vbar(self, xc, width_px, value, colour, tip):
        top, base = self.py(value), self.py(self.yr[0] if not self.ylog else 10 ** self.yr[0])
        self.parts.append(f'<rect class="bar" x="{xc - width_px / 2:.1f}" y="{min(top, base):.1f}" '
                          f'width="{width_px:.1f}" height="{abs(base - top):.1f}" fill="{colour}">'
                          f'<title>{esc(tip)}</title></rect>')

    def hline(self, y, colour, label, dash="5 4"):
        yp = self.py(y)
        self.parts.append(f'<line class="ref" x1="{self.x0}" y1="{yp:.1f}" x2="{self.x1}" y2="{yp:.1f}" '
                          f'stroke="{colour}" stroke-width="1.5" stroke-dasharray="{dash}"/>')
        self.text(self.x1 - 4, yp - 5, label, 11, "end", colour)

    def legend(self, items):
        """Rows under the title, wrapped at the chart's width (three rows fit above the plot)."""
        x, y = self.x0, 34
        for name, colour in items:
            need = 26 + 7 * len(name)
            if x > self.x0 and x + need > self.w - 8:
                x, y = self.x0, y + 16
            self.parts.append(f'<rect x="{x}" y="{y}" width="12" height="12" fill="{colour}"/>')
            self.text(x + 16, y + 10, name, 12, "start", INK)
            x += need

    def svg(self):
        head = (f'<svg xmlns="http://www.w3.org/2000/svg" viewBox="0 0 {self.w} {self.h}" width="{self.w}" '
                f'height="{self.h}" role="img" aria-labelledby="t d" font-family="Helvetica, Arial, sans-serif">'
                f'<title id="t">{esc(self.title)}</title><desc id="d">{esc(self.desc)}</desc>'
                f'<rect width="{self.w}" height="{self.h}" fill="{BG}"/>')
        t = f'<text x="{self.x0}" y="24" font-size="15" font-weight="600" fill="{INK}">{esc(self.title)}</text>'
        return head + t + "".join(self.parts) + "</svg>\n"


def grouped_bars(title, desc, groups, series, values, ylabel, ref=None, ymax=None):
    """values[g][s] (None = absent); one bar per present value, groups along x; ref = (y, label)."""
    c = Chart(title, desc, w=max(560, 120 * len(groups) + 120))
    top = max([v for row in values for v in row if v is not None] + ([ref[0]] if ref else []))
    c.axes((0, len(groups)), (0, ymax if ymax else top * 1.12), ylabel=ylabel, xticks=False)
    gw = (c.x1 - c.x0) / len(groups)
    bw = min(44.0, 0.8 * gw / max(1, len(series)))
    for g, name in enumerate(groups):
        xc = c.x0 + gw * (g + 0.5)
        present = [s for s in range(len(series)) if values[g][s] is not None]
        for k, s in enumerate(present):
            x = xc + (k - (len(present) - 1) / 2) * bw
            c.vbar(x, bw * 0.9, values[g][s], COLOURS[s % len(COLOURS)], f"{name} {series[s]}: {fmt(values[g][s])}")
            c.text(x, c.py(values[g][s]) - 4, fmt(values[g][s]), 10, "middle", INK2)
        for i, line in enumerate(str(name).split("\n")):
            c.text(xc, c.y0 + 17 + 13 * i, line, 11, "middle", INK2)
    if ref:
        c.hline(ref[0], INK2, ref[1])
    c.legend([(s, COLOURS[i % len(COLOURS)]) for i, s in enumerate(series)])
    return c


# ---------------------------------------------------------------------------------------------
# bench.py's line
# ---------------------------------------------------------------------------------------------
def bench_charts(line, src):
    out = {}
    groups, rows = [], []
    mk = line.get("megakernel", {}).get("mrays_per_s")
    groups.append("sphere_grid\n(headline)")
    rows.append([line["value"], mk])
    for cap, p in sorted((line.get("pools") or {}).items(), key=lambda kv: -int(kv[0])):
        groups.append(f"sphere_grid\n2^{int(math.log2(int(cap)))} slots")
        rows.append([p["mrays_per_s"], None])
    cfg = line.get("configs") or {}
    if "2" in cfg:
        groups.append("heightfield")
        rows.append([cfg["2"]["mrays_per_s"], cfg["2"].get("megakernel", {}).get("mrays_per_s")])
    if "3" in cfg:
        groups.append("64-job batch")
        rows.append([cfg["3"]["mrays_per_s"], cfg["3"].get("megakernel_mrays_per_s")])
    if "4" in cfg:
        groups.append(f"serving at\n{cfg['4'].get('arrival_rate_per_s', 0):g} req/s")
        rows.append([cfg["4"]["mrays_per_s"], None])
    out["bench-throughput.svg"] = grouped_bars(
        "Throughput per workload (MI355X, one GPU)", f"Mrays/s from {src}; megakernel where measured",
        groups, ["persistent wavefront", "megakernel"], rows, "Mrays/s").svg()
    roof = line.get("roofline") or {}
    st = roof.get("stages")
    if st:
        names = list(st)
        c = grouped_bars("Headline render: time per stage", f"stage-timed render from {src}",
                         names, ["ms per render"], [[st[k]["ms"]] for k in names], "ms")
        for i, k in enumerate(names):
            gbs, wps = st[k].get("gbs"), st[k].get("waves_per_simd")
            note = f"{fmt(gbs)} GB/s" if gbs else ""
            note += f", {wps:g} waves/SIMD" if wps else ""
            gw = (c.x1 - c.x0) / len(names)
            c.text(c.x0 + gw * (i + 0.5), c.y0 + 30, note, 10, "middle", INK2)
        out["bench-stages.svg"] = c.svg()
        peak = roof.get("peak", 8000.0)
        c = grouped_bars("Algorithmic bandwidth per stage against the HBM peak", f"from {src}",
                         names, ["GB/s"], [[st[k]["gbs"] or 0.0] for k in names], "GB/s",
                         ref=(peak, f"HBM peak {fmt(peak)} GB/s"), ymax=peak * 1.08)
        out["bench-roofline.svg"] = c.svg()
    pools = line.get("pools")
    if pools:
        xs = [int(k) for k in pools]
        used = (line.get("config") or {}).get("pool_paths_used")
        pts = sorted([(x, pools[str(x)]["ms_per_step"]) for x in xs] + ([(used, line["ms_per_step"])] if used else []))
        c = Chart("Render time against the path pool's size", f"sphere_grid 1280x720x64 from {src}")
        c.axes((pts[0][0] / 2, pts[-1][0] * 2), (0, max(p[1] for p in pts) * 1.15), xlog=True,
               xlabel="path slots (log scale)", ylabel="ms per render")
        c.polyline([p[0] for p in pts], [p[1] for p in pts], COLOURS[0], label="ms per render")
        c.points([p[0] for p in pts], [p[1] for p in pts], COLOURS[0], label="ms per render")
        for x, y in pts:
            c.text(c.px(x), c.py(y) - 8, f"{y:.2f} ms", 10, "middle", INK2)
        out["bench-pools.svg"] = c.svg()
    return out


# ---------------------------------------------------------------------------------------------
# wfbench documents
# ---------------------------------------------------------------------------------------------
def wfbench_charts(docs, srcs):
    arms = ["megakernel", "wavefront", "persistent"]
    groups, cond = [], []
    for d in docs:
        by = {c["mode"]: c for c in d["conditions"]}
        if "mi355x" in d:
            by["persistent"] = d["mi355x"]["persistent"]
        groups.append(f"{d['scene_name']}\n{d['width']}x{d['height']}x{d['spp']}")
        cond.append(by)
    top = max(c[a]["p90"] for c in cond for a in arms if a in c)
    c = Chart("Render time per arm: median, p10-p90", "from " + ", ".join(srcs), w=max(560, 170 * len(groups) + 120))
    c.axes((0, len(groups)), (0, top * 1.15), ylabel="ms per render", xticks=False)
    gw = (c.x1 - c.x0) / len(groups)
    for g, by in enumerate(cond):
        xc = c.x0 + gw * (g + 0.5)
        for k, a in enumerate(arms):
            if a not in by:
                continue
            x = xc + (k - 1) * min(40.0, gw / 4)
            c.vbar(x, min(32.0, gw / 5), by[a]["median"], COLOURS[k], f"{a}: median {fmt(by[a]['median'])} ms")
            c.whisker(x, by[a]["p10"], by[a]["p90"])
        for i, line in enumerate(groups[g].split("\n")):
            c.text(xc, c.y0 + 17 + 13 * i, line, 11, "middle", INK2)
    c.legend([(a, COLOURS[k]) for k, a in enumerate(arms)])
    out = {"wfbench-timing.svg": c.svg()}
    rows = []
    for d in docs:
        rows.append([d.get("speedup"), d.get("mi355x", {}).get("persistent_speedup_vs_megakernel")])
    out["wfbench-speedup.svg"] = grouped_bars(
        "Speedup over the megakernel", "from " + ", ".join(srcs), groups,
        ["plain wavefront", "persistent wavefront"], rows, "x megakernel", ref=(1.0, "megakernel")).svg()
    return out


# ---------------------------------------------------------------------------------------------
# wfstats sections
# ---------------------------------------------------------------------------------------------
def stats_charts(s, src):
    out = {}
    cv = s.get("convergence")
    if cv:
        c = Chart("Convergence: RMSE against a high-spp reference",
                  f"{cv['scene']} {cv['width']}x{cv['height']}, reference {cv['reference_spp']} spp, from {src}")
        spp, rm = cv["spp"], cv["rmse_mean"]
        ys = [v for v in rm if v > 0] + [v for v in cv.get("cpu_rmse_mean", []) if v > 0]
        c.axes((min(spp) / 1.5, max(spp) * 1.5), (min(ys) / 2, max(ys) * 2), xlog=True, ylog=True,
               xlabel="samples per pixel", ylabel="RMSE")
        c.polyline(spp, rm, COLOURS[0], label="GPU")
        c.points(spp, rm, COLOURS[0], label="GPU")
        if cv.get("cpu_rmse_mean"):
            n = len(cv["cpu_rmse_mean"])
            c.points(spp[:n], cv["cpu_rmse_mean"], COLOURS[1], r=4, label="CPU arm")
        f = cv.get("loglog_fit")   # log2(RMSE) = a + b log2(spp)
        if f:
            c.polyline([spp[0], spp[-1]], [2.0 ** f["a"] * spp[0] ** f["b"], 2.0 ** f["a"] * spp[-1] ** f["b"]],
                       COLOURS[2], 1.5, "6 4", "fit")
            c.text(c.x1 - 6, c.y1 + 14, f"fit: RMSE ~ spp^{f['b']:.3f} (R^2 {f['r2']:.4f}); ideal -0.5", 12, "end")
        c.legend([("GPU", COLOURS[0]), ("CPU arm", COLOURS[1]), ("log-log fit", COLOURS[2])])
        out["stats-convergence.svg"] = c.svg()
    fu = s.get("furnace")
    if fu:
        d = fu["depth"]
        arms = [a for a in ("cpu", "megakernel", "wavefront", "persistent") if a in fu]
        lo = min(min(fu[a]) for a in arms)
        c = Chart("White furnace: mean radiance against the depth cap",
                  f"albedo {fu['albedo']}, emission {fu['emission']}; limit {fu['limit']}; from {src}")
        c.axes((0, max(d) + 1), (lo * 0.95, fu["limit"] * 1.03), xlabel="depth cap", ylabel="mean radiance")
        c.hline(fu["limit"], INK2, f"limit {fu['limit']:g}")
        c.polyline(d, fu["series_to_depth"], COLOURS[5], 1.5, "4 3", "geometric series")
        for k, a in enumerate(arms):
            c.points([x + 0.12 * (k - 1.5) for x in d], fu[a], COLOURS[k], r=3, label=a)
        c.legend([(a, COLOURS[k]) for k, a in enumerate(arms)] + [("series", COLOURS[5])])
        out["stats-furnace.svg"] = c.svg()
    va = s.get("variance")
    if va:
        c = Chart("Variance concentration: Lorenz curves of per-pixel variance", f"from {src}")
        c.axes((0, 1), (0, 1), xlabel="fraction of pixels (lowest variance first)", ylabel="share of variance")
        c.polyline([0, 1], [0, 1], INK2, 1.0, "4 3", "equal share")
        items = []
        for k, (name, v) in enumerate(va.items()):
            L = v["lorenz"]
            c.polyline([i / (len(L) - 1) for i in range(len(L))], L, COLOURS[k], label=name)
            items.append((f"{name} (Gini {v['gini']:.2f})", COLOURS[k]))
        c.legend(items)
        out["stats-variance.svg"] = c.svg()
    sv = s.get("survival")
    if sv:
        c = Chart("Path survival: extension rays per path by depth cap", f"from {src}")
        caps = max(len(v["extension_rays_by_cap"]) for v in sv.values())
        top = max(v["extension_rays_by_cap"][-1] / v["paths"] for v in sv.values())
        c.axes((0, caps - 1), (0, top * 1.1), xlabel="depth cap", ylabel="extension rays per path")
        for k, (name, v) in enumerate(sv.items()):
            ys = [e / v["paths"] for e in v["extension_rays_by_cap"]]
            c.polyline(list(range(len(ys))), ys, COLOURS[k], label=name)
            c.points(list(range(len(ys))), ys, COLOURS[k], r=2.5, label=name)
        c.legend([(n, COLOURS[k]) for k, n in enumerate(sv)])
        out["stats-survival.svg"] = c.svg()
    dv = s.get("divergence")
    if dv:
        names = list(dv)
        rows = [[dv[n]["warp32"]["mean_distinct_per_tile"], dv[n]["wave64"]["mean_distinct_per_tile"],
                 dv[n]["wave64"]["expected_distinct_iid"]] for n in names]
        out["stats-divergence.svg"] = grouped_bars(
            "Material divergence of primary hits: distinct classes per tile", f"from {src}", names,
            ["32-lane 16x2 tile", "64-lane 8x8 tile", "64 independent pixels"], rows, "distinct classes").svg()
    sc = s.get("scaling")
    if sc:
        name = next(iter(sc))
        v = sc[name]
        arms = [a for a in ("megakernel", "wavefront", "persistent") if f"{a}_ms" in v]
        mp = v["megapixels"]
        top = max(max(v[f"{a}_ms"]) for a in arms)
        c = Chart(f"Render time against image size ({name}, {v['spp']} spp)", f"from {src}")
        c.axes((0, max(mp) * 1.05), (0, top * 1.15), xlabel="megapixels", ylabel="ms per render")
        items = []
        for k, a in enumerate(arms):
            c.points(mp, v[f"{a}_ms"], COLOURS[k], label=a)
            f = v.get(f"{a}_fit")
            if f:
                c.polyline([0, max(mp)], [f["a"], f["a"] + f["b"] * max(mp)], COLOURS[k], 1.5, "5 3", a)
                items.append((f"{a}: {f['a']:.2f} + {f['b']:.2f} ms/MP", COLOURS[k]))
            else:
                items.append((a, COLOURS[k]))
        c.legend(items)
        out["stats-scaling.svg"] = c.svg()
    bv = s.get("bvh")
    if bv:
        names = list(bv)
        depth = max(len(bv[n]["leaf_depth_hist"]) for n in names)
        c = Chart("BVH leaf depth", f"share of leaves per depth; from {src}")
        c.axes((-0.5, depth - 0.5), (0, 1.05), xlabel="leaf depth", ylabel="share of leaves")
        for k, n in enumerate(names):
            h = bv[n]["leaf_depth_hist"]
            tot = max(1, sum(h))
            c.polyline(list(range(len(h))), [x / tot for x in h], COLOURS[k], label=n)
            c.points(list(range(len(h))), [x / tot for x in h], COLOURS[k], r=2.5, label=n)
        c.legend([(f"{n} ({bv[n]['nodes']} nodes, SAH {bv[n]['sah_cost']:.1f})", COLOURS[k])
                  for k, n in enumerate(names)])
        out["stats-bvh.svg"] = c.svg()
    rn = s.get("rng")
    if rn and rn.get("acf"):
        acf = rn["acf"]
        band = rn.get("acf_band95", 0.0)
        lim = max(max(abs(a) for a in acf), band) * 1.3 or 1.0
        c = Chart("RNG stream autocorrelation", f"{rn.get('stream_samples', 0)} samples; KS D {rn.get('ks_d', 0):.2g}; "
                  f"pair chi-square {rn.get('pair_chi2', 0):.1f} ({rn.get('pair_dof', 0)} dof); from {src}")
        c.axes((0, len(acf) + 1), (-lim, lim), xlabel="lag", ylabel="autocorrelation")
        c.hline(band, COLOURS[1], "95% band")
        c.hline(-band, COLOURS[1], "")
        c.points(list(range(1, len(acf) + 1)), acf, COLOURS[0], r=2.5, label="acf")
        out["stats-rng.svg"] = c.svg()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("out_dir")
    ap.add_argument("--bench")
    ap.add_argument("--wfbench", nargs="*", default=[])
    ap.add_argument("--stats")
    a = ap.parse_args(argv)
    if not a.bench and not a.wfbench and not a.stats:
        ap.error("nothing to plot: give --bench, --wfbench and/or --stats")
    charts = {}
    if a.bench:
        lines = [ln for ln in open(a.bench).read().splitlines() if ln.strip().startswith("{")]
        if not lines:
            ap.error(f"{a.bench}: no JSON line")
        charts.update(bench_charts(json.loads(lines[-1]), os.path.basename(a.bench)))
    if a.wfbench:
        docs = [json.load(open(p)) for p in a.wfbench]
        charts.update(wfbench_charts(docs, [os.path.basename(p) for p in a.wfbench]))
    if a.stats:
        charts.update(stats_charts(json.load(open(a.stats)), os.path.basename(a.stats)))
    os.makedirs(a.out_dir, exist_ok=True)
    for name, svg in charts.items():
        with open(os.path.join(a.out_dir, name), "w") as f:
            f.write(svg)
    print(f"wrote {len(charts)} charts to {a.out_dir}: {', '.join(sorted(charts))}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
