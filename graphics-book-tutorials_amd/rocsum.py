"""rocsum: summarises rocprofv3 CSV output per kernel -- the MI355X counterpart of the reference's
Nsight Compute summariser (tools/ncusum.cpp: per-kernel metric means, a summary object).

  python graphics-book-tutorials_amd/rocsum.py [--trace DIR] [--counters DIR ...]
         [--calibrate COPY_FETCH_DIR COPY_WRITE_DIR BYTES] [--json OUT] [--md OUT]

--trace     a `rocprofv3 --kernel-trace --stats --output-format csv` output directory (the
            *_kernel_stats.csv there): calls, total, mean, min and max duration per kernel.
--counters  `rocprofv3 --pmc ... --output-format csv` output directories (*_counter_collection.csv):
            per kernel, the number of dispatches and the mean value per dispatch of every counter.
--calibrate the FETCH_SIZE and WRITE_SIZE counter directories of a plain copy of BYTES bytes (a
            kernel that reads and writes BYTES each): the factors that turn the KiB those counters
            report into device-memory bytes (FETCH_SIZE counts one side of the read traffic on this
            device; the factors are measured rather than assumed). With them, every kernel that has
            both counters gets `hbm_bytes_per_dispatch` = FETCH_SIZE * f + WRITE_SIZE * w.
Kernel names are shortened to the function name (template arguments kept). A counter value that
is not a number is reported as null and left out of the means.
"""
import argparse
import csv
import glob
import json
import math
import os
import re
import sys
from collections import defaultdict


def short_name(full):
    """`ns::(anonymous namespace)::foo_kernel<2>(args...)` -> `foo_kernel<2>`."""
    s = full.strip().strip('"')
    depth, cut = 0, len(s)
    for i, ch in enumerate(s):   # the argument list starts at the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and not s.startswith("(anonymous", i):
            cut = i
            break
    head = s[:cut]
    head = re.sub(r"^void ", "", head)
    return head.split("::")[-1] if "::" in head else head


def _csvs(path, suffix):
    if os.path.isfile(path):
        return [path]
    found = sorted(glob.glob(os.path.join(path, "**", "*" + suffix), recursive=True))
    if not found:
        raise FileNotFoundError(f"no *{suffix} under {path}")
    return found


def _num(v):
    try:
        x = float(v)
    except (TypeError, ValueError):
        return None
    return x if math.isfinite(x) else None


def kernel_stats(trace_dir):
    out = {}
    for f in _csvs(trace_dir, "_kernel_stats.csv"):
        for r in csv.DictReader(open(f, newline="")):
            name = short_name(r["Name"])
            calls = int(float(r["Calls"]))
            total = _num(r["TotalDurationNs"]) or 0.0
            e = out.setdefault(name, {"calls": 0, "total_ns": 0.0, "min_ns": None, "max_ns": None})
            e["calls"] += calls
            e["total_ns"] += total
            for k, agg in (("min_ns", min), ("max_ns", max)):
                v = _num(r.get(k[:3].capitalize() + "Ns"))
                if v is not None:
                    e[k] = v if e[k] is None else agg(e[k], v)
    for e in out.values():
        e["mean_ns"] = e["total_ns"] / e["calls"] if e["calls"] else None
    return out


def counters(dirs):
    """{kernel: {counter: {"dispatches": n, "total": t, "per_dispatch": t / n}}}."""
    acc = defaultdict(lambda: defaultdict(lambda: [0, 0.0]))
    for d in dirs:
        for f in _csvs(d, "_counter_collection.csv"):
            for r in csv.DictReader(open(f, newline="")):
                v = _num(r["Counter_Value"])
                if v is None:
                    continue
                a = acc[short_name(r["Kernel_Name"])][r["Counter_Name"]]
                a[0] += 1
                a[1] += v
    return {k: {c: {"dispatches": n, "total": t, "per_dispatch": t / n} for c, (n, t) in cs.items()}
            for k, cs in acc.items()}


def calibrate(fetch_dir, write_dir, nbytes):
    """Bytes per counted KiB of FETCH_SIZE and WRITE_SIZE from a plain copy of nbytes: the copy
    kernel is the dispatch with the largest FETCH_SIZE."""
    f = counters([fetch_dir])
    w = counters([write_dir])
    kf = max(f, key=lambda k: f[k].get("FETCH_SIZE", {}).get("per_dispatch", 0.0))
    kw = max(w, key=lambda k: w[k].get("WRITE_SIZE", {}).get("per_dispatch", 0.0))
    return {"copy_kernel": kf, "copy_bytes": nbytes,
            "fetch_factor": nbytes / (f[kf]["FETCH_SIZE"]["per_dispatch"] * 1024.0),
            "write_factor": nbytes / (w[kw]["WRITE_SIZE"]["per_dispatch"] * 1024.0)}


def summarise(trace=None, counter_dirs=(), calib=None):
    doc = {"kernels": {}, "summary": {}}
    ks = kernel_stats(trace) if trace else {}
    cs = counters(counter_dirs) if counter_dirs else {}
    for name in sorted(set(ks) | set(cs)):
        e = dict(ks.get(name, {}))
        if name in cs:
            e["counters"] = cs[name]
            fs, ws = cs[name].get("FETCH_SIZE"), cs[name].get("WRITE_SIZE")
            if calib and fs and ws:
                e["hbm_bytes_per_dispatch"] = (fs["per_dispatch"] * calib["fetch_factor"] +
                                               ws["per_dispatch"] * calib["write_factor"]) * 1024.0
                if e.get("mean_ns"):
                    e["hbm_gbs"] = e["hbm_bytes_per_dispatch"] / e["mean_ns"]
        doc["kernels"][name] = e
    total = sum(e.get("total_ns", 0.0) for e in doc["kernels"].values())
    doc["summary"] = {"kernels": len(doc["kernels"]), "total_kernel_ns": total, "calibration": calib}
    if total > 0:
        top = max(doc["kernels"], key=lambda k: doc["kernels"][k].get("total_ns", 0.0))
        doc["summary"]["dominant_kernel"] = top
        doc["summary"]["dominant_share"] = doc["kernels"][top]["total_ns"] / total
    return doc


def markdown(doc):
    lines = ["| kernel | calls | mean µs | total ms | share | HBM MB / dispatch | HBM GB/s |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    total = doc["summary"]["total_kernel_ns"] or 1.0
    rows = sorted(doc["kernels"].items(), key=lambda kv: -kv[1].get("total_ns", 0.0))
    for name, e in rows:
        mean = e.get("mean_ns")
        hb = e.get("hbm_bytes_per_dispatch")
        gbs = e.get("hbm_gbs")
        lines.append(f"| `{name}` | {e.get('calls', '')} | {'' if mean is None else f'{mean / 1e3:.1f}'} | "
                     f"{e.get('total_ns', 0.0) / 1e6:.3f} | {e.get('total_ns', 0.0) / total:.1%} | "
                     f"{'' if hb is None else f'{hb / 1e6:.2f}'} | {'' if gbs is None else f'{gbs:.0f}'} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--trace")
    ap.add_argument("--counters", nargs="*", default=[])
    ap.add_argument("--calibrate", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "BYTES"))
    ap.add_argument("--json")
    ap.add_argument("--md")
    a = ap.parse_args(argv)
    if not a.trace and not a.counters:
        ap.error("nothing to summarise: give --trace and/or --counters")
    calib = calibrate(a.calibrate[0], a.calibrate[1], float(a.calibrate[2])) if a.calibrate else None
    doc = summarise(a.trace, a.counters, calib)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(doc, f, indent=1)
    md = markdown(doc)
    if a.md:
        with open(a.md, "w") as f:
            f.write(md)
    if not a.json and not a.md:
        sys.stdout.write(md)
    return 0


if __name__ == "__main__":
    sys.exit(main())
