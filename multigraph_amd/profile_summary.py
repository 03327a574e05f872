#!/usr/bin/env python3
"""Summaries of rocprofv3 CSV output for the engine's kernels.

    # per-kernel durations from a kernel-trace run (rocprofv3 --kernel-trace --stats
    # --output-format csv -d DIR -o NAME -- ...)
    python3 multigraph_amd/profile_summary.py stats DIR [--top 25] [--kernel K --grid G]

    # HBM bytes per launch from two counter runs of the workload and two of a plain
    # copy of known size (multigraph_amd/bin/hbm_copy MiB ...): every counter is scaled
    # by what it reads for the copy
    python3 multigraph_amd/profile_summary.py traffic --fetch DIR --write DIR \\
        --copy-fetch DIR --copy-write DIR --copy-mib 1024 [--json OUT] [--kernels a,b] [--grid G]

Output is markdown on stdout (and JSON with --json).  Only the Python standard
library is used.
"""
import argparse
import csv
import glob
import json
import os
import sys


def _find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        sys.exit(f"no *{suffix} under {d}")
    return hits


def _short(name):
    """kernel function name without return type, namespace, template arguments or
    parameters ("void (anonymous namespace)::k_pull<true>(int, ...)" -> "k_pull<true>"
    keeps the template arguments, which tell instantiations apart)"""
    n = name.replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    n = n.split("(")[0].strip()
    base, sep, targs = n.partition("<")
    return base.split("::")[-1] + sep + targs.replace("mg::", "")


def stats(d, top, kernel="", grid=0):
    """per-kernel durations: from kernel_stats.csv, or (with `kernel` and `grid`) from
    the dispatches of kernel_trace.csv of that kernel with that grid size only"""
    agg = {}
    if kernel:
        for path in _find(d, "kernel_trace.csv"):
            with open(path) as f:
                for r in csv.DictReader(f):
                    k = _short(r["Kernel_Name"])
                    if k != kernel or (grid and int(r["Grid_Size_X"]) != grid):
                        continue
                    ns = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
                    a = agg.setdefault(k, {"calls": 0, "total_ns": 0.0, "min_ns": float("inf"), "max_ns": 0.0})
                    a["calls"] += 1
                    a["total_ns"] += ns
                    a["min_ns"] = min(a["min_ns"], ns)
                    a["max_ns"] = max(a["max_ns"], ns)
    else:
        for path in _find(d, "kernel_stats.csv"):
            with open(path) as f:
                for r in csv.DictReader(f):
                    k = _short(r["Name"])
                    a = agg.setdefault(k, {"calls": 0, "total_ns": 0.0, "min_ns": float("inf"), "max_ns": 0.0})
                    a["calls"] += int(r["Calls"])
                    a["total_ns"] += float(r["TotalDurationNs"])
                    a["min_ns"] = min(a["min_ns"], float(r["MinNs"]))
                    a["max_ns"] = max(a["max_ns"], float(r["MaxNs"]))
    tot = sum(a["total_ns"] for a in agg.values()) or 1.0
    print("| kernel | calls | avg µs | min µs | max µs | total ms | % of GPU time |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for k, a in sorted(agg.items(), key=lambda kv: -kv[1]["total_ns"])[:top]:
        print(f"| {k} | {a['calls']} | {a['total_ns'] / a['calls'] / 1e3:.2f} | {a['min_ns'] / 1e3:.2f} | "
              f"{a['max_ns'] / 1e3:.2f} | {a['total_ns'] / 1e6:.3f} | {100 * a['total_ns'] / tot:.1f} |")
    return agg


def counters(d, grid=0):
    """{kernel: [value per dispatch]} of the single counter collected in run dir d
    (only dispatches of grid size `grid`, when given)"""
    out = {}
    for path in _find(d, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if grid and int(r["Grid_Size"]) != grid:
                    continue
                k = _short(r["Kernel_Name"])
                out.setdefault(k, {}).setdefault(r["Dispatch_Id"], 0.0)
                out[k][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in out.items()}


def split_names(spec):
    """comma-separated kernel names; commas inside template arguments belong to the name"""
    out, depth, cur = [], 0, ""
    for ch in spec:
        depth += (ch == "<") - (ch == ">")
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def traffic(a):
    fetch, write = counters(a.fetch, a.grid), counters(a.write, a.grid)
    cf, cw = counters(a.copy_fetch), counters(a.copy_write)
    known = a.copy_mib * 1024 * 1024
    # FETCH_SIZE / WRITE_SIZE are in KiB per dispatch; the copy reads and writes `known`
    # bytes per launch (its first launch warms up, so the median is used)
    med = lambda v: sorted(v)[len(v) // 2]
    f_scale = known / (med(cf["k_copy"]) * 1024.0)
    w_scale = known / (med(cw["k_copy"]) * 1024.0)
    print(f"copy calibration: {a.copy_mib} MiB read + written per launch; FETCH_SIZE reads "
          f"{med(cf['k_copy']) * 1024 / known:.4f}x and WRITE_SIZE {med(cw['k_copy']) * 1024 / known:.4f}x of it "
          f"(scale factors {f_scale:.4f}, {w_scale:.4f})\n")
    want = split_names(a.kernels) if a.kernels else sorted(set(fetch) | set(write))
    res = {"copy_mib": a.copy_mib, "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    print("| kernel | dispatches | fetched MB / launch | written MB / launch | HBM MB / launch |")
    print("|---|---:|---:|---:|---:|")
    for k in want:
        fv, wv = fetch.get(k, []), write.get(k, [])
        if not fv and not wv:
            continue
        fb = sum(fv) / max(len(fv), 1) * 1024.0 * f_scale
        wb = sum(wv) / max(len(wv), 1) * 1024.0 * w_scale
        res["kernels"][k] = {"dispatches": max(len(fv), len(wv)), "fetch_bytes": fb, "write_bytes": wb,
                             "bytes": fb + wb}
        print(f"| {k} | {max(len(fv), len(wv))} | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} |")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=25)
    s.add_argument("--kernel", default="", help="one kernel, from the dispatch trace")
    s.add_argument("--grid", type=int, default=0, help="with --kernel: only dispatches of this grid size")
    t = sub.add_parser("traffic")
    for n in ("fetch", "write", "copy_fetch", "copy_write"):
        t.add_argument("--" + n.replace("_", "-"), dest=n, required=True)
    t.add_argument("--copy-mib", type=int, default=1024)
    t.add_argument("--kernels", default="")
    t.add_argument("--json", default="")
    t.add_argument("--grid", type=int, default=0, help="only dispatches of this grid size (workload's)")
    a = ap.parse_args()
    if a.cmd == "stats":
        stats(a.dir, a.top, a.kernel, a.grid)
    else:
        traffic(a)


if __name__ == "__main__":
    main()
