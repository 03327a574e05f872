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

    # HBM bytes per launch from the L2-to-memory request counts, by request size:
    #   pass A: --pmc TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_BUBBLE_sum TCC_EA0_WRREQ_sum
    #   pass B: --pmc TCC_EA0_WRREQ_64B_sum TCC_EA0_RDREQ_DRAM_sum TCC_EA0_WRREQ_DRAM_sum
    python3 multigraph_amd/profile_summary.py requests --a DIR --b DIR [--copy-a DIR --copy-mib 1024]
        [--json OUT] [--kernels a,b]

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


def counters_by_name(d):
    """{kernel: {counter: [value per dispatch]}} of a multi-counter run dir d"""
    out = {}
    for path in _find(d, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                k = _short(r["Kernel_Name"])
                c = out.setdefault(k, {}).setdefault(r["Counter_Name"], {})
                c[r["Dispatch_Id"]] = c.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    return {k: {c: list(v.values()) for c, v in cs.items()} for k, cs in out.items()}


def _mean(v):
    return sum(v) / len(v) if v else 0.0


def requests(a):
    """Bytes from request counts.  Writes: 64 B per TCC_EA0_WRREQ_64B, 32 B per other
    write request (the WRITE_SIZE rule).  Reads: 32 B per TCC_EA0_RDREQ_32B; the other
    read requests are 64 B or 128 B, and on gfx950 TCC_BUBBLE (documented as the
    128-byte requests) reads 0, so the size of each cannot be told apart: both bounds are
    reported, and `bytes` takes 128 B (an L2 line; a streaming copy of known size
    issues exactly one such request per 128 B, which the --copy-a run checks)."""
    A, B = counters_by_name(a.a), counters_by_name(a.b)
    res = {"kernels": {}}
    if a.copy_a:
        C = counters_by_name(a.copy_a)["k_copy"]
        rd = sorted(C["TCC_EA0_RDREQ_sum"])[len(C["TCC_EA0_RDREQ_sum"]) // 2]
        r32 = sorted(C["TCC_EA0_RDREQ_32B_sum"])[len(C["TCC_EA0_RDREQ_32B_sum"]) // 2]
        known = a.copy_mib * 1024 * 1024
        res["copy_check"] = {"copy_bytes": known, "read_requests": rd, "bytes_per_request": known / max(rd - r32, 1)}
        print(f"copy check: {a.copy_mib} MiB read by {rd:.0f} read requests ({r32:.0f} of 32 B): "
              f"{known / max(rd - r32, 1):.1f} B per request; TCC_BUBBLE reads {_mean(C['TCC_BUBBLE_sum']):.0f}\n")
    want = split_names(a.kernels) if a.kernels else sorted(set(A) | set(B))
    print("| kernel | dispatches | read req / launch | of 32 B | read MB (64 B .. 128 B) | from DRAM | write MB | "
          "HBM MB / launch (128 B reads) |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|")
    for k in want:
        ca, cb = A.get(k, {}), B.get(k, {})
        if not ca:
            continue
        rd = _mean(ca.get("TCC_EA0_RDREQ_sum", []))
        r32 = _mean(ca.get("TCC_EA0_RDREQ_32B_sum", []))
        wr = _mean(ca.get("TCC_EA0_WRREQ_sum", []))
        w64 = _mean(cb.get("TCC_EA0_WRREQ_64B_sum", [])) if cb else wr
        rdram = _mean(cb.get("TCC_EA0_RDREQ_DRAM_sum", [])) if cb else rd
        lo = 32.0 * r32 + 64.0 * (rd - r32)
        hi = 32.0 * r32 + 128.0 * (rd - r32)
        wb = 64.0 * w64 + 32.0 * max(wr - w64, 0.0)
        n = len(ca.get("TCC_EA0_RDREQ_sum", []))
        res["kernels"][k] = {"dispatches": n, "read_requests": rd, "read_requests_32b": r32, "read_bytes_lo": lo,
                             "read_bytes_hi": hi, "write_bytes": wb, "dram_read_share": rdram / rd if rd else 0.0,
                             "bytes": hi + wb}
        print(f"| {k} | {n} | {rd:.0f} | {r32:.0f} | {lo / 1e6:.1f} .. {hi / 1e6:.1f} | "
              f"{100 * rdram / rd if rd else 0:.0f} % | {wb / 1e6:.1f} | {(hi + wb) / 1e6:.1f} |")
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
    q = sub.add_parser("requests")
    q.add_argument("--a", required=True, help="pass A run dir")
    q.add_argument("--b", default="", help="pass B run dir")
    q.add_argument("--copy-a", dest="copy_a", default="", help="pass A of hbm_copy (check of the 128 B rule)")
    q.add_argument("--copy-mib", type=int, default=1024)
    q.add_argument("--kernels", default="")
    q.add_argument("--json", default="")
    a = ap.parse_args()
    if a.cmd == "stats":
        stats(a.dir, a.top, a.kernel, a.grid)
    elif a.cmd == "requests":
        if not a.b:
            a.b = a.a
        requests(a)
    else:
        traffic(a)


if __name__ == "__main__":
    main()
