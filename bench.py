#!/usr/bin/env python3
"""Headline benchmark: GTEPS and %-of-HBM-roofline of the frontier engine on RMAT-22.

One step = SSSP from the highest-degree vertex (weights uniform int in [1,64], seed 7)
plus connected components, on a symmetrized RMAT scale-22 graph (a,b,c,d =
.57,.19,.19,.05, edge factor 16, seed 1, self-loops and duplicates removed, isolated
vertices compacted).  value = (E + E_u) / step time, the reference's TEPS conventions
(SSSP counts directed edges E, CC the src<dst edges E_u; code/main.cu:46,124).

The JSON line also carries
  * roofline: the dominant kernel of the step (k_ps_gather), algorithmic bytes per
    launch over its average duration from HIP events recorded around its launches on
    the engine stream during the timed steps (only its: events around every kernel
    cost ~0.1 ms per step); an extra untimed step with every kernel profiled fills
    `kernels` and checks that it is the dominant one;
  * cpu_baseline: the oracle's single-thread restatement of code/validate.cu (SSSP
    Bellman-Ford + CC label propagation) on the same graph, timed on this host;
  * per_algo: SSSP, CC, BFS, BC, PR_T, PR_D on RMAT-22 with GTEPS, the single-thread
    CPU restatement of the same algorithm (code/validate.cu) timed beside each, and
    equiv_frac: one full pass's algorithmic bytes (SSSP 12E+8V, CC 16E_u+8V, BFS 8E+8V,
    BC 36E+24V) over the run time and the 8 TB/s peak.  It is a yardstick, not a
    bandwidth: a direction-optimised BFS or an Afforest CC touches far less than a full
    pass (measured bandwidths come from rocprofv3 counters, see profiles/).  A PageRank
    iteration is one full pass, so its hbm_frac_per_iter (8E+16V) is a bandwidth;
  * batched: 64 sources (uniform over the vertices, seed 3) run as one batch vs 64
    single-source runs, BFS, BC and SSSP, on RMAT-22, every row checked against the
    oracle;
  * pagerank_rmat24: PR_T and PR_D at damping 0.85, tolerance 1e-7, at most 100
    iterations on RMAT-24, with relative L1 and max-abs error against FP64 run for the
    same number of iterations;
  * uneven_batch: 32 graphs (RMAT-14..20, grids 256^2..2048^2) in one batched run,
    4 BFS / SSSP sources each, default options vs autotuned, every slot checked against
    the oracle (--no-uneven skips it).
Inputs are resident in HBM before every timed region.  The oracle (oracle/) is used
only as the checker and for the CPU columns.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS + % HBM roofline per algo on RMAT-22; batched vs per-source GTEPS"
DOMINANT = "k_ps_gather"  # the step's dominant kernel: one SSSP sweep over the partitioned layout
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_table(engines):
    tab = {}
    for eng in engines:
        for k, v in eng.kernel_stats().items():
            t = tab.setdefault(k, {"launches": 0, "ms": 0.0, "bytes": 0.0})
            for f in t:
                t[f] += v[f]
    return tab


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def frac(bytes_, ms):
    return bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS


def gteps(edges, ms):
    return edges / (ms * 1e-3) / 1e9


def measured_traffic(kernel, scale):
    """HBM bytes per launch of `kernel` on this workload from the committed rocprofv3
    counter summary (L2-to-memory request counts by size), or None"""
    path = os.path.join(ROOT, "profiles", "traffic_rmat22_sssp_cc.json")
    if scale != 22 or not os.path.exists(path):
        return None
    with open(path) as f:
        ks = json.load(f).get("kernels", {})
    # the trace names template instantiations ("k_ps_gather<unsigned char, unsigned short>")
    k = ks.get(kernel) or next((v for n, v in ks.items() if n.split("<")[0] == kernel), None)
    return float(k["bytes"]) if k else None


def rmat(mg, scale):
    from multigraph_amd import workloads

    t0 = time.time()
    g = workloads.rmat(mg, scale)
    log(f"RMAT-{scale}: nv={g.nv} ne={g.ne} built on the host in {time.time() - t0:.1f}s")
    return g


def cpu_col(edges, seconds, what):
    """single-thread CPU figure of the oracle's restatement of code/validate.cu"""
    return {"gteps": edges / seconds / 1e9, "s": seconds, "cores": 1, "kind": "port", "what": what}


def per_algo_extras(mg, np, orc, g, src, nv, ne, row, col):
    out = {}
    e = mg.Engine(g, mg.BFS)
    e.run(src)
    t = float(np.median([e.run(src) for _ in range(5)]))
    ok = bool(np.array_equal(e.result(mg.RES_LEVEL), orc.bfs(row, col, src)))
    t0 = time.perf_counter()
    orc.bfs(row, col, src)
    tc = time.perf_counter() - t0
    out["bfs"] = {"ms": t, "gteps": gteps(ne, t), "levels": e.iterations(), "equiv_frac": frac(8.0 * ne + 8.0 * nv, t),
                  "matches_oracle": ok, "cpu": cpu_col(ne, tc, "queue BFS, validate.cu:106-127")}
    e.close()
    e = mg.Engine(g, mg.BC)
    e.run(src)
    t = float(np.median([e.run(src) for _ in range(3)]))
    t0 = time.perf_counter()
    lv, sg, dl = orc.bc_fp64(row, col, src)
    tc = time.perf_counter() - t0
    d = e.result(mg.RES_DELTA).astype(np.float64)
    rel = float((np.abs(d - dl) / np.maximum(1.0, np.abs(dl))).max())
    ok = bool(np.array_equal(e.result(mg.RES_LEVEL), lv) and np.array_equal(e.result(mg.RES_SIGMA), sg)) and rel <= 1e-6
    out["bc"] = {"ms": t, "gteps": gteps(2 * ne, t), "equiv_frac": frac(36.0 * ne + 24.0 * nv, t),
                 "delta_max_rel_err": rel, "matches_oracle": ok,
                 "cpu": cpu_col(2 * ne, tc, "Brandes forward + backward, validate.cu:149-189 (FP64)")}
    e.close()
    for name, algo, variant in (("pr_t", mg.PR_T, 0), ("pr_d", mg.PR_D, 1)):
        e = mg.Engine(g, algo, pr_tolerance=1e-7, pr_max_iter=100)
        e.run(0)
        ms = e.run(0)
        its = max(e.iterations(), 1)
        # CPU: the reference's float rule for a bounded sample of 5 iterations
        t0 = time.perf_counter()
        orc.pr_f32(row, col, variant, 0.85, 1e-7, -5)
        tc = (time.perf_counter() - t0) / 5
        out[name] = {"ms": ms, "iterations": its, "ms_per_iter": ms / its, "gteps": gteps(ne, ms),
                     "gteps_per_iter": gteps(ne, ms / its), "hbm_frac_per_iter": frac(8.0 * ne + 16.0 * nv, ms / its),
                     "cpu": cpu_col(ne, tc, "one iteration (mean of 5) of the float rule, validate.cu:34-91")}
        e.close()
    return out


def batched(mg, np, orc, g, nv, ne, row, col, w):
    """64 sources (seed 3) as one batch vs 64 single-source runs; every row of the batch
    against the oracle (levels, sigma and the summed dependencies for BC)"""
    from multigraph_amd import workloads

    src = workloads.batch_sources(nv)
    out = {"sources": "64 uniform over the compacted vertices, numpy default_rng(3)"}
    for name, kind, algo, te in (("bfs", mg.BATCH_BFS, mg.BFS, 1), ("bc", mg.BATCH_BC, mg.BC, 2),
                                 ("sssp", mg.BATCH_SSSP, mg.SSSP, 1)):
        b = mg.Batch(g, kind)
        b.run(src)
        tb = float(np.median([b.run(src) for _ in range(3)]))
        levels = b.sweeps()
        lv = b.levels()
        rec = {}
        if name == "bfs":
            ok = bool(np.array_equal(lv, orc.bfs_multi(row, col, src)))
        elif name == "sssp":
            ok = bool(np.array_equal(lv, orc.sssp_dijkstra_multi(row, col, w, src)))
        else:
            rl, rs, rbc = orc.bc_fp64_multi(row, col, src)
            bc = b.bc().astype(np.float64)
            err = float((np.abs(bc - rbc) / np.maximum(1.0, np.abs(rbc))).max())
            ok = bool(np.array_equal(lv, rl) and np.array_equal(b.sigma(), rs)) and err <= 1e-5
            rec["bc_sum_max_rel_err"] = err
            del rl, rs
        e = mg.Engine(g, algo)
        e.run(int(src[0]))
        ts = [e.run(int(s)) for s in src]
        tsum = float(np.sum(ts))
        rec.update({"batch_ms": tb, "singles_ms": tsum, "batched_gteps": gteps(64 * te * ne, tb),
                    "per_source_gteps": gteps(64 * te * ne, tsum), "speedup": tsum / tb, "levels": levels,
                    "rows_match_oracle": ok})
        if name == "bfs":
            rec["equiv_frac"] = frac((12.0 * ne + 20.0 * nv) * levels, tb)
        b.close()
        e.close()
        out[name] = rec
        log(f"batched {name}: {tb:.2f} ms vs {tsum:.2f} ms of singles ({tsum / tb:.2f}x), oracle {'ok' if ok else 'MISMATCH'}")
    return out


def uneven_oracle(np, orc, graphs, offs, src, slots, lv, kind):
    """every slot of every member graph against the oracle"""
    for gi, p in enumerate(graphs):
        row, col, w = p.csr()
        o = offs[gi]
        loc = src[4 * gi:4 * gi + 4] - o
        ref = orc.bfs_multi(row, col, loc) if kind == "bfs" else orc.sssp_dijkstra_multi(row, col, w, loc)
        for j in range(4):
            if not np.array_equal(lv[slots[4 * gi + j], o:o + p.nv], ref[j]):
                return False
    return True


def uneven(mg, np, orc):
    """32 graphs in one batch: RMAT-14..20 (4 seeds each) and 4-neighbour grids 256^2 ..
    2048^2 joined as a disjoint union, 4 sources per graph sharing 4 slots of one run;
    default options vs the autotuner's"""
    from multigraph_amd import workloads

    t0 = time.time()
    u, graphs, offs, src, slots = workloads.uneven(mg)
    ne = u.ne
    log(f"uneven batch: {len(graphs)} graphs, nv={u.nv} ne={ne} ({time.time() - t0:.1f}s)")
    out = {"graphs": workloads.UNEVEN_DESC, "nv": u.nv, "ne": ne,
           "sources": "4 per graph, numpy default_rng(5), slots 0..3 shared across graphs"}
    for name, kind in (("bfs", mg.BATCH_BFS), ("sssp", mg.BATCH_SSSP)):
        b = mg.Batch(u, kind)
        b.run(src, slots)
        t_def = float(np.median([b.run(src, slots) for _ in range(3)]))
        lv_def = b.levels()
        levels = b.sweeps()
        st_def = b.stats()
        b.close()
        log(f"uneven {name}: untuned {t_def:.2f} ms ({st_def}), tuning")
        t1 = time.time()
        opt, _, _ = mg.batch_autotune(u, kind, src, slots, reps=2)
        tune_s = time.time() - t1
        b = mg.Batch(u, kind, options=opt)
        b.run(src, slots)
        t_tun = float(np.median([b.run(src, slots) for _ in range(3)]))
        same = bool(np.array_equal(b.levels(), lv_def))
        b.close()
        ok = uneven_oracle(np, orc, graphs, offs, src, slots, lv_def, name)
        out[name] = {"untuned_ms": t_def, "tuned_ms": t_tun, "untuned_gteps": gteps(4 * ne, t_def),
                     "tuned_gteps": gteps(4 * ne, t_tun), "levels": levels, "tuning_s": tune_s,
                     "results_equal": same, "slots_match_oracle": ok,
                     "tuned_options": {"alpha": opt.alpha, "beta": opt.beta, "batch_chunk": opt.batch_chunk,
                                       "grid_blocks": opt.grid_blocks, "graph_levels": opt.graph_levels,
                                       "persist_levels": opt.persist_levels, "sssp_delta": opt.sssp_delta,
                                       "batch_light": opt.batch_light, "sparse_grid": opt.sparse_grid},
                     "untuned_stats": st_def}
        log(f"uneven {name}: {t_def:.2f} ms untuned, {t_tun:.2f} ms tuned ({tune_s:.1f}s of tuning)")
    return out


def pagerank_big(mg, np, orc, scale):
    """time per iteration, and the ranks' relative L1 and max-abs error against the FP64
    oracle run for the same number of iterations"""
    g = rmat(mg, scale)
    nv, ne = g.nv, g.ne
    row, col, _ = g.csr()
    out = {"graph": f"RMAT-{scale}", "nv": nv, "ne": ne}
    for name, algo, variant in (("pr_t", mg.PR_T, 0), ("pr_d", mg.PR_D, 1)):
        e = mg.Engine(g, algo, pr_tolerance=1e-7, pr_max_iter=100)
        e.run(0)
        e.set_profiling(True)
        ms_prof = e.run(0)
        ks = e.kernel_stats()
        e.set_profiling(False)
        ms = e.run(0)
        its = max(e.iterations(), 1)
        r = e.result(mg.RES_RANK).astype(np.float64)
        ref, _ = orc.pr_fp64_pull(row, col, variant, 0.85, 1e-7, -its)
        l1 = float(np.abs(r - ref).sum() / np.abs(ref).sum())
        mx = float(np.abs(r - ref).max())
        del r, ref
        rec = {"ms": ms, "iterations": its, "ms_per_iter": ms / its, "gteps_per_iter": gteps(ne, ms / its),
               "hbm_frac_per_iter": frac(8.0 * ne + 16.0 * nv, ms / its), "l1_rel_vs_fp64": l1,
               "max_abs_vs_fp64": mx,
               "kernels": {k: {"avg_ms": v["ms"] / max(v["launches"], 1), "launches": v["launches"]}
                           for k, v in ks.items()}, "profiled_ms": ms_prof}
        e.close()
        out[name] = rec
        log(f"{name} RMAT-{scale}: {ms / its:.3f} ms/iter over {its} iterations, L1 {l1:.2e}, max abs {mx:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--pr-scale", type=int, default=24, help="graph scale of the PageRank figures (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip per-algorithm, batched and PageRank figures")
    ap.add_argument("--no-uneven", action="store_true", help="skip the 32-graph uneven batch (RMAT-14..20 + grids)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the engine has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import multigraph_amd as mg

    if mg.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    g = rmat(mg, args.scale)
    nv, ne, _ = g.info()
    row, col, w = g.csr()
    ne_u = int(np.count_nonzero(col > np.repeat(np.arange(nv, dtype=np.int32), np.diff(row))))
    src = g.max_degree_vertex()

    sssp = mg.Engine(g, mg.SSSP)
    cc = mg.Engine(g, mg.CC)
    for _ in range(args.warmup):
        sssp.run(src)
        cc.run(0)
    log("warmup done")
    # HIP events bracket the launches of the step's dominant kernel only (the rocprofv3
    # trace in profiles/k_ps_gather_rmat22_sssp_cc.md; an untimed profiled step below
    # confirms it): events around every kernel cost ~0.1 ms per step
    sssp.set_profiling(True, only=DOMINANT)
    cc.set_profiling(False)
    ms_sssp, ms_cc = [], []
    mg.lib().mg_sync(sssp._h)
    mg.lib().mg_sync(cc._h)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ms_sssp.append(sssp.run(src))
        ms_cc.append(cc.run(0))
    mg.lib().mg_sync(sssp._h)
    mg.lib().mg_sync(cc._h)
    elapsed = time.perf_counter() - t_start
    log(f"timed {args.steps} steps in {elapsed:.3f}s")
    ms_per_step = elapsed * 1000.0 / args.steps
    value = (ne + ne_u) / (ms_per_step * 1e-3) / 1e9
    dom_name = DOMINANT
    dom = kernel_table([sssp])[dom_name]
    avg_ms = dom["ms"] / dom["launches"]
    avg_bytes = dom["bytes"] / dom["launches"]
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    sssp_iters = sssp.iterations()
    # one more step, untimed, with every kernel profiled: the per-kernel table
    sssp.set_profiling(True)
    cc.set_profiling(True)
    sssp.run(src)
    cc.run(0)
    kt = kernel_table([sssp, cc])
    dominant_ok = max(kt, key=lambda k: kt[k]["ms"]) == DOMINANT

    # correctness of what was timed, against the oracle: SSSP distances (Dijkstra) and
    # CC labels (union-find, min label per component), bit-exact
    from oracle import oracle as orc

    d = sssp.result(mg.RES_LEVEL)
    lab = cc.result(mg.RES_LEVEL)
    ok = bool(np.array_equal(d, orc.sssp_dijkstra(row, col, w, src)))
    ok = ok and bool(np.array_equal(lab, orc.cc_unionfind(row, col)[0]))
    sssp.set_profiling(False)
    cc.set_profiling(False)

    t_sssp, t_cc = float(np.mean(ms_sssp)), float(np.mean(ms_cc))
    per_algo = {
        "sssp": {"ms": t_sssp, "gteps": gteps(ne, t_sssp), "iterations": sssp_iters,
                 "equiv_frac": frac(12.0 * ne + 8.0 * nv, t_sssp)},
        "cc": {"ms": t_cc, "gteps": gteps(ne_u, t_cc), "equiv_frac": frac(16.0 * ne_u + 8.0 * nv, t_cc)},
    }
    sssp.close()
    cc.close()

    batch = None
    pr_big = None
    uneven_batch = None
    if not args.no_extra:
        per_algo.update(per_algo_extras(mg, np, orc, g, src, nv, ne, row, col))
        log("per-algorithm figures done")
        batch = batched(mg, np, orc, g, nv, ne, row, col, w)

    cpu = None
    if not args.no_cpu:
        t1 = time.perf_counter()
        ref, sweeps = orc.sssp_bf(row, col, w, src)
        t2 = time.perf_counter()
        lp, ncomp, _ = orc.cc_labelprop(row, col)
        t3 = time.perf_counter()
        ok = ok and bool(np.array_equal(ref, d)) and bool(np.array_equal(orc.canon(lp), orc.canon(lab)))
        cpu = {"value": (ne + ne_u) / (t3 - t1) / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"the full step on the same RMAT-{args.scale} graph: Bellman-Ford SSSP ({sweeps} sweeps, "
                         f"{t2 - t1:.2f}s) + label-propagation CC ({t3 - t2:.2f}s), code/validate.cu restated, "
                         f"1 thread on {cpu_model()}"}
        per_algo["sssp"]["cpu"] = cpu_col(ne, t2 - t1, "in-place Bellman-Ford, validate.cu:131-146")
        per_algo["cc"]["cpu"] = cpu_col(ne_u, t3 - t2, "label propagation + relabel, validate.cu:223-306")
        log("cpu baseline done")

    del g, row, col, w, d, lab
    if not args.no_extra and args.pr_scale > 0:
        pr_big = pagerank_big(mg, np, orc, args.pr_scale)
    if not args.no_extra and not args.no_uneven:
        uneven_batch = uneven(mg, np, orc)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"SSSP (source = max-degree vertex, int weights U[1,64] seed 7) + CC on symmetrized "
                               f"RMAT-{args.scale} (a=.57 b=.19 c=.19, edge factor 16, seed 1)",
                   "scale": args.scale, "edge_factor": 16, "nv": nv, "ne": ne, "ne_u": ne_u, "source": int(src)},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom_name, args.scale),
                     "traffic_source": "profiles/traffic_rmat22_sssp_cc.json (rocprofv3 L2-to-memory requests: 128 B per read "
                                       "request, the size a copy of known bytes issues; 64 B per 64-byte write request, "
                                       "32 B per other)",
                     "avg_launch_ms": avg_ms,
                     "bytes_per_launch": avg_bytes, "launches": dom["launches"],
                     "dominant_in_profiled_step": dominant_ok},
        "cpu_baseline": cpu,
        "per_algo": per_algo,
        "batched": batch,
        "pagerank_rmat24": pr_big,
        "uneven_batch": uneven_batch,
        "kernels": {k: {"launches": v["launches"], "avg_ms": v["ms"] / max(v["launches"], 1),
                        "gbs": v["bytes"] / max(v["ms"], 1e-9) / 1e6} for k, v in kt.items()},
        "verified": ok,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
