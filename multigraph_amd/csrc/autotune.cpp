// Autotuner: picks mg_options for one graph and algorithm by timing candidates on the
// graph itself.  The reference re-decides push vs pull every level from wall-clock
// samples of both (process.cu:1161-1255, pseudo_* kernels) and fixes its degree tiers
// and block sizes at compile time (common.h:10-61, process.cu:522-537); here the
// decision rule is deterministic and its thresholds, the bin edges of the degree tiers
// (push split, pull split, BC lane / wave / chunk tiers, batched lane / wave split), the
// workgroup size of the level kernels, the grid shape, the levels per captured graph,
// the PageRank partition size, the batched chunk size and lane order, and whether
// batched sparse levels run in a persistent launch, are searched once.
//
// Search: coordinate descent from the starting options, one pass over the
// coordinates in the order listed; each candidate is built, run once to warm up, then
// timed `reps` times (median, HIP events inside mg_run / mg_batch_run).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "mg_internal.hpp"

namespace {

struct Coord {
    std::function<void(mg_options &, double)> set;
    std::vector<double> values;
};

float median(std::vector<float> v) {
    std::sort(v.begin(), v.end());
    return v.empty() ? 0.f : v[v.size() / 2];
}

// cost of one option set: summed median times over the sources (engines) or of the
// batch run; < 0 on error
using CostFn = std::function<int(const mg_options &, float *)>;

int descend(const std::vector<Coord> &coords, const CostFn &cost, mg_options *best, float *ms_default, float *ms_tuned) {
    float base = 0.f;
    int rc = cost(*best, &base);
    if (rc != MG_OK) return rc;
    if (ms_default) *ms_default = base;
    float cur = base;
    for (const Coord &c : coords) {
        mg_options keep = *best;
        for (double v : c.values) {
            mg_options o = *best;
            c.set(o, v);
            if (memcmp(&o, best, sizeof(o)) == 0) continue;
            float t = 0.f;
            if (cost(o, &t) != MG_OK) continue;  // a candidate that does not fit is skipped
            if (t < cur) { cur = t; keep = o; }
        }
        *best = keep;
    }
    if (ms_tuned) *ms_tuned = cur;
    return MG_OK;
}

}  // namespace

extern "C" int mg_autotune(const mg_graph *g, int algo, int nsrc, const int32_t *sources, int reps, mg_options *best,
                           float *ms_default, float *ms_tuned) {
    if (!g || !best || (nsrc > 0 && !sources)) return mg_fail(MG_ERR_ARG, "null argument");
    if (algo < MG_BFS || algo > MG_PR_D) return mg_fail(MG_ERR_ARG, "unknown algorithm %d", algo);
    if (reps < 1) reps = 1;
    std::vector<int> src(sources, sources + std::max(0, nsrc));
    if (src.empty()) src.push_back(g->start_p);
    CostFn cost = [&](const mg_options &o, float *out) -> int {
        mg_engine *e = nullptr;
        int rc = mg_build(g, algo, &o, &e);
        if (rc != MG_OK) return rc;
        float total = 0.f;
        for (int s : src) {
            float ms = 0.f;
            if ((rc = mg_run(e, s, &ms)) != MG_OK) break;  // warm-up (graph capture, caches)
            std::vector<float> t;
            for (int r = 0; r < reps && rc == MG_OK; ++r) { rc = mg_run(e, s, &ms); t.push_back(ms); }
            total += median(t);
            if (algo == MG_CC || algo == MG_PR_T || algo == MG_PR_D) break;  // no source
        }
        mg_release(e);
        *out = total;
        return rc;
    };
    std::vector<Coord> coords;
    auto grid = Coord{[](mg_options &o, double v) { o.grid_blocks = (int)v; }, {1024, 2048, 4096}};
    if (algo == MG_BFS || algo == MG_SSSP || algo == MG_BC) {
        coords.push_back({[](mg_options &o, double v) { o.alpha = (float)v; }, {4, 8, 15, 30, 60}});
        coords.push_back({[](mg_options &o, double v) { o.beta = (float)v; }, {6, 12, 18, 36, 72}});
        coords.push_back(grid);
        coords.push_back({[](mg_options &o, double v) { o.graph_levels = (int)v; }, {0, 4, 16}});
        // bin edges and block shape: push split of high-degree vertices, the workgroup
        // size of the level kernels, the pull's chunk split and (BC) the lane / wave /
        // chunk tiers of the pull and the backward pass
        coords.push_back({[](mg_options &o, double v) { o.huge_degree = (int)v; }, {4096, 8192, 16384, 32768}});
        coords.push_back({[](mg_options &o, double v) { o.block_threads = (int)v; }, {256, 512, 1024}});
        if (algo != MG_SSSP)
            coords.push_back({[](mg_options &o, double v) { o.pull_split = (int)v; }, {1024, 2048, 4096, 8192}});
        if (algo == MG_BC) {
            coords.push_back({[](mg_options &o, double v) { o.lane_edges = (int)v; }, {16, 32, 64}});
            coords.push_back({[](mg_options &o, double v) { o.back_split = (int)v; }, {1024, 2048, 4096}});
            coords.push_back({[](mg_options &o, double v) { o.back_chunk = (int)v; }, {512, 1024, 2048}});
        }
    } else if (algo == MG_CC) {
        coords.push_back(grid);
    } else {
        coords.push_back({[](mg_options &o, double v) { o.part_log2 = (int)v; }, {0, 11, 12, 13}});
        coords.push_back({[](mg_options &o, double v) { o.graph_levels = (int)v; }, {4, 8, 16}});
    }
    return descend(coords, cost, best, ms_default, ms_tuned);
}

extern "C" int mg_batch_autotune(const mg_graph *g, int kind, int n, const int32_t *sources, const int32_t *slots, int reps,
                                 mg_options *best, float *ms_default, float *ms_tuned) {
    if (!g || !best || n < 1 || !sources) return mg_fail(MG_ERR_ARG, "null argument");
    if (reps < 1) reps = 1;
    CostFn cost = [&](const mg_options &o, float *out) -> int {
        mg_batch *b = nullptr;
        int rc = mg_batch_build(g, kind, 64, &o, &b);
        if (rc != MG_OK) return rc;
        float ms = 0.f;
        auto run = [&]() { return slots ? mg_batch_run_slots(b, n, sources, slots, &ms) : mg_batch_run(b, n, sources, &ms); };
        rc = run();  // warm-up (graph capture)
        std::vector<float> t;
        for (int r = 0; r < reps && rc == MG_OK; ++r) { rc = run(); t.push_back(ms); }
        mg_batch_release(b);
        *out = median(t);
        return rc;
    };
    std::vector<Coord> coords;
    coords.push_back({[](mg_options &o, double v) { o.alpha = (float)v; }, {4, 8, 15, 30, 60, 120}});
    if (kind != MG_BATCH_BC)  // the sparse-level threshold (queue levels below (nv + E) / beta)
        coords.push_back({[](mg_options &o, double v) { o.beta = (float)v; }, {4, 9, 18, 36, 72}});
    coords.push_back({[](mg_options &o, double v) { o.batch_chunk = (int)v; }, {64, 128, 256, 512, 1024, 4096}});
    coords.push_back({[](mg_options &o, double v) { o.grid_blocks = (int)v; }, {1024, 2048, 4096}});
    if (kind == MG_BATCH_BFS)  // lane / wave split of the pull's in-edge rows
        coords.push_back({[](mg_options &o, double v) { o.batch_light = (int)v; }, {16, 32, 64}});
    if (kind != MG_BATCH_BC) {
        coords.push_back({[](mg_options &o, double v) { o.graph_levels = (int)v; }, {2, 8, 32}});
        coords.push_back({[](mg_options &o, double v) { o.persist_levels = (int)v; }, {0, -1}});
        coords.push_back({[](mg_options &o, double v) { o.sparse_grid = (int)v; }, {0, 64, 128}});
    }
    if (kind == MG_BATCH_SSSP) {
        // near-far band (0: a quarter of the mean weight; -1: off), in steps of the mean
        // weight.  Levels fall as the band widens until it passes the weight that shortest
        // paths typically step by, where re-lowered pairs multiply (the 32-graph uneven
        // batch, weights 1..64: 6742 levels at 8, 3853 at 14, but 2.7x the queued pairs at 15)
        double sw = 0.0;
        for (long long p = 0; p < g->ne; ++p) sw += (double)g->w[p];
        const double mw = g->ne > 0 ? sw / (double)g->ne : 1.0;
        std::vector<double> dv = {-1, 0};
        for (const double f : {0.125, 0.3125, 0.375, 0.4375, 0.5}) {
            const double d = std::max(1.0, std::floor(mw * f));
            if (std::find(dv.begin(), dv.end(), d) == dv.end()) dv.push_back(d);
        }
        coords.push_back({[](mg_options &o, double v) { o.sssp_delta = (int)v; }, dv});
    }
    if (!slots)  // lane order of the sources (a slot list fixes the lanes itself)
        coords.push_back({[](mg_options &o, double v) { o.batch_lanes = (int)v; }, {0, 1}});
    return descend(coords, cost, best, ms_default, ms_tuned);
}
