// C ABI of the device engine (include/multigraph_amd.h): build, reset, step, run,
// results, release, profiling.  Dispatches to the per-algorithm code in
// frontier.hip (BFS/SSSP/BC) and cc_pr.hip (CC, PageRank).
#include <algorithm>
#include <chrono>
#include <cstring>
#include "engine.hpp"

using namespace mg;

void mg_free_dev(void *p) {
    if (p) (void)hipFree(p);
}

extern "C" void mg_default_options(mg_options *o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->mode = MG_MODE_HYBRID;
    o->alpha = 0.0f;           // 0: the path's own (kAlpha / kBeta for engines, 15 / 18 for batched runs)
    o->beta = 0.0f;
    o->huge_degree = 8192;
    o->grid_blocks = 0;        // 2048 for single-source engines (4096 for BC and CC), 4096 for batched runs
    o->graph_levels = 0;
    o->pr_damping = 0.85f;     // common.h:59
    o->pr_tolerance = 0.01f;   // common.h:60
    o->pr_max_iter = 0;        // the reference runs until no vertex changes
    o->part_log2 = 0;          // automatic
}

static void fill_defaults(mg_options *o, int algo) {
    mg_options d;
    mg_default_options(&d);
    if (o->alpha <= 0) o->alpha = kAlpha;
    if (o->beta <= 0) o->beta = kBeta;
    if (o->huge_degree <= 0) o->huge_degree = d.huge_degree;
    if (o->huge_degree < kHugeChunk) o->huge_degree = kHugeChunk;
    // BC's pulls and backward tiers and CC's chunk hooks fill a 4096-workgroup grid
    // better (RMAT-22: BC from 8 random sources 1.79 -> 1.70 ms, CC 0.141 -> 0.139 ms);
    // BFS and SSSP levels are a little faster on 2048
    if (o->grid_blocks <= 0) o->grid_blocks = (algo == MG_BC || algo == MG_CC) ? 4096 : 2048;
    if (o->pr_damping <= 0) o->pr_damping = d.pr_damping;
    if (o->pr_tolerance <= 0) o->pr_tolerance = d.pr_tolerance;
    if (o->lane_edges <= 0) o->lane_edges = 32;
    o->lane_edges = std::min(o->lane_edges, 1024);
    if (o->pull_split <= 0) o->pull_split = kPullHuge;
    o->pull_split = std::max(o->pull_split, 64);
    if (o->back_chunk <= 0) o->back_chunk = 1024;
    o->back_chunk = std::max(o->back_chunk, 64);
    if (o->back_split <= 0) o->back_split = 2048;
    o->back_split = std::max(o->back_split, o->lane_edges + 1);
    if (o->block_threads <= 0) o->block_threads = kBlock;
    if (o->batch_light <= 0) o->batch_light = 32;
}

extern "C" int mg_device_count(int *n) {
    if (!n) return mg_fail(MG_ERR_ARG, "null argument");
    int c = 0;
    hipError_t err = hipGetDeviceCount(&c);
    if (err != hipSuccess) { *n = 0; return mg_fail(MG_ERR_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(err)); }
    *n = c;
    return MG_OK;
}

static void release_all(mg_engine *e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->gexec) hipGraphExecDestroy(e->gexec);
    void *ptrs[] = {e->pull_ch, e->pull_hv, e->pull_found, e->pull_sum, e->part_n, e->part_mf, e->eoff, e->tsum, e->cstart, e->pcstart, e->row, e->row64, e->col, e->w, e->trow, e->tcol, e->tw, e->level, e->vfc, e->sigma, e->delta, e->dacc, e->lev8, e->bc_cw, e->q, e->bm[0], e->bm[1],
                    e->vis, e->chunks, e->lvl_off, e->lvl_mf, e->aux, e->ctl, e->urow, e->ucol, e->rank, e->cval, e->deg, e->psum,
                    e->pb.a_src, e->pb.b_dst, e->pb.amap, e->pb.msg, e->pb.gbase, e->pb.unitsA,
                    e->pb.unitsB, e->pb.split_parts, e->pb.apart, e->pb.b_w, e->pb.inv, e->pb.unitsS};
    for (void *p : ptrs) mg_free_dev(p);
    if (e->h_ctl) (void)hipHostFree(e->h_ctl);
    for (auto &ev : e->ev_pool) hipEventDestroy(ev);
    for (auto &p : e->pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int mg_build(const mg_graph *g, int algo, const mg_options *opt, mg_engine **out) {
    if (!g || !out) return mg_fail(MG_ERR_ARG, "null argument");
    if (algo < MG_BFS || algo > MG_PR_D) return mg_fail(MG_ERR_ARG, "unknown algorithm %d", algo);
    *out = nullptr;
    if (opt && opt->block_threads != 0 && opt->block_threads != 256 && opt->block_threads != 512 && opt->block_threads != 1024)
        return mg_fail(MG_ERR_ARG, "block_threads %d: 256, 512 or 1024", opt->block_threads);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return mg_fail(MG_ERR_DEVICE, "no HIP device available");
    mg_engine *e = new mg_engine();
    e->algo = algo;
    if (opt) e->opt = *opt; else mg_default_options(&e->opt);
    fill_defaults(&e->opt, algo);
    e->nv = g->nv;
    e->ne = g->ne;
    auto fail = [&](int rc) { release_all(e); return rc; };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return fail(mg_fail(MG_ERR_DEVICE, "hipStreamCreate failed"));
    const long long nv1 = (long long)g->nv + 1;
    const long long ne = g->ne > 0 ? g->ne : 1;
    e->wide = g->wide() || e->opt.wide_offsets == 1;
    if (e->wide) {
        // 64-bit row offsets: the BFS / SSSP frontier kernels have them (symmetric
        // graphs: pull levels read a row as in-edges); every push level's edges must
        // stay below 2^31 (the edge-balanced scan counts them in int), so pushes stop
        // at E / alpha with alpha >= 4 (E <= 2^32) and a push-only mode is refused
        if (algo != MG_BFS && algo != MG_SSSP)
            return fail(mg_fail(MG_ERR_LIMIT, "64-bit row offsets (%lld edges): BFS and SSSP only", (long long)g->ne));
        if (!g->symmetric) return fail(mg_fail(MG_ERR_LIMIT, "64-bit row offsets need a symmetric graph"));
        if (e->opt.mode == MG_MODE_SPARSE || e->opt.alpha < 4.0f)
            return fail(mg_fail(MG_ERR_LIMIT, "64-bit row offsets: push-only levels, or alpha below 4, are not supported"));
        std::vector<long long> r((size_t)nv1);
        for (long long v = 0; v < nv1; ++v) r[(size_t)v] = g->off(v);
        if (hipMalloc(&e->row64, sizeof(long long) * nv1) != hipSuccess || hipMalloc(&e->col, sizeof(int) * ne) != hipSuccess)
            return fail(mg_fail(MG_ERR_NOMEM, "device allocation of the CSR failed"));
        if (hipMemcpy(e->row64, r.data(), sizeof(long long) * nv1, hipMemcpyHostToDevice) != hipSuccess ||
            (g->ne > 0 && hipMemcpy(e->col, g->col.data(), sizeof(int) * g->ne, hipMemcpyHostToDevice) != hipSuccess))
            return fail(mg_fail(MG_ERR_DEVICE, "CSR upload failed"));
    } else {
        if (hipMalloc(&e->row, sizeof(int) * nv1) != hipSuccess || hipMalloc(&e->col, sizeof(int) * ne) != hipSuccess)
            return fail(mg_fail(MG_ERR_NOMEM, "device allocation of the CSR failed"));
        if (hipMemcpy(e->row, g->row.data(), sizeof(int) * nv1, hipMemcpyHostToDevice) != hipSuccess ||
            (g->ne > 0 && hipMemcpy(e->col, g->col.data(), sizeof(int) * g->ne, hipMemcpyHostToDevice) != hipSuccess))
            return fail(mg_fail(MG_ERR_DEVICE, "CSR upload failed"));
    }
    if (algo == MG_SSSP) {
        if (hipMalloc(&e->w, sizeof(int) * ne) != hipSuccess) return fail(mg_fail(MG_ERR_NOMEM, "weight allocation failed"));
        if (g->ne > 0 && hipMemcpy(e->w, g->w.data(), sizeof(int) * g->ne, hipMemcpyHostToDevice) != hipSuccess)
            return fail(mg_fail(MG_ERR_DEVICE, "weight upload failed"));
    }
    int rc;
    switch (algo) {
    case MG_CC: rc = mg_cc_build(e, g); break;
    case MG_PR_T:
    case MG_PR_D: rc = mg_pr_build(e, g); break;
    default: rc = mg_frontier_build(e, g); break;
    }
    if (rc != MG_OK) return fail(rc);
    if (hipStreamSynchronize(e->stream) != hipSuccess) return fail(mg_fail(MG_ERR_DEVICE, "build failed on the device"));
    e->source = g->start_p;
    *out = e;
    return MG_OK;
}

extern "C" int mg_release(mg_engine *e) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    release_all(e);
    return MG_OK;
}

extern "C" int mg_reset(mg_engine *e, int source) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    int rc;
    switch (e->algo) {
    case MG_CC: rc = mg_cc_reset(e); break;
    case MG_PR_T:
    case MG_PR_D: rc = mg_pr_reset(e); break;
    default:
        if (e->nv > 0 && (source < 0 || source >= e->nv)) return mg_fail(MG_ERR_ARG, "source %d outside [0,%d)", source, e->nv);
        rc = mg_frontier_reset(e, source);
        break;
    }
    if (rc == MG_OK) e->ready = true;
    return rc;
}

extern "C" int mg_step(mg_engine *e, int *iter) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    if (!e->ready) return mg_fail(MG_ERR_STATE, "mg_step before mg_reset");
    int more = 0;
    switch (e->algo) {
    case MG_CC: {
        e->last_mode = 0;
        if (e->done) { more = 0; break; }
        int rc = mg_cc_step(e, &more);
        if (rc != MG_OK) return rc;
        break;
    }
    case MG_PR_T:
    case MG_PR_D: {
        e->last_mode = 0;
        int rc = mg_pr_step(e, &more);
        if (rc != MG_OK) return rc;
        break;
    }
    default: {
        e->last_mode = e->done ? -1 : (e->h_ctl->mode == mg::MODE_BU ? 0 : 1);
        int rc = mg_frontier_level(e);
        if (rc != MG_OK) return rc;
        more = e->done ? 0 : 1;
        break;
    }
    }
    if (iter) *iter = e->iter;
    return more;
}

extern "C" int mg_step_mode(const mg_engine *e, int *mode) {
    if (!e || !mode) return mg_fail(MG_ERR_ARG, "null argument");
    *mode = e->last_mode;
    return MG_OK;
}

extern "C" int mg_bc_backward(mg_engine *e, int iter) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    if (e->algo != MG_BC) return mg_fail(MG_ERR_STATE, "mg_bc_backward needs a BC engine");
    if (!e->done) return mg_fail(MG_ERR_STATE, "mg_bc_backward before the forward pass finished");
    (void)iter;
    int rc = mg_bc_backward_impl(e);
    if (rc != MG_OK) return rc;
    MG_HIP(hipStreamSynchronize(e->stream));
    return MG_OK;
}

extern "C" int mg_run(mg_engine *e, int source, float *ms) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    MG_TRY(mg_reset(e, source));
    MG_HIP(hipStreamSynchronize(e->stream));
    hipEvent_t a, b;
    MG_HIP(hipEventCreate(&a));
    MG_HIP(hipEventCreate(&b));
    MG_HIP(hipEventRecord(a, e->stream));
    const bool frontier = e->algo == MG_BFS || e->algo == MG_SSSP || e->algo == MG_BC;
    // PageRank over the partitioned layout always replays captured iterations (8 per
    // host check unless graph_levels says otherwise); profiling times launches eagerly
    const bool pr_graph = (e->algo == MG_PR_T || e->algo == MG_PR_D) && e->pr_pb;
    const int K = frontier ? e->opt.graph_levels : (pr_graph ? (e->opt.graph_levels > 0 ? e->opt.graph_levels : 8) : 0);
    int rc = MG_OK;
    if ((frontier || pr_graph) && K > 0 && !e->prof) {
        if (!e->gexec || e->gexec_levels != K) {
            if (e->gexec) { hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
            rc = frontier ? mg_frontier_capture(e, K, &e->gexec) : mg_pr_capture(e, K, &e->gexec);
            e->gexec_levels = K;
        }
        while (rc == MG_OK) {
            if (hipGraphLaunch(e->gexec, e->stream) != hipSuccess) { rc = mg_fail(MG_ERR_DEVICE, "hipGraphLaunch failed"); break; }
            if (hipMemcpyAsync(e->h_ctl, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
                hipStreamSynchronize(e->stream) != hipSuccess) { rc = mg_fail(MG_ERR_DEVICE, "graph replay failed"); break; }
            if (e->h_ctl->mode == MODE_DONE) break;
        }
        if (rc == MG_OK) {
            e->done = true;
            e->iter = e->h_ctl->iter;
            e->levels = e->h_ctl->levels;
            e->edges_traversed = e->h_ctl->edges;
        }
    } else {
        int it = 1;
        for (;;) {
            int m = mg_step(e, &it);
            if (m < 0) { rc = m; break; }
            if (m == 0) break;
        }
    }
    if (rc == MG_OK && e->algo == MG_BC) rc = mg_bc_backward_impl(e);
    hipEventRecord(b, e->stream);
    hipError_t se = hipEventSynchronize(b);
    float t = 0.f;
    hipEventElapsedTime(&t, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc != MG_OK) return rc;
    if (se != hipSuccess) return mg_fail(MG_ERR_DEVICE, "run failed: %s", hipGetErrorString(se));
    if (e->prof) MG_TRY(mg_prof_flush(e));
    if (ms) *ms = t;
    return MG_OK;
}

extern "C" int mg_sync(mg_engine *e) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    MG_HIP(hipStreamSynchronize(e->stream));
    return MG_OK;
}

extern "C" int mg_iterations(const mg_engine *e, int *iters) {
    if (!e || !iters) return mg_fail(MG_ERR_ARG, "null argument");
    *iters = e->levels;
    return MG_OK;
}

extern "C" int mg_get_result(mg_engine *e, int which, void *dst, int64_t bytes) {
    if (!e || !dst) return mg_fail(MG_ERR_ARG, "null argument");
    const void *src = nullptr;
    int64_t need = 0;
    switch (which) {
    case MG_RES_LEVEL:
        if (e->algo == MG_PR_T || e->algo == MG_PR_D) return mg_fail(MG_ERR_ARG, "PageRank has no level result");
        src = e->level; need = (int64_t)sizeof(int) * e->nv; break;
    case MG_RES_SIGMA:
        if (e->algo != MG_BC) return mg_fail(MG_ERR_ARG, "sigma needs a BC engine");
        src = e->sigma; need = (int64_t)sizeof(double) * e->nv; break;
    case MG_RES_DELTA:
        if (e->algo != MG_BC) return mg_fail(MG_ERR_ARG, "delta needs a BC engine");
        src = e->delta; need = (int64_t)sizeof(float) * e->nv; break;
    case MG_RES_RANK:
        if (e->algo != MG_PR_T && e->algo != MG_PR_D) return mg_fail(MG_ERR_ARG, "rank needs a PageRank engine");
        src = e->rank; need = (int64_t)sizeof(float) * e->nv; break;
    default: return mg_fail(MG_ERR_ARG, "unknown result %d", which);
    }
    if (bytes < need) return mg_fail(MG_ERR_ARG, "buffer of %lld bytes, need %lld", (long long)bytes, (long long)need);
    MG_HIP(hipStreamSynchronize(e->stream));
    if (which == MG_RES_RANK && e->pr_pb) {  // internal numbering -> compacted ids
        std::vector<float> tmp(e->pr_nvi);
        MG_HIP(hipMemcpy(tmp.data(), e->rank, sizeof(float) * e->pr_nvi, hipMemcpyDeviceToHost));
        float *out = (float *)dst;
        for (int v = 0; v < e->nv; ++v) out[v] = tmp[e->pr_perm_h[v]];
        return MG_OK;
    }
    if (need > 0) MG_HIP(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
    return MG_OK;
}

// --------------------------------------------------------------------------- profiling
int mg_prof_begin(mg_engine *e, const char *name, double bytes, int *slot) {
    *slot = -1;
    if (!e->prof_only.empty() && strncmp(name, e->prof_only.c_str(), e->prof_only.size()) != 0) return MG_OK;
    int idx = -1;
    for (size_t i = 0; i < e->stats.size(); ++i) if (e->stats[i].name == name) { idx = (int)i; break; }
    if (idx < 0) { e->stats.push_back(KStat{name, 0, 0.0, 0.0}); idx = (int)e->stats.size() - 1; }
    hipEvent_t a, b;
    if (e->ev_pool.size() >= 2) {
        a = e->ev_pool.back(); e->ev_pool.pop_back();
        b = e->ev_pool.back(); e->ev_pool.pop_back();
    } else {
        MG_HIP(hipEventCreate(&a));
        MG_HIP(hipEventCreate(&b));
    }
    MG_HIP(hipEventRecord(a, e->stream));
    e->pending.push_back(mg_engine::Pending{idx, a, b, bytes});
    *slot = (int)e->pending.size() - 1;
    return MG_OK;
}

int mg_prof_end(mg_engine *e, int slot) {
    if (slot < 0 || slot >= (int)e->pending.size()) return MG_OK;
    MG_HIP(hipEventRecord(e->pending[slot].b, e->stream));
    return MG_OK;
}

int mg_prof_flush(mg_engine *e) {
    MG_HIP(hipStreamSynchronize(e->stream));
    for (auto &p : e->pending) {
        float t = 0.f;
        hipEventElapsedTime(&t, p.a, p.b);
        KStat &k = e->stats[p.stat];
        k.launches += 1;
        k.ms += t;
        k.bytes += p.bytes;
        e->ev_pool.push_back(p.a);
        e->ev_pool.push_back(p.b);
    }
    e->pending.clear();
    return MG_OK;
}

extern "C" int mg_set_profiling(mg_engine *e, int on) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    MG_TRY(mg_prof_flush(e));
    e->prof = on != 0;
    e->stats.clear();
    return MG_OK;
}

extern "C" int mg_set_profile_filter(mg_engine *e, const char *prefix) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    MG_TRY(mg_prof_flush(e));
    e->prof_only = prefix ? prefix : "";
    e->stats.clear();
    return MG_OK;
}

extern "C" int mg_kernel_stats(mg_engine *e, int idx, char *name, int name_len, int64_t *launches, double *ms, double *bytes) {
    if (!e) return mg_fail(MG_ERR_ARG, "null engine");
    MG_TRY(mg_prof_flush(e));
    if (idx < 0 || idx >= (int)e->stats.size()) return mg_fail(MG_ERR_ARG, "no kernel statistic %d", idx);
    const KStat &k = e->stats[idx];
    if (name && name_len > 0) { strncpy(name, k.name.c_str(), name_len - 1); name[name_len - 1] = 0; }
    if (launches) *launches = k.launches;
    if (ms) *ms = k.ms;
    if (bytes) *bytes = k.bytes;
    return MG_OK;
}

// batched runs are implemented in batch.hip
