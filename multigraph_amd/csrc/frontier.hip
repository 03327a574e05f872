// Frontier engine for BFS, SSSP and the BC forward pass (the role of process.cu) and
// the BC backward pass (the role of process2.cu), hand-written for gfx950.
//
// Per level the engine runs
//   * a push level (sparse queue): expand + contract, one workgroup per 256 frontier
//     vertices with a block-wide exclusive scan of their degrees so every lane takes
//     edges of the block in turn (edge-balanced regardless of degree); vertices at or
//     above `huge` edges are cut into 4096-edge chunks that the whole grid shares;
//     winners are appended with one atomic per wave (ballot + mbcnt);
//   * or a pull level (dense bitmap): one wave per 64 vertices, the frontier bitmap
//     word of a vertex's neighbour is tested per edge, and the wave's 64 results are
//     the next frontier word itself (ballot), so no atomics build the bitmap;
//   * then a one-thread control kernel that chooses push or pull for the next level
//     from frontier degree sums (deterministic, replacing the reference's wall-clock
//     sampling, process.cu:1161-1255).
// All kernels are grid-stride and read their sizes from the device control block,
// so a run of levels can be captured in a hipGraph and replayed with no host sync.
#include "engine.hpp"
#include "user_function.hpp"

using namespace mg;

namespace {

__global__ void k_fill_i32(int *p, long long n, int v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double *p, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32(float *p, long long n, float v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

// seed a run: state[src] = 0, queue = {src}, control block at level 1
__global__ void k_seed(Ctl *ctl, int *state, int *q, double *sigma, int src, int deg_src, long long ne, int out_begin) {
    state[src] = 0;
    if (sigma) sigma[src] = 1.0;
    q[0] = src;
    Ctl c;
    c.iter = 1; c.mode = MODE_TD; c.in_begin = 0; c.in_n = 1; c.out_begin = out_begin; c.out_n = 0;
    c.nchunks = 0; c.convert = 0; c.bm_cur = 0; c.levels = 1; c.changed = 0; c.pad0 = 0;
    c.out_mf = 0; c.mf = deg_src; c.mu = ne - deg_src; c.resid = 0.0; c.edges = 0;
    *ctl = c;
}

__global__ void k_ctl(Ctl *ctl, Params p, int *lvl_off) {
    Ctl c = *ctl;
    if (c.mode == MODE_DONE) return;
    if (lvl_off) {
        lvl_off[c.iter] = c.out_begin;
        lvl_off[c.iter + 1] = c.out_begin + c.out_n;
    }
    c.edges += c.mf;
    advance(c, p);
    *ctl = c;
}

// relax one edge src->u (edge index p); returns 1 when u joins the next frontier
template <class H>
struct Relax;

template <>
struct Relax<BFSHooks> {
    __device__ static int go(int sval, double, int u, int, const int *, int *state, int *, double *, int iter) {
        int msg;
        BFSHooks::get_partial_result(&msg, sval);
        return BFSHooks::fused_update_condition(&state[u], &state[u], msg, iter, nullptr);
    }
};
template <>
struct Relax<SSSPHooks> {
    __device__ static int go(int sval, double, int u, int p, const int *wgt, int *state, int *vfc, double *, int iter) {
        int msg;
        SSSPHooks::get_partial_result(&msg, sval, wgt[p]);
        return SSSPHooks::fused_update_condition(&state[u], &state[u], msg, iter, &vfc[u]);
    }
};
template <>
struct Relax<BCHooks> {
    __device__ static int go(int sval, double s2, int u, int, const int *, int *state, int *, double *sigma, int iter) {
        int msg;
        double msg2;
        BCHooks::get_partial_result(&msg, sval, &msg2, s2);
        return BCHooks::fused_update_condition(&state[u], &state[u], msg, iter, nullptr, &sigma[u], &sigma[u], msg2);
    }
};

template <bool kTrackDeg>
__device__ __forceinline__ void append(Ctl *ctl, int *q, int out_begin, const int *row, bool win, int u) {
    int pos = wave_append(&ctl->out_n, win);
    if (win) q[out_begin + pos] = u;
    if (kTrackDeg) {
        long long d = win ? (long long)(row[u + 1] - row[u]) : 0;
        if (ballot(win)) {
            d = wave_sum(d);
            if (lane() == 0) atomicAdd((unsigned long long *)&ctl->out_mf, (unsigned long long)d);
        }
    }
}

// Push level: expand the frontier's edges, contract winners into the next queue.
template <class H, bool kTrackDeg>
__global__ __launch_bounds__(kBlock) void k_push(Ctl *ctl, const Params P, const int *__restrict__ row,
                                                 const int *__restrict__ col, const int *__restrict__ wgt,
                                                 int *__restrict__ state, int *__restrict__ vfc,
                                                 double *__restrict__ sigma, int *__restrict__ q,
                                                 HugeChunk *__restrict__ chunks) {
    if (ctl->mode != MODE_TD) return;
    const int iter = ctl->iter, in_begin = ctl->in_begin, n = ctl->in_n, out_begin = ctl->out_begin;
    __shared__ int s_off[kBlock];
    __shared__ int s_rb[kBlock];
    __shared__ int s_val[kBlock];
    __shared__ double s_val2[kBlock];
    __shared__ int s_tmp[kBlock / MG_WAVE + 1];
    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const int i = base + threadIdx.x;
        int v = -1, rb = 0, deg = 0, val = 0;
        double val2 = 0.0;
        if (i < n) {
            v = q[in_begin + i];
            rb = row[v];
            deg = row[v + 1] - rb;
            val = state[v];
            if (sigma) val2 = sigma[v];
            if (deg >= P.huge) {
                int nch = (deg + kHugeChunk - 1) / kHugeChunk;
                int c0 = atomicAdd(&ctl->nchunks, nch);
                for (int j = 0; j < nch; ++j) chunks[c0 + j] = HugeChunk{v, j};
                deg = 0;
            }
        }
        int total;
        int off = block_excl_scan<kBlock / MG_WAVE>(deg, s_tmp, &total);
        s_off[threadIdx.x] = off;
        s_rb[threadIdx.x] = rb;
        s_val[threadIdx.x] = val;
        s_val2[threadIdx.x] = val2;
        __syncthreads();
        for (int e0 = 0; e0 < total; e0 += kBlock) {
            const int e = e0 + threadIdx.x;
            int win = 0, u = 0;
            if (e < total) {
                int lo = 0, hi = kBlock - 1;
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= e) lo = mid; else hi = mid - 1;
                }
                const int p = s_rb[lo] + (e - s_off[lo]);
                u = ld_stream(&col[p]);
                win = Relax<H>::go(s_val[lo], s_val2[lo], u, p, wgt, state, vfc, sigma, iter);
            }
            append<kTrackDeg>(ctl, q, out_begin, row, win != 0, u);
        }
        __syncthreads();
    }
}

// Push level, split high-degree vertices: grid-stride over 4096-edge chunks.
template <class H, bool kTrackDeg>
__global__ __launch_bounds__(kBlock) void k_push_huge(Ctl *ctl, const int *__restrict__ row, const int *__restrict__ col,
                                                      const int *__restrict__ wgt, int *__restrict__ state,
                                                      int *__restrict__ vfc, double *__restrict__ sigma,
                                                      int *__restrict__ q, const HugeChunk *__restrict__ chunks) {
    if (ctl->mode != MODE_TD) return;
    const int nch = ctl->nchunks;
    const int iter = ctl->iter, out_begin = ctl->out_begin;
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const HugeChunk hc = chunks[c];
        const int v = hc.v;
        const int vb = row[v], ve = row[v + 1];
        const int b = vb + hc.j * kHugeChunk;
        const int e = min(ve, b + kHugeChunk);
        const int val = state[v];
        const double val2 = sigma ? sigma[v] : 0.0;
        for (int p0 = b; p0 < e; p0 += kBlock) {
            const int p = p0 + threadIdx.x;
            int win = 0, u = 0;
            if (p < e) {
                u = ld_stream(&col[p]);
                win = Relax<H>::go(val, val2, u, p, wgt, state, vfc, sigma, iter);
            }
            append<kTrackDeg>(ctl, q, out_begin, row, win != 0, u);
        }
    }
}

// Build the visited and frontier bitmaps from the level array before a pull level.
__global__ __launch_bounds__(kBlock) void k_level2bm(Ctl *ctl, int nv, const int *__restrict__ level,
                                                     uint64_t *__restrict__ vis, uint64_t *__restrict__ bm0,
                                                     uint64_t *__restrict__ bm1) {
    if (ctl->mode != MODE_BU || !ctl->convert) return;
    const int fl = ctl->iter - 1;
    uint64_t *fr = ctl->bm_cur ? bm1 : bm0;
    const int words = (nv + 63) >> 6;
    const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * kBlock) >> 6;
    for (int w = wave; w < words; w += nwaves) {
        const int v = (w << 6) + lane();
        int lv = v < nv ? level[v] : 0;
        uint64_t mv = ballot(v >= nv || lv != MG_INF);
        uint64_t mf = ballot(v < nv && lv == fl);
        if (lane() == 0) { vis[w] = mv; fr[w] = mf; }
    }
}

__device__ __forceinline__ bool bm_test(const uint64_t *bm, int u) { return (bm[u >> 6] >> (u & 63)) & 1ull; }

// Pull level for BFS (early exit at the first parent) and for the BC forward pass
// (all parents, summing their path counts).
template <bool kBC>
__global__ __launch_bounds__(kBlock) void k_pull(Ctl *ctl, int nv, const int *__restrict__ row,
                                                 const int *__restrict__ col, int *__restrict__ level,
                                                 double *__restrict__ sigma, uint64_t *__restrict__ vis,
                                                 uint64_t *__restrict__ bm0, uint64_t *__restrict__ bm1,
                                                 int *__restrict__ q) {
    if (ctl->mode != MODE_BU) return;
    const int iter = ctl->iter, out_begin = ctl->out_begin;
    const uint64_t *fr = ctl->bm_cur ? bm1 : bm0;
    uint64_t *nx = ctl->bm_cur ? bm0 : bm1;
    const int words = (nv + 63) >> 6;
    const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * kBlock) >> 6;
    for (int w = wave; w < words; w += nwaves) {
        const uint64_t vw = vis[w];
        const int v = (w << 6) + lane();
        bool found = false;
        double s = 0.0;
        if (vw != ~0ull && !((vw >> lane()) & 1ull)) {
            const int b = row[v], e = row[v + 1];
            for (int p = b; p < e; ++p) {
                const int u = ld_stream(&col[p]);
                if (bm_test(fr, u)) {
                    found = true;
                    if (!kBC) break;
                    s += sigma[u];
                }
            }
        }
        const uint64_t m = ballot(found);
        if (lane() == 0) { nx[w] = m; if (m) vis[w] = vw | m; }
        if (found) {
            level[v] = iter;
            if (kBC) sigma[v] = s;
        }
        append<true>(ctl, q, out_begin, row, found, v);
    }
}

// -------------------------------------------------------------------- BC backward
// delta[v] = sum over neighbours w one level deeper of sigma[v]/sigma[w] * (1 + delta[w])
// (process2.cu:5-123, validate.cu:182-189), for the vertices of one level.  Degree
// tiers: a lane per vertex below 32 edges, a wave per vertex below 2048, and the
// whole grid (4096-edge chunks, FP64 partials) above.  Sums are Neumaier-compensated
// FP32 over terms split into two floats, so each term keeps ~48 bits.
__device__ __forceinline__ void bc_term(KSum &acc, double sv, double sw, float dw) {
    double t = (sv / sw) * (1.0 + (double)dw);
    float hi = (float)t;
    float lo = (float)(t - (double)hi);
    acc.add(hi);
    acc.add(lo);
}

__global__ __launch_bounds__(kBlock) void k_bc_back(int l, const int *__restrict__ lvl_off, const int *__restrict__ q,
                                                    const int *__restrict__ row, const int *__restrict__ col,
                                                    const int *__restrict__ level, const double *__restrict__ sigma,
                                                    float *__restrict__ delta, HugeChunk *__restrict__ chunks,
                                                    int *__restrict__ nchunks) {
    const int b = lvl_off[l], n = lvl_off[l + 1] - b;
    __shared__ int s_wv[kBlock];
    __shared__ int s_wn;
    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        if (threadIdx.x == 0) s_wn = 0;
        __syncthreads();
        const int i = base + threadIdx.x;
        if (i < n) {
            const int v = q[b + i];
            const int rb = row[v], re = row[v + 1], deg = re - rb;
            if (deg < 32) {
                const double sv = sigma[v];
                KSum acc;
                for (int p = rb; p < re; ++p) {
                    const int wv = col[p];
                    if (level[wv] == l + 1) bc_term(acc, sv, sigma[wv], delta[wv]);
                }
                delta[v] = acc.value();
            } else if (deg < 2048) {
                s_wv[atomicAdd(&s_wn, 1)] = v;
            } else {
                const int nch = (deg + kHugeChunk - 1) / kHugeChunk;
                const int c0 = atomicAdd(nchunks, nch);
                for (int j = 0; j < nch; ++j) chunks[c0 + j] = HugeChunk{v, j};
            }
        }
        __syncthreads();
        const int wn = s_wn;
        for (int k = threadIdx.x >> 6; k < wn; k += kBlock / MG_WAVE) {
            const int v = s_wv[k];
            const int rb = row[v], re = row[v + 1];
            const double sv = sigma[v];
            KSum acc;
            for (int p = rb + lane(); p < re; p += MG_WAVE) {
                const int wv = ld_stream(&col[p]);
                if (level[wv] == l + 1) bc_term(acc, sv, sigma[wv], delta[wv]);
            }
            acc = wave_ksum(acc);
            if (lane() == 0) delta[v] = acc.value();
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_bc_back_huge(int l, const int *__restrict__ row, const int *__restrict__ col,
                                                         const int *__restrict__ level, const double *__restrict__ sigma,
                                                         const float *__restrict__ delta, const HugeChunk *__restrict__ chunks,
                                                         const int *__restrict__ nchunks, double *__restrict__ dacc) {
    const int nch = *nchunks;
    __shared__ float s_s[kBlock / MG_WAVE], s_c[kBlock / MG_WAVE];
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const HugeChunk hc = chunks[c];
        const int v = hc.v;
        const int b = row[v] + hc.j * kHugeChunk, e = min(row[v + 1], b + kHugeChunk);
        const double sv = sigma[v];
        KSum acc;
        for (int p = b + threadIdx.x; p < e; p += kBlock) {
            const int wv = ld_stream(&col[p]);
            if (level[wv] == l + 1) bc_term(acc, sv, sigma[wv], delta[wv]);
        }
        acc = wave_ksum(acc);
        if (lane() == 0) { s_s[threadIdx.x >> 6] = acc.s; s_c[threadIdx.x >> 6] = acc.c; }
        __syncthreads();
        if (threadIdx.x == 0) {
            KSum t;
            for (int k = 0; k < kBlock / MG_WAVE; ++k) { t.add(s_s[k]); t.add(s_c[k]); }
            atomicAdd(&dacc[v], (double)t.s + (double)t.c);
        }
        __syncthreads();
    }
}

__global__ void k_bc_back_fin(const HugeChunk *__restrict__ chunks, int *__restrict__ nchunks, double *__restrict__ dacc,
                              float *__restrict__ delta) {
    const int nch = *nchunks;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nch; c += gridDim.x * blockDim.x) {
        if (chunks[c].j != 0) continue;  // the first chunk of each vertex finalizes it
        const int v = chunks[c].v;
        delta[v] = (float)dacc[v];
        dacc[v] = 0.0;
    }
}

__global__ void k_zero_i32(int *p) { *p = 0; }

int grid_for(long long items, int per_block, int cap) {
    long long g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

int max_grid(const mg_engine *e) { return e->opt.grid_blocks > 0 ? e->opt.grid_blocks : 2048; }

}  // namespace

// ---------------------------------------------------------------------------- host side

int mg_frontier_build(mg_engine *e, const mg_graph *g) {
    const int nv = e->nv;
    const long long ne = e->ne;
    const long long padv = (long long)nv + 64;
    MG_HIP(hipMalloc(&e->level, sizeof(int) * padv));
    e->qcap = (e->algo == MG_SSSP) ? 2 * (long long)nv + 2 : (long long)nv + 2;
    MG_HIP(hipMalloc(&e->q, sizeof(int) * e->qcap));
    e->chunk_cap = 2 * (ne / kHugeChunk) + nv / 1024 + 64;
    MG_HIP(hipMalloc(&e->chunks, sizeof(HugeChunk) * e->chunk_cap));
    MG_HIP(hipMalloc(&e->ctl, sizeof(Ctl)));
    MG_HIP(hipMalloc(&e->aux, sizeof(int) * 16));
    MG_HIP(hipHostMalloc(&e->h_ctl, sizeof(Ctl), hipHostMallocDefault));
    if (e->algo == MG_SSSP) {
        MG_HIP(hipMalloc(&e->vfc, sizeof(int) * padv));
    } else {
        const long long words = (nv + 63) / 64 + 1;
        MG_HIP(hipMalloc(&e->bm[0], sizeof(uint64_t) * words));
        MG_HIP(hipMalloc(&e->bm[1], sizeof(uint64_t) * words));
        MG_HIP(hipMalloc(&e->vis, sizeof(uint64_t) * words));
    }
    if (e->algo == MG_BC) {
        MG_HIP(hipMalloc(&e->sigma, sizeof(double) * padv));
        MG_HIP(hipMalloc(&e->delta, sizeof(float) * padv));
        MG_HIP(hipMalloc(&e->dacc, sizeof(double) * padv));
        MG_HIP(hipMalloc(&e->lvl_off, sizeof(int) * ((long long)nv + 4)));
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(padv, 256, 4096)), dim3(256), 0, e->stream, e->dacc, padv, 0.0);
    }
    Params &P = e->P;
    P.nv = nv;
    P.ne = ne;
    P.log_queue = (e->algo == MG_SSSP) ? 0 : 1;
    P.forced = (e->algo == MG_SSSP) ? MG_MODE_SPARSE : e->opt.mode;
    P.huge = e->opt.huge_degree;
    P.alpha = e->opt.alpha;
    P.beta = e->opt.beta;
    (void)g;
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int mg_frontier_reset(mg_engine *e, int source) {
    const int nv = e->nv;
    const long long padv = (long long)nv + 64;
    hipStream_t s = e->stream;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(padv, 256, 4096)), dim3(256), 0, s, e->level, padv, (int)MG_INF);
    if (e->vfc) hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(padv, 256, 4096)), dim3(256), 0, s, e->vfc, padv, -1);
    if (e->sigma) {
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(padv, 256, 4096)), dim3(256), 0, s, e->sigma, padv, 0.0);
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(padv, 256, 4096)), dim3(256), 0, s, e->delta, padv, 0.0f);
    }
    int deg_src = 0;
    if (nv > 0) {
        int rr[2];
        MG_HIP(hipMemcpy(rr, e->row + source, sizeof(int) * 2, hipMemcpyDeviceToHost));
        deg_src = rr[1] - rr[0];
    }
    const int out_begin = (e->P.log_queue) ? 1 : nv;
    if (nv > 0)
        hipLaunchKernelGGL(k_seed, dim3(1), dim3(1), 0, s, e->ctl, e->level, e->q, e->sigma, source, deg_src, e->ne, out_begin);
    MG_HIP(hipGetLastError());
    e->source = source;
    e->iter = 1;
    e->done = (nv == 0);
    e->levels = 1;
    if (nv > 0) {
        MG_HIP(hipMemcpyAsync(e->h_ctl, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        MG_HIP(hipStreamSynchronize(s));
    }
    return MG_OK;
}

// launch the kernels of one level; `known` = host knows the control block (eager mode)
static int launch_level(mg_engine *e, bool known) {
    const Ctl &c = *e->h_ctl;
    hipStream_t s = e->stream;
    const int G = max_grid(e);
    const int nv = e->nv;
    const bool td = !known || c.mode == MODE_TD;
    const bool bu = !known || c.mode == MODE_BU;
    int slot = -1;
    if (bu && (!known || c.convert)) {
        const int gb = grid_for((nv + 63) / 64, kBlock / MG_WAVE, G);
        if (e->prof) mg_prof_begin(e, "k_level2bm", 4.0 * nv + nv / 4.0, &slot);
        hipLaunchKernelGGL(k_level2bm, dim3(gb), dim3(kBlock), 0, s, e->ctl, nv, e->level, e->vis, e->bm[0], e->bm[1]);
        if (e->prof) mg_prof_end(e, slot);
    }
    if (td) {
        const int gt = known ? grid_for(c.in_n, kBlock, G) : G;
        // algorithmic bytes: queue + row pair per frontier vertex, column + state per edge
        const double bytes = known ? (12.0 * c.in_n + 8.0 * (double)c.mf + (e->algo == MG_SSSP ? 4.0 * c.mf : 0.0)) : 0.0;
        if (e->prof) mg_prof_begin(e, "k_push", bytes, &slot);
        switch (e->algo) {
        case MG_BFS:
            hipLaunchKernelGGL((k_push<BFSHooks, true>), dim3(gt), dim3(kBlock), 0, s, e->ctl, e->P, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
            break;
        case MG_SSSP:
            hipLaunchKernelGGL((k_push<SSSPHooks, true>), dim3(gt), dim3(kBlock), 0, s, e->ctl, e->P, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
            break;
        default:
            hipLaunchKernelGGL((k_push<BCHooks, true>), dim3(gt), dim3(kBlock), 0, s, e->ctl, e->P, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
            break;
        }
        if (e->prof) mg_prof_end(e, slot);
        const int gh = known ? grid_for((long long)(c.mf / kHugeChunk) + 1, 1, G) : G;
        if (!known || c.mf >= e->P.huge) {
            if (e->prof) mg_prof_begin(e, "k_push_huge", 0.0, &slot);
            switch (e->algo) {
            case MG_BFS:
                hipLaunchKernelGGL((k_push_huge<BFSHooks, true>), dim3(gh), dim3(kBlock), 0, s, e->ctl, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
                break;
            case MG_SSSP:
                hipLaunchKernelGGL((k_push_huge<SSSPHooks, true>), dim3(gh), dim3(kBlock), 0, s, e->ctl, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
                break;
            default:
                hipLaunchKernelGGL((k_push_huge<BCHooks, true>), dim3(gh), dim3(kBlock), 0, s, e->ctl, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
                break;
            }
            if (e->prof) mg_prof_end(e, slot);
        }
    }
    if (bu) {
        const int gb = grid_for((nv + 63) / 64, kBlock / MG_WAVE, G);
        const double bytes = known ? (8.0 * nv / 64.0 * 3 + 4.0 * (double)c.mu + 8.0 * nv) : 0.0;
        if (e->prof) mg_prof_begin(e, "k_pull", bytes, &slot);
        if (e->algo == MG_BC)
            hipLaunchKernelGGL((k_pull<true>), dim3(gb), dim3(kBlock), 0, s, e->ctl, nv, e->row, e->col, e->level, e->sigma, e->vis, e->bm[0], e->bm[1], e->q);
        else
            hipLaunchKernelGGL((k_pull<false>), dim3(gb), dim3(kBlock), 0, s, e->ctl, nv, e->row, e->col, e->level, e->sigma, e->vis, e->bm[0], e->bm[1], e->q);
        if (e->prof) mg_prof_end(e, slot);
    }
    hipLaunchKernelGGL(k_ctl, dim3(1), dim3(1), 0, s, e->ctl, e->P, e->lvl_off);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int mg_frontier_level(mg_engine *e) {
    if (e->done) return MG_OK;
    MG_TRY(launch_level(e, true));
    MG_HIP(hipMemcpyAsync(e->h_ctl, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream));
    MG_HIP(hipStreamSynchronize(e->stream));
    e->iter = e->h_ctl->iter;
    e->levels = e->h_ctl->levels;
    e->edges_traversed = e->h_ctl->edges;
    if (e->h_ctl->mode == MODE_DONE) e->done = true;
    return MG_OK;
}

int mg_frontier_capture(mg_engine *e, int levels, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    MG_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = MG_OK;
    for (int i = 0; i < levels && rc == MG_OK; ++i) rc = launch_level(e, false);
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc != MG_OK) { if (g) hipGraphDestroy(g); return rc; }
    if (ce != hipSuccess) return mg_fail(MG_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) return mg_fail(MG_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    return MG_OK;
}

int mg_bc_backward_impl(mg_engine *e) {
    const int L = e->levels;  // levels 0 .. L-1 hold vertices
    hipStream_t s = e->stream;
    const int G = max_grid(e);
    std::vector<int> off(L + 2, 0);
    // level offsets live on the device (written by the control kernel); level 0 is {source}
    MG_HIP(hipMemcpyAsync(off.data(), e->lvl_off, sizeof(int) * (L + 1), hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    off[0] = 0;
    off[1] = 1;
    MG_HIP(hipMemcpyAsync(e->lvl_off, off.data(), sizeof(int) * 2, hipMemcpyHostToDevice, s));
    int *nch = e->aux;  // split-vertex chunk counter
    for (int l = L - 1; l >= 0; --l) {
        const int n = off[l + 1] - off[l];
        if (n <= 0) continue;
        int slot = -1;
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, nch);
        if (e->prof) mg_prof_begin(e, "k_bc_back", 0.0, &slot);
        hipLaunchKernelGGL(k_bc_back, dim3(grid_for(n, kBlock, G)), dim3(kBlock), 0, s, l, e->lvl_off, e->q, e->row, e->col,
                           e->level, e->sigma, e->delta, e->chunks, nch);
        if (e->prof) mg_prof_end(e, slot);
        hipLaunchKernelGGL(k_bc_back_huge, dim3(G), dim3(kBlock), 0, s, l, e->row, e->col, e->level, e->sigma, e->delta,
                           e->chunks, nch, e->dacc);
        hipLaunchKernelGGL(k_bc_back_fin, dim3(64), dim3(256), 0, s, e->chunks, nch, e->dacc, e->delta);
    }
    MG_HIP(hipGetLastError());
    return MG_OK;
}
