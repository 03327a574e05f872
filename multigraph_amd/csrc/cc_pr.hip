// Connected components (the role of MultiGraph_E_E + MultiGraph_V_V, process.cu:940-1077)
// and PageRank (the role of MultiGraph_V_E_V_NOTIDEM, processi.cu) for gfx950.
#include "engine.hpp"
#include "hook_calls.hpp"
#include "lb.hpp"

using namespace mg;

namespace {

int grid_for(long long items, int per_block, int cap) {
    long long g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ------------------------------------------------------------------ scans
// Exclusive scan of n int counts into long long offsets (out has n+1 entries).
constexpr int kScanBlock = 256, kScanItems = 4;

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(const int *__restrict__ in, long long n, long long *__restrict__ out,
                                                            long long *__restrict__ bsum) {
    __shared__ long long tmp[kScanBlock / MG_WAVE + 1];
    const long long base = (long long)blockIdx.x * kScanBlock * kScanItems + (long long)threadIdx.x * kScanItems;
    long long v[kScanItems];
    long long local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? (long long)in[base + k] : 0;
        local += v[k];
    }
    long long total;
    long long pre = block_excl_scan<kScanBlock / MG_WAVE>(local, tmp, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks64(const long long *__restrict__ in, long long n,
                                                              long long *__restrict__ out, long long *__restrict__ bsum) {
    __shared__ long long tmp[kScanBlock / MG_WAVE + 1];
    const long long base = (long long)blockIdx.x * kScanBlock * kScanItems + (long long)threadIdx.x * kScanItems;
    long long v[kScanItems];
    long long local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        local += v[k];
    }
    long long total;
    long long pre = block_excl_scan<kScanBlock / MG_WAVE>(local, tmp, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void k_scan_add(long long *__restrict__ out, long long n, const long long *__restrict__ boff) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += boff[i / (kScanBlock * kScanItems)];
}

__global__ void k_set_last(long long *__restrict__ out, long long n, const long long *__restrict__ bsum_scanned,
                           long long nb, const long long *__restrict__ bsum) {
    out[n] = bsum_scanned[nb - 1] + bsum[nb - 1];
}

}  // namespace

// out[0..n] = exclusive scan of in[0..n) (out[n] = total); on the engine stream.
int mg_scan64(hipStream_t s, const long long *in64, const int *in32, long long n, long long *out) {
    const long long per = (long long)kScanBlock * kScanItems;
    const long long nb = (n + per - 1) / per;
    if (n == 0) {
        long long z = 0;
        MG_HIP(hipMemcpyAsync(out, &z, sizeof(long long), hipMemcpyHostToDevice, s));
        MG_HIP(hipStreamSynchronize(s));
        return MG_OK;
    }
    long long *bsum = nullptr, *boff = nullptr;
    MG_HIP(hipMalloc(&bsum, sizeof(long long) * nb));
    MG_HIP(hipMalloc(&boff, sizeof(long long) * (nb + 1)));
    if (in32) hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, s, in32, n, out, bsum);
    else hipLaunchKernelGGL(k_scan_blocks64, dim3(nb), dim3(kScanBlock), 0, s, in64, n, out, bsum);
    int rc = MG_OK;
    if (nb > 1) rc = mg_scan64(s, bsum, nullptr, nb, boff);
    else { long long z = 0; hipMemcpyAsync(boff, &z, sizeof(long long), hipMemcpyHostToDevice, s); }
    if (rc == MG_OK) {
        if (nb > 1) hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, s, out, n, boff);
        hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, out, n, boff, nb, bsum);
    }
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(s);
    hipFree(bsum);
    hipFree(boff);
    if (rc != MG_OK) return rc;
    if (le != hipSuccess || se != hipSuccess) return mg_fail(MG_ERR_DEVICE, "scan: %s", hipGetErrorString(le != hipSuccess ? le : se));
    return MG_OK;
}

namespace {

// ------------------------------------------------------------------ CC
// Upper-triangular CSR (edges u < v), the role of the reference's src<dst COO
// (read_input.cu:111-113, gen_structure.cu:34-139).
// the hooked edge list: (u, v) with v > u of a symmetric CSR (each undirected edge once);
// every non-loop edge of a directed one (`all`), whose components are the weak ones
__global__ void k_cc_count(int nv, const int *__restrict__ row, const int *__restrict__ col, int all, int *__restrict__ cnt) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int u = wave; u < nv; u += nw) {
        int c = 0;
        for (int p = row[u] + lane(); p < row[u + 1]; p += MG_WAVE) c += all ? col[p] != u : col[p] > u;
        c = wave_sum(c);
        if (lane() == 0) cnt[u] = c;
    }
}

__global__ void k_cc_fill(int nv, const int *__restrict__ row, const int *__restrict__ col, int all, const long long *__restrict__ uoff,
                          int *__restrict__ urow, int *__restrict__ ucol) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int u = wave; u < nv; u += nw) {
        long long o = uoff[u];
        if (lane() == 0) urow[u] = (int)o;
        for (int p0 = row[u]; p0 < row[u + 1]; p0 += MG_WAVE) {
            const int p = p0 + lane();
            const bool keep = p < row[u + 1] && (all ? col[p] != u : col[p] > u);
            const uint64_t m = ballot(keep);
            if (keep) ucol[o + mbcnt(m)] = col[p];
            o += __builtin_popcountll(m);
        }
        if (u == nv - 1 && lane() == 0) urow[nv] = (int)uoff[nv];
    }
}

__global__ void k_cc_init(int nv, int *__restrict__ parent) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) parent[v] = v;
}

// Edge-balanced hooking: workgroup c takes edges [c*kEPB, (c+1)*kEPB) of the
// upper-triangular CSR; owners come from chunk_owners (lb.hpp).
constexpr int kGiantSamples = 512;  // of k_cc_sample's 1024: the giant tree is "most" of the graph
// gate (optional, k_cc_sample's output): nothing to do when the sampled giant tree
// held at least half the samples (k_cc_link_rest hooked the edges instead)
__global__ __launch_bounds__(kBlock) void k_cc_hook_lb(int nch, long long ne, const int *__restrict__ off,
                                                       const int *__restrict__ adj, const int *__restrict__ cstart,
                                                       int *__restrict__ parent, const int *__restrict__ gate) {
    if (gate && gate[1] >= kGiantSamples) return;
    __shared__ int s_own[kEPB];
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const int e0 = c * kEPB;
        const int e1 = (int)min((long long)e0 + kEPB, ne);
        chunk_owners(off, cstart[c], cstart[c + 1], e0, e1, s_own);
        for (int e = e0 + threadIdx.x; e < e1; e += kBlock) MG_CC_HOOKS::fused_update_condition(parent, s_own[e - e0], ld_stream(&adj[e]));
        __syncthreads();
    }
}

// Sampling pass without atomics: rows are sorted, so col[row[v]] is v's smallest
// neighbour; parent[v] = min(v, that neighbour) joins v with a vertex of its own
// component, keeps parent[v] <= v, and is written only by v's thread.
__global__ void k_cc_link_min(int nv, const int *__restrict__ row, const int *__restrict__ col, int *__restrict__ parent) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        const int b = row[v];
        const int m = (row[v + 1] > b) ? col[b] : v;
        parent[v] = m < v ? m : v;
    }
}

// hook the second neighbour of every vertex (the first was linked by k_cc_link_min)
__global__ void k_cc_link_second(int nv, const int *__restrict__ row, const int *__restrict__ col, int *__restrict__ parent) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        const int b = row[v] + 1;
        if (b < row[v + 1]) MG_CC_HOOKS::fused_update_condition(parent, v, col[b]);
    }
}

// the most frequent root among 1024 evenly spaced vertices (bitonic sort in LDS)
__global__ __launch_bounds__(1024) void k_cc_sample(int nv, int *__restrict__ parent, int *__restrict__ out) {
    __shared__ int s[1024];
    const int t = threadIdx.x;
    const long long v = ((long long)t * 2654435761ll) % nv;
    s[t] = MG_CC_HOOKS::find_ro(parent, (int)v);
    __syncthreads();
    for (int k = 2; k <= 1024; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            const int i = t ^ j;
            if (i > t) {
                const bool up = (t & k) == 0;
                const int a = s[t], b = s[i];
                if ((a > b) == up) { s[t] = b; s[i] = a; }
            }
            __syncthreads();
        }
    // the thread at the start of each run of equal roots finds the run's end by binary
    // search; the block keeps the longest run, the smallest root among equally long ones
    long long key = -1;
    if (t == 0 || s[t - 1] != s[t]) {
        const int a = s[t];
        int lo = t + 1, hi = 1024;  // first index past t holding a larger root
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s[mid] > a) hi = mid;
            else lo = mid + 1;
        }
        key = ((long long)(lo - t) << 32) | (long long)(unsigned)(0x7FFFFFFF - a);
    }
    __shared__ long long s_key[1024 / MG_WAVE];
    key = block_max<1024 / MG_WAVE>(key, s_key);
    if (t == 0) {
        out[0] = 0x7FFFFFFF - (int)(key & 0xFFFFFFFFll);
        out[1] = (int)(key >> 32);
    }
}

// hook the remaining neighbours of every vertex outside the sampled giant tree, when
// that tree held at least half the samples (giant = {root, samples}); otherwise
// k_cc_hook_lb hooks every edge
__global__ __launch_bounds__(kBlock) void k_cc_link_rest(int nv, int k, const int *__restrict__ row,
                                                         const int *__restrict__ col, const int *__restrict__ giant,
                                                         int *__restrict__ parent) {
    if (giant[1] < kGiantSamples) return;
    __shared__ int s_wv[kBlock];
    __shared__ int s_wn;
    const int c = giant[0];
    for (int base = blockIdx.x * kBlock; base < nv; base += gridDim.x * kBlock) {
        if (threadIdx.x == 0) s_wn = 0;
        __syncthreads();
        const int v = base + threadIdx.x;
        if (v < nv && MG_CC_HOOKS::find(parent, v) != c) {
            const int r0 = row[v], b = r0 + k, e = row[v + 1];
            // k_cc_link_min joined v with its smallest neighbour only if that one is smaller
            if (e > r0 && col[r0] > v) MG_CC_HOOKS::fused_update_condition(parent, v, col[r0]);
            if (e - b >= 64) s_wv[atomicAdd(&s_wn, 1)] = v;
            else for (int p = b; p < e; ++p) MG_CC_HOOKS::fused_update_condition(parent, v, col[p]);
        }
        __syncthreads();
        const int wn = s_wn;
        for (int i = threadIdx.x >> 6; i < wn; i += kBlock / MG_WAVE) {
            const int u = s_wv[i];
            for (int p = row[u] + k + lane(); p < row[u + 1]; p += MG_WAVE) MG_CC_HOOKS::fused_update_condition(parent, u, col[p]);
        }
        __syncthreads();
    }
}

__global__ void k_cc_compress(int nv, int *__restrict__ parent) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x)
        MG_CC_HOOKS::update_condition(parent, v);
}

// ------------------------------------------------------------------ PageRank (pull)
// psum[v] = sum over neighbours u of c[u] (c = rank/deg), FP64, degree tiers as in
// the BC backward pass; then apply the PR_T / PR_D rule.
// CSR pull for PageRank: sums start at initialize's identity, every in-edge adds its
// source's message (accumulate_nonAtomics in registers, lanes merged the same way), a
// split vertex's chunks add theirs with accumulate; k_pr_apply applies update_condition
// and computes the next messages with get_partial_result.
template <class H>
__global__ __launch_bounds__(kBlock) void k_pr_gather(int nv, int huge, const int *__restrict__ row,
                                                      const int *__restrict__ col, const float *__restrict__ c,
                                                      double *__restrict__ psum) {
    __shared__ int s_wv[kBlock];
    __shared__ int s_wn;
    for (int base = blockIdx.x * kBlock; base < nv; base += gridDim.x * kBlock) {
        if (threadIdx.x == 0) s_wn = 0;
        __syncthreads();
        const int v = base + threadIdx.x;
        if (v < nv) {
            const int rb = row[v], re = row[v + 1], deg = re - rb;
            if (deg < 32) {
                double s;
                H::initialize(&s);
                for (int p = rb; p < re; ++p) H::accumulate_nonAtomics(&s, (double)c[ld_stream(&col[p])]);
                psum[v] = s;
            } else if (deg < huge) {
                s_wv[atomicAdd(&s_wn, 1)] = v;
            } else {
                H::initialize(&psum[v]);
            }
        }
        __syncthreads();
        const int wn = s_wn;
        for (int k = threadIdx.x >> 6; k < wn; k += kBlock / MG_WAVE) {
            const int u = s_wv[k];
            double s;
            H::initialize(&s);
            for (int p = row[u] + lane(); p < row[u + 1]; p += MG_WAVE) H::accumulate_nonAtomics(&s, (double)c[ld_stream(&col[p])]);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) H::accumulate_nonAtomics(&s, __shfl_xor(s, o));
            if (lane() == 0) psum[u] = s;
        }
        __syncthreads();
    }
}

template <class H>
__global__ __launch_bounds__(kBlock) void k_pr_gather_huge(const HugeChunk *__restrict__ chunks, int nch,
                                                           const int *__restrict__ row, const int *__restrict__ col,
                                                           const float *__restrict__ c, double *__restrict__ psum) {
    __shared__ double s_t[kBlock / MG_WAVE];
    for (int k = blockIdx.x; k < nch; k += gridDim.x) {
        const HugeChunk hc = chunks[k];
        const int b = row[hc.v] + hc.j * kHugeChunk, e = min(row[hc.v + 1], b + kHugeChunk);
        double s;
        H::initialize(&s);
        for (int p = b + threadIdx.x; p < e; p += kBlock) H::accumulate_nonAtomics(&s, (double)c[ld_stream(&col[p])]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) H::accumulate_nonAtomics(&s, __shfl_xor(s, o));
        if (lane() == 0) s_t[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t;
            H::initialize(&t);
            for (int i = 0; i < kBlock / MG_WAVE; ++i) H::accumulate_nonAtomics(&t, s_t[i]);
            H::accumulate(&psum[hc.v], t);
        }
        __syncthreads();
    }
}

template <class H>
__global__ __launch_bounds__(kBlock) void k_pr_apply(int nv, int variant, float damping, float tol,
                                                     const int *__restrict__ deg, const double *__restrict__ psum,
                                                     float *__restrict__ rank, float *__restrict__ c,
                                                     int *__restrict__ active, Ctl *__restrict__ ctl) {
    __shared__ double s_tmp[kBlock / MG_WAVE];
    __shared__ int s_ch;
    if (threadIdx.x == 0) s_ch = 0;
    __syncthreads();
    double res = 0.0;
    int ch = 0;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < nv; v += gridDim.x * kBlock) {
        float f1;
        const float old = rank[v];
        ch |= H::update_condition(&rank[v], &active[v], psum[v], variant, damping, tol, &f1);
        res += fabs((double)f1 - (double)old);
        H::get_partial_result(&c[v], rank[v], deg[v]);
    }
    res = block_sum<kBlock / MG_WAVE>(res, s_tmp);
    if (ch) s_ch = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&ctl->resid, res);
        if (s_ch) atomicOr(&ctl->changed, 1);
    }
}

__global__ void k_pr_init(int nv, float *__restrict__ rank, float *__restrict__ c, int *__restrict__ active,
                          Ctl *__restrict__ ctl) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        rank[v] = 0.f; c[v] = 0.f; active[v] = 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Ctl k{};
        k.iter = 1; k.mode = MODE_TD; k.changed = 0; k.resid = 0.0;
        *ctl = k;
    }
}

// end of one iteration: stop when nothing changed or at the cap; clear the flags
__global__ void k_pr_ctl(Ctl *ctl, int max_iter, long long ne) {
    Ctl c = *ctl;
    if (c.mode == MODE_DONE) return;
    c.edges += ne;
    c.levels = c.iter;
    if (!c.changed || (max_iter > 0 && c.iter >= max_iter)) c.mode = MODE_DONE;
    else { c.iter += 1; c.changed = 0; c.mf = 0; c.resid = 0.0; }
    *ctl = c;
}

__global__ void k_pr_begin(Ctl *ctl) {
    if (ctl->mode == MODE_DONE) return;
    ctl->resid = 0.0;
    ctl->changed = 0;
}

__global__ void k_degrees(int nv, const int *__restrict__ row, int *__restrict__ deg) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) deg[v] = row[v + 1] - row[v];
}

// the PageRank kernels exit when the run is over, so a captured graph of several
// iterations can be replayed past convergence
__global__ void k_pr_gate(Ctl *ctl, int *go) { *go = ctl->mode != MODE_DONE; }

}  // namespace

static int build_huge_chunks(mg_engine *e, const std::vector<int32_t> &row, const int nv, HugeChunk **out, int *nch) {
    std::vector<HugeChunk> v;
    const int huge = e->opt.huge_degree;
    for (int u = 0; u < nv; ++u) {
        int d = row[u + 1] - row[u];
        if (d >= huge)
            for (int j = 0; j * kHugeChunk < d; ++j) v.push_back(HugeChunk{u, j});
    }
    *nch = (int)v.size();
    *out = nullptr;
    if (!v.empty()) {
        MG_HIP(hipMalloc(out, sizeof(HugeChunk) * v.size()));
        MG_HIP(hipMemcpy(*out, v.data(), sizeof(HugeChunk) * v.size(), hipMemcpyHostToDevice));
    }
    return MG_OK;
}

// ---------------------------------------------------------------------------- CC host
int mg_cc_build(mg_engine *e, const mg_graph *g) {
    const int nv = e->nv;
    hipStream_t s = e->stream;
    int *cnt = nullptr;
    long long *uoff = nullptr;
    MG_HIP(hipMalloc(&cnt, sizeof(int) * ((long long)nv + 1)));
    MG_HIP(hipMalloc(&uoff, sizeof(long long) * ((long long)nv + 1)));
    hipLaunchKernelGGL(k_cc_count, dim3(grid_for(nv, 4, 8192)), dim3(256), 0, s, nv, e->row, e->col, g->symmetric ? 0 : 1, cnt);
    MG_TRY(mg_scan64(s, nullptr, cnt, nv, uoff));
    long long ne_u = 0;
    MG_HIP(hipMemcpy(&ne_u, uoff + nv, sizeof(long long), hipMemcpyDeviceToHost));
    e->ne_u = ne_u;
    MG_HIP(hipMalloc(&e->urow, sizeof(int) * ((long long)nv + 1)));
    MG_HIP(hipMalloc(&e->ucol, sizeof(int) * (ne_u > 0 ? ne_u : 1)));
    hipLaunchKernelGGL(k_cc_fill, dim3(grid_for(nv, 4, 8192)), dim3(256), 0, s, nv, e->row, e->col, g->symmetric ? 0 : 1, uoff, e->urow, e->ucol);
    if (nv == 0) {
        int z = 0;
        MG_HIP(hipMemcpy(e->urow, &z, sizeof(int), hipMemcpyHostToDevice));
    }
    // chunk owners of the upper-triangular edge list (static: computed once)
    const int nch = (int)((ne_u + kEPB - 1) / kEPB);
    e->chunk_cap = nch;
    MG_HIP(hipMalloc(&e->cstart, sizeof(int) * (nch + 2)));
    if (nch > 0) hipLaunchKernelGGL(k_chunk_part<int>, dim3(grid_for(nch + 1, 256, 4096)), dim3(256), 0, s, nv, ne_u, e->urow, nch, e->cstart);
    MG_HIP(hipStreamSynchronize(s));
    MG_HIP(hipGetLastError());
    hipFree(cnt);
    hipFree(uoff);
    MG_HIP(hipMalloc(&e->level, sizeof(int) * ((long long)nv + 64)));
    MG_HIP(hipMalloc(&e->aux, sizeof(int) * 16));
    MG_HIP(hipMalloc(&e->ctl, sizeof(Ctl)));
    MG_HIP(hipHostMalloc(&e->h_ctl, sizeof(Ctl), hipHostMallocDefault));
    e->cc_afforest = g->symmetric && nv >= 1024;
    return MG_OK;
}

int mg_cc_reset(mg_engine *e) {
    hipLaunchKernelGGL(k_cc_init, dim3(grid_for(e->nv, 256, 4096)), dim3(256), 0, e->stream, e->nv, e->level);
    MG_HIP(hipGetLastError());
    e->iter = 1;
    e->done = e->nv == 0;
    e->levels = 0;
    return MG_OK;
}

// One call does all the work (hook + compress), so MultiGraph_E_E-style loops end
// after their first call.  With a symmetric CSR the sampling pass runs first: link
// every vertex to its smallest neighbour (no atomics), compress, hook its second
// neighbour, compress, take the most frequent root of 1024 samples as the
// giant tree, and hook the remaining edges only of vertices outside it; every edge
// with an endpoint outside the giant tree is hooked from that endpoint's row, and an
// edge with both endpoints inside joins vertices already in one tree.
int mg_cc_step(mg_engine *e, int *more) {
    hipStream_t s = e->stream;
    const int G = e->opt.grid_blocks > 0 ? e->opt.grid_blocks : 2048;
    const int nv = e->nv;
    int slot = -1, rest_slot = -1;
    bool hooked = false;
    if (e->cc_afforest) {  // no host round trip: the last two hooks read the sample on the device
        const int k = 2;
        if (e->prof) mg_prof_begin(e, "k_cc_link_min", 16.0 * nv, &slot);
        hipLaunchKernelGGL(k_cc_link_min, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, s, nv, e->row, e->col, e->level);
        if (e->prof) mg_prof_end(e, slot);
        hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, s, nv, e->level);
        if (e->prof) mg_prof_begin(e, "k_cc_link_second", 16.0 * nv, &slot);
        hipLaunchKernelGGL(k_cc_link_second, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, s, nv, e->row, e->col, e->level);
        if (e->prof) mg_prof_end(e, slot);
        hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, s, nv, e->level);
        hipLaunchKernelGGL(k_cc_sample, dim3(1), dim3(1024), 0, s, nv, e->level, e->aux);
        if (e->prof) mg_prof_begin(e, "k_cc_link_rest", 8.0 * nv, &rest_slot);
        hipLaunchKernelGGL(k_cc_link_rest, dim3(grid_for(nv, kBlock, G)), dim3(kBlock), 0, s, nv, k, e->row, e->col, e->aux, e->level);
        if (e->prof) mg_prof_end(e, rest_slot);
        hooked = true;
    }
    int hook_slot = -1;
    if (e->chunk_cap > 0) {  // after the sampling pass: only when its giant tree was too small
        if (e->prof) mg_prof_begin(e, "k_cc_hook_lb", 4.0 * e->ne_u + 4.0 * (nv + 1) + 8.0 * e->ne_u, &hook_slot);
        hipLaunchKernelGGL(k_cc_hook_lb, dim3(grid_for(e->chunk_cap, 1, G)), dim3(kBlock), 0, s, (int)e->chunk_cap, e->ne_u,
                           e->urow, e->ucol, e->cstart, e->level, hooked ? (const int *)e->aux : (const int *)nullptr);
        if (e->prof) mg_prof_end(e, hook_slot);
    }
    if (e->prof) mg_prof_begin(e, "k_cc_compress", 8.0 * nv, &slot);
    hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, s, nv, e->level);
    if (e->prof) mg_prof_end(e, slot);
    MG_HIP(hipGetLastError());
    MG_HIP(hipStreamSynchronize(s));
    if (e->prof && hooked) {
        // of k_cc_link_rest and k_cc_hook_lb, the sample lets one work and the other
        // return at once: the one that returned moved no bytes
        int aux2[2] = {0, 0};
        MG_HIP(hipMemcpy(aux2, e->aux, sizeof(aux2), hipMemcpyDeviceToHost));
        const bool giant = aux2[1] >= kGiantSamples;
        if (giant && hook_slot >= 0) e->pending[hook_slot].bytes = 0.0;
        if (!giant && rest_slot >= 0) e->pending[rest_slot].bytes = 0.0;
    }
    e->iter += 1;
    e->levels = 1;
    e->edges_traversed = e->ne_u;
    e->done = true;
    *more = 0;
    return MG_OK;
}

// ---------------------------------------------------------------------------- PR host
int mg_pb_build_alloc(mg_engine *e, const mg_graph *g) {
    MG_TRY(mg_pb_build(e, g));  // sets pr_nvi and the degrees of the internal ids
    const long long padv = e->pr_nvi + 64;
    MG_HIP(hipMalloc(&e->rank, sizeof(float) * padv));
    MG_HIP(hipMalloc(&e->cval, sizeof(float) * padv));
    MG_HIP(hipMalloc(&e->vfc, sizeof(int) * padv));
    return MG_OK;
}
// PageRank: the partitioned layout (default, the reference's DENSE_MODE build) or a
// pull over the CSR (MG_MODE_SPARSE)
int mg_pr_build(mg_engine *e, const mg_graph *g) {
    const int nv = e->nv;
    MG_HIP(hipMalloc(&e->ctl, sizeof(Ctl)));
    MG_HIP(hipHostMalloc(&e->h_ctl, sizeof(Ctl), hipHostMallocDefault));
    // the pull kernels read a CSR row as in-edges, so a directed graph always takes the
    // partitioned layout, which scatters along out-edges and gathers by destination
    if ((e->opt.mode != MG_MODE_SPARSE || !g->symmetric) && nv > 0) {
        e->pr_pb = true;
        MG_TRY(mg_pb_build_alloc(e, g));  // internal ids are padded to whole partitions
        return MG_OK;
    }
    const long long padv = (long long)nv + 64;
    MG_HIP(hipMalloc(&e->rank, sizeof(float) * padv));
    MG_HIP(hipMalloc(&e->cval, sizeof(float) * padv));
    MG_HIP(hipMalloc(&e->deg, sizeof(int) * padv));
    MG_HIP(hipMalloc(&e->vfc, sizeof(int) * padv));
    MG_HIP(hipMalloc(&e->psum, sizeof(double) * padv));
    hipLaunchKernelGGL(k_degrees, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, e->stream, nv, e->row, e->deg);
    int nch = 0;
    MG_TRY(build_huge_chunks(e, g->row, nv, &e->chunks, &nch));
    e->chunk_cap = nch;
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int mg_pr_reset(mg_engine *e) {
    const long long n = e->pr_pb ? e->pr_nvi : e->nv;
    hipLaunchKernelGGL(k_pr_init, dim3(grid_for(n, 256, 4096)), dim3(256), 0, e->stream, (int)n, e->rank, e->cval, e->vfc, e->ctl);
    if (e->pr_pb) MG_TRY(mg_pb_reset(e));
    MG_HIP(hipGetLastError());
    MG_HIP(hipMemcpyAsync(e->h_ctl, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream));
    MG_HIP(hipStreamSynchronize(e->stream));
    e->iter = 1;
    e->done = e->nv == 0;
    e->levels = 0;
    return MG_OK;
}

static int pr_launch_iteration(mg_engine *e) {
    hipStream_t s = e->stream;
    if (e->pr_pb) {  // the control kernel clears the flags for the next iteration
        MG_TRY(mg_pb_iteration(e));
        hipLaunchKernelGGL(k_pr_ctl, dim3(1), dim3(1), 0, s, e->ctl, e->opt.pr_max_iter, e->ne);
        MG_HIP(hipGetLastError());
        return MG_OK;
    }
    const int G = e->opt.grid_blocks > 0 ? e->opt.grid_blocks : 2048;
    const int variant = e->algo == MG_PR_T ? 0 : 1;
    int slot = -1;
    hipLaunchKernelGGL(k_pr_begin, dim3(1), dim3(1), 0, s, e->ctl);
    // algorithmic bytes of one pull iteration: row + col + contribution gather per edge
    if (e->prof) mg_prof_begin(e, "k_pr_gather", 4.0 * (e->nv + 1) + 8.0 * e->ne + 8.0 * e->nv, &slot);
    hipLaunchKernelGGL(k_pr_gather<MG_PR_HOOKS>, dim3(grid_for(e->nv, kBlock, G)), dim3(kBlock), 0, s, e->nv, e->opt.huge_degree, e->row,
                       e->col, e->cval, e->psum);
    if (e->prof) mg_prof_end(e, slot);
    if (e->chunk_cap > 0)
        hipLaunchKernelGGL(k_pr_gather_huge<MG_PR_HOOKS>, dim3(grid_for(e->chunk_cap, 1, G)), dim3(kBlock), 0, s, e->chunks,
                           (int)e->chunk_cap, e->row, e->col, e->cval, e->psum);
    if (e->prof) mg_prof_begin(e, "k_pr_apply", 28.0 * e->nv, &slot);
    hipLaunchKernelGGL(k_pr_apply<MG_PR_HOOKS>, dim3(grid_for(e->nv, kBlock, 4096)), dim3(kBlock), 0, s, e->nv, variant, e->opt.pr_damping,
                       e->opt.pr_tolerance, e->deg, e->psum, e->rank, e->cval, e->vfc, e->ctl);
    if (e->prof) mg_prof_end(e, slot);
    hipLaunchKernelGGL(k_pr_ctl, dim3(1), dim3(1), 0, s, e->ctl, e->opt.pr_max_iter, e->ne);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int mg_pr_step(mg_engine *e, int *more) {
    if (e->done) { *more = 0; return MG_OK; }
    MG_TRY(pr_launch_iteration(e));
    MG_HIP(hipMemcpyAsync(e->h_ctl, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream));
    MG_HIP(hipStreamSynchronize(e->stream));
    e->iter = e->h_ctl->iter;
    e->levels = e->h_ctl->levels;
    e->edges_traversed = e->h_ctl->edges;
    e->done = e->h_ctl->mode == MODE_DONE;
    *more = e->done ? 0 : 1;
    return MG_OK;
}

// K iterations of the partitioned layout in one hipGraph: every kernel of an
// iteration returns at once when the control block says the run is over, so the graph
// can be replayed past convergence and the host checks the control block once per K
int mg_pr_capture(mg_engine *e, int iters, hipGraphExec_t *out) {
    if (!e->pr_pb) return mg_fail(MG_ERR_STATE, "only the partitioned PageRank layout runs from a captured graph");
    hipGraph_t g = nullptr;
    MG_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = MG_OK;
    for (int i = 0; i < iters && rc == MG_OK; ++i) rc = pr_launch_iteration(e);
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc != MG_OK) { if (g) hipGraphDestroy(g); return rc; }
    if (ce != hipSuccess) return mg_fail(MG_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) return mg_fail(MG_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    return MG_OK;
}
