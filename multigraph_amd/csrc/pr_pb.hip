// PageRank over the partitioned layout: the role of gen_structure.cu (layout build)
// and of _phase1 / _phase2 in processi.cu (one dense iteration), re-designed for gfx950.
//
// Layout.  Vertices are cut into np partitions of P ids.  An edge of row x with
// column y carries rank mass from y (source) to x (destination); it belongs to tile
// (s, d) = (y / P, x / P).  Every tile is stored twice, with the same slot order:
//   stream A (source-major: s, then d) holds the local source id (y mod P, 16 bit),
//   stream B (destination-major: d, then s) holds the local destination id (x mod P).
// Tiles are padded to a multiple of G = 16 slots (padding id 0xFFFF), and amap[g]
// gives, for the g-th group of 16 A-slots, the position of the matching 16 B-slots.
//
// One iteration = two streaming passes, no random HBM access and no global atomics:
//   scatter: per source partition, stage c = rank / degree of its P vertices in LDS,
//            stream A and write msg[B position] = c[local source] (16-slot runs);
//   gather:  per destination partition, accumulate msg into P FP64 LDS accumulators
//            (native LDS FP64 atomic add) while streaming B, then apply the PR_T /
//            PR_D rule to the partition's vertices and write rank and the next c.
// Work is cut into units of about equal slot counts; a destination partition whose
// stream is cut into several units adds its partial sums into an FP64 vector instead
// and is applied by a separate pass.  Traffic: 2 + 4 + 0.25 B per slot in the scatter,
// 2 + 4 B per slot in the gather, plus 20 B per vertex.
#include <algorithm>
#include <vector>
#include "engine.hpp"
#include "user_function.hpp"

using namespace mg;

namespace {

constexpr int kG = 16;          // slots per mapper group
constexpr int kBT = 512;        // threads per workgroup of the layout and phase kernels
constexpr int kBuildEdges = 1 << 16;

int grid_for(long long items, int per_block, int cap) {
    long long g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

__global__ void k_fill_u16(uint16_t *p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0xFFFF;
}

// build work item: rows [x0, x1) of destination partition d
struct BItem { int d, x0, x1, pad; };

// rank of an edge's row inside the item: largest k with rowoff[k] <= e
__device__ __forceinline__ int row_of(const int *s_row, int n, int e) {
    int lo = 0, hi = n - 1;
    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (s_row[mid] <= e) lo = mid; else hi = mid - 1; }
    return lo;
}

// pass 1: edges per tile (destination-major index d*np + s)
__global__ __launch_bounds__(kBT) void k_pb_count(const int *__restrict__ col, const int *__restrict__ row,
                                                  const BItem *__restrict__ items, int nitems, int P, int np,
                                                  int *__restrict__ count) {
    extern __shared__ int s_hist[];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BItem b = items[it];
        for (int k = threadIdx.x; k < np; k += kBT) s_hist[k] = 0;
        __syncthreads();
        const int e0 = row[b.x0], e1 = row[b.x1];
        for (int e = e0 + threadIdx.x; e < e1; e += kBT) atomicAdd(&s_hist[col[e] / P], 1);
        __syncthreads();
        for (int k = threadIdx.x; k < np; k += kBT)
            if (s_hist[k]) atomicAdd(&count[(long long)b.d * np + k], s_hist[k]);
        __syncthreads();
    }
}

// padded tile sizes in destination-major (pb) and source-major (pa) order
__global__ void k_pb_pad(const int *__restrict__ count, int np, long long *__restrict__ pb, long long *__restrict__ pa) {
    const long long n = (long long)np * np;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long c = ((count[t] + kG - 1) / kG) * kG;
        const long long d = t / np, s = t % np;
        pb[t] = c;
        pa[s * np + d] = c;
    }
}

// pass 2: place every edge; ranks inside a tile come from a per-tile cursor reserved
// once per (item, tile) and LDS counters inside the item
__global__ __launch_bounds__(kBT) void k_pb_scatter(const int *__restrict__ col, const int *__restrict__ row,
                                                    const BItem *__restrict__ items, int nitems, int P, int np,
                                                    int *__restrict__ cursor, const long long *__restrict__ boff,
                                                    const long long *__restrict__ aoff, uint16_t *__restrict__ a_src,
                                                    uint16_t *__restrict__ b_dst, int *__restrict__ amap) {
    extern __shared__ int s_mem[];
    int *s_hist = s_mem;              // np
    int *s_row = s_mem + np;          // item rows + 1 (<= P + 1)
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BItem b = items[it];
        const int nr = b.x1 - b.x0;
        for (int k = threadIdx.x; k < np; k += kBT) s_hist[k] = 0;
        for (int k = threadIdx.x; k <= nr; k += kBT) s_row[k] = row[b.x0 + k];
        __syncthreads();
        const int e0 = s_row[0], e1 = s_row[nr];
        for (int e = e0 + threadIdx.x; e < e1; e += kBT) atomicAdd(&s_hist[col[e] / P], 1);
        __syncthreads();
        for (int k = threadIdx.x; k < np; k += kBT) {
            const int h = s_hist[k];
            s_hist[k] = h ? atomicAdd(&cursor[(long long)b.d * np + k], h) : 0;
        }
        __syncthreads();
        for (int e = e0 + threadIdx.x; e < e1; e += kBT) {
            const int y = col[e];
            const int s = y / P;
            const int x = b.x0 + row_of(s_row, nr + 1, e);
            const int k = atomicAdd(&s_hist[s], 1);
            const long long tb = (long long)b.d * np + s, ta = (long long)s * np + b.d;
            const long long jb = boff[tb] + k, ia = aoff[ta] + k;
            b_dst[jb] = (uint16_t)(x - b.d * P);
            a_src[ia] = (uint16_t)(y - s * P);
            if ((k % kG) == 0) amap[ia / kG] = (int)(boff[tb] + k);
        }
        __syncthreads();
    }
}

// ---- one iteration ---------------------------------------------------------------
// unit = {partition, first slot, end slot, split flag}
// partition p holds local ids [0, part_count(p)) under the round-robin numbering
__device__ __forceinline__ int part_count(int p, int nreal, int np) { return nreal / np + (p < nreal % np ? 1 : 0); }

__global__ __launch_bounds__(kBT) void k_pb_scatter_msg(Ctl *ctl, const int4 *__restrict__ units, int nunits, int P, int nv, int np,
                                                        const float *__restrict__ rank, const int *__restrict__ deg,
                                                        const uint16_t *__restrict__ a_src, const int *__restrict__ amap,
                                                        float *__restrict__ msg) {
    if (ctl->mode == MODE_DONE) return;
    extern __shared__ float s_c[];
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int4 un = units[u];
        const int v0 = un.x * P;
        const int pv = part_count(un.x, nv, np);
        for (int k = threadIdx.x; k < pv; k += kBT) {
            float c;
            PRHooks::get_partial_result(&c, rank[v0 + k], deg[v0 + k]);
            s_c[k] = c;
        }
        __syncthreads();
        // 8 slots per lane: one 16-byte id load, two 16-byte message stores into the
        // matching B run (8 slots never straddle a 16-slot group); padding writes 0
        for (int i0 = un.y + threadIdx.x * 8; i0 < un.z; i0 += kBT * 16) {
            const int i1 = i0 + kBT * 8;
            const bool has1 = i1 < un.z;
            const uint4 ida = ld_stream_u4(&a_src[i0]);
            const int basea = amap[i0 / kG] + (i0 % kG);
            uint4 idb = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
            int baseb = 0;
            if (has1) { idb = ld_stream_u4(&a_src[i1]); baseb = amap[i1 / kG] + (i1 % kG); }
            const uint32_t w[8] = {ida.x, ida.y, ida.z, ida.w, idb.x, idb.y, idb.z, idb.w};
            float m[16];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t a = w[k] & 0xFFFFu, b = w[k] >> 16;
                m[2 * k] = a != 0xFFFFu ? s_c[a] : 0.f;
                m[2 * k + 1] = b != 0xFFFFu ? s_c[b] : 0.f;
            }
            *(float4 *)&msg[basea] = make_float4(m[0], m[1], m[2], m[3]);
            *(float4 *)&msg[basea + 4] = make_float4(m[4], m[5], m[6], m[7]);
            if (has1) {
                *(float4 *)&msg[baseb] = make_float4(m[8], m[9], m[10], m[11]);
                *(float4 *)&msg[baseb + 4] = make_float4(m[12], m[13], m[14], m[15]);
            }
        }
        __syncthreads();
    }
}

// apply the PR rule at v from its FP64 sum; returns (changed, |f1 - old|)
__device__ __forceinline__ int pr_apply(int v, double sum, int variant, float damping, float tol,
                                        const int *__restrict__ deg, float *__restrict__ rank, int *__restrict__ active,
                                        double *res) {
    float f1;
    const float old = rank[v];
    const int ch = PRHooks::update_condition(&rank[v], &active[v], sum, variant, damping, tol, &f1);
    *res += fabs((double)f1 - (double)old);
    return ch;
}

__global__ __launch_bounds__(kBT) void k_pb_gather(Ctl *ctl, const int4 *__restrict__ units, int nunits, int P, int nv, int np,
                                                   int variant, float damping, float tol,
                                                   const uint16_t *__restrict__ b_dst, const float *__restrict__ msg,
                                                   const int *__restrict__ deg, float *__restrict__ rank,
                                                   int *__restrict__ active, double *__restrict__ psum) {
    if (ctl->mode == MODE_DONE) return;
    extern __shared__ double s_acc[];
    __shared__ double s_tmp[kBT / MG_WAVE];
    double res = 0.0;
    int ch = 0;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int4 un = units[u];
        const int v0 = un.x * P;
        const int pv = part_count(un.x, nv, np);
        for (int k = threadIdx.x; k < P; k += kBT) s_acc[k] = 0.0;
        __syncthreads();
        // 8 slots per lane; the next vector's loads are issued before this vector's
        // LDS atomics, so HBM requests stay in flight while the atomics run
        int j = un.y + threadIdx.x * 8;
        bool have = j < un.z;
        uint4 nid = make_uint4(0u, 0u, 0u, 0u);
        float4 nm0 = make_float4(0.f, 0.f, 0.f, 0.f), nm1 = nm0;
        if (have) { nid = ld_stream_u4(&b_dst[j]); nm0 = ld_stream_f4(&msg[j]); nm1 = ld_stream_f4(&msg[j + 4]); }
        while (have) {
            const uint4 id = nid;
            const float4 m0 = nm0, m1 = nm1;
            const int jn = j + kBT * 8;
            const bool have_n = jn < un.z;
            if (have_n) { nid = ld_stream_u4(&b_dst[jn]); nm0 = ld_stream_f4(&msg[jn]); nm1 = ld_stream_f4(&msg[jn + 4]); }
            const uint32_t w[4] = {id.x, id.y, id.z, id.w};
            const float m[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t a = w[k] & 0xFFFFu, b = w[k] >> 16;
                if (a != 0xFFFFu) atomicAdd(&s_acc[a], (double)m[2 * k]);
                if (b != 0xFFFFu) atomicAdd(&s_acc[b], (double)m[2 * k + 1]);
            }
            j = jn;
            have = have_n;
        }
        __syncthreads();
        if (un.w) {  // partition split over several units: add partials, applied later
            for (int k = threadIdx.x; k < pv; k += kBT) {
                const double a = s_acc[k];
                if (a != 0.0) atomicAdd(&psum[v0 + k], a);
            }
        } else {
            for (int k = threadIdx.x; k < pv; k += kBT)
                ch |= pr_apply(v0 + k, s_acc[k], variant, damping, tol, deg, rank, active, &res);
        }
        __syncthreads();
    }
    res = block_sum<kBT / MG_WAVE>(res, s_tmp);
    const int any = __syncthreads_or(ch);
    if (threadIdx.x == 0) {
        if (res != 0.0) atomicAdd(&ctl->resid, res);
        if (any) atomicOr(&ctl->changed, 1);
    }
}

// vertices of split partitions: apply from the FP64 partial sums and clear them
__global__ __launch_bounds__(kBT) void k_pb_apply_split(Ctl *ctl, const int *__restrict__ parts, int nparts, int P, int nv, int np,
                                                        int variant, float damping, float tol, const int *__restrict__ deg,
                                                        float *__restrict__ rank, int *__restrict__ active,
                                                        double *__restrict__ psum) {
    if (ctl->mode == MODE_DONE) return;
    __shared__ double s_tmp[kBT / MG_WAVE];
    double res = 0.0;
    int ch = 0;
    const long long total = (long long)nparts * P;
    for (long long i = (long long)blockIdx.x * kBT + threadIdx.x; i < total; i += (long long)gridDim.x * kBT) {
        const int p = parts[i / P], k = (int)(i % P);
        if (k >= part_count(p, nv, np)) continue;
        const int v = p * P + k;
        ch |= pr_apply(v, psum[v], variant, damping, tol, deg, rank, active, &res);
        psum[v] = 0.0;
    }
    res = block_sum<kBT / MG_WAVE>(res, s_tmp);
    const int any = __syncthreads_or(ch);
    if (threadIdx.x == 0) {
        if (res != 0.0) atomicAdd(&ctl->resid, res);
        if (any) atomicOr(&ctl->changed, 1);
    }
}

__global__ void k_pb_zero_f64(double *p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0.0;
}

// internal CSR under the degree round-robin numbering: degree of every internal id
__global__ void k_pb_ideg(int nv, const int *__restrict__ row, const int *__restrict__ perm, int *__restrict__ ideg) {
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += gridDim.x * blockDim.x) ideg[perm[x]] = row[x + 1] - row[x];
}
__global__ void k_pb_relabel(int nv, const int *__restrict__ row, const int *__restrict__ col, const int *__restrict__ perm,
                             const long long *__restrict__ irow64, int *__restrict__ icol) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int x = wave; x < nv; x += nw) {
        const long long o = irow64[perm[x]];
        const int b = row[x], e = row[x + 1];
        for (int p = b + lane(); p < e; p += MG_WAVE) icol[o + (p - b)] = perm[col[p]];
    }
}
__global__ void k_pb_narrow(long long n, const long long *__restrict__ a, int *__restrict__ b) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (long long)gridDim.x * blockDim.x) b[i] = (int)a[i];
}

}  // namespace

int mg_scan64(hipStream_t s, const long long *in64, const int *in32, long long n, long long *out);
int mg_pb_layout_build(mg_engine *e, const std::vector<int> &hrow, const int *irow, const int *icol);

// Build the partitioned layout on the device.  Vertices are renumbered first: sorted
// by degree (descending, ties by id) and dealt round-robin into the np partitions, so
// every partition receives about E / np edges and one workgroup per partition is
// balanced (the degree-group renumbering the reference's itable is meant for,
// gen_structure.cu:152-254).  internal id = (rank mod np) * P + rank / np.
int mg_pb_build(mg_engine *e, const mg_graph *g) {
    mg_pb_layout &L = e->pb;
    hipStream_t s = e->stream;
    const int nv = e->nv;
    // Partition count and size.  With the round-robin numbering every partition holds
    // ceil(nv / np) vertices and about E / np edges, so one workgroup per partition is
    // balanced and np can be any count: it is chosen as a multiple of the CU count,
    // trading tile padding (np^2 tiles, 7.5 padding slots on average) against the
    // rounds of workgroups each CU runs (LDS of the FP64 accumulators bounds how many
    // are resident).  part_log2 > 0 forces P = 2^part_log2 instead.
    int cus = 256;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    int np, P;
    if (e->opt.part_log2 > 0) {
        P = 1 << std::min(e->opt.part_log2, 14);
        np = nv > 0 ? (nv + P - 1) / P : 1;
    } else {
        double best = 1e300;
        np = 1; P = std::max(1, nv);
        for (int k = 1; k <= 64; ++k) {
            const long long cnp = (long long)cus * k;
            const long long cP = ((long long)nv + cnp - 1) / cnp;
            if (cP > 16384) continue;
            if (cP < 64 && k > 1) break;
            const int res = (int)std::min<long long>(4, (160 * 1024) / (8 * cP + 2048));  // resident per CU
            if (res < 1) continue;
            const long long rounds = (k + res - 1) / res;
            const double per_block = ((double)e->ne + 7.5 * (double)cnp * (double)cnp) / (double)cnp;
            const double cost = (double)rounds * per_block;
            if (cost < best) { best = cost; np = (int)cnp; P = (int)std::max<long long>(1, cP); }
        }
        if (nv < cus) { np = 1; P = std::max(1, nv); }
    }
    L.plog = 0;
    L.P = P;
    L.G = kG;
    L.np = np;
    const long long nvi = (long long)np * P;
    e->pr_nvi = nvi;
    MG_HIP(hipMalloc(&e->psum, sizeof(double) * (nvi + 64)));
    if ((long long)np * np > (1ll << 31)) return mg_fail(MG_ERR_LIMIT, "%d partitions: tile table too large", np);
    // The internal rows are in-edges (a destination gathers from its sources).  A
    // symmetric CSR is its own transpose; a directed one is transposed here on the host.
    std::vector<int> trow, tcol;
    const int *hr = g->row.data();
    const int *crow = e->row, *ccol = e->col;
    int *d_trow = nullptr, *d_tcol = nullptr, *ilen = nullptr;
    if (!g->symmetric) {
        trow.assign((size_t)nv + 1, 0);
        tcol.resize((size_t)std::max<long long>(1, e->ne));
        for (long long p = 0; p < e->ne; ++p) trow[(size_t)g->col[p] + 1]++;
        for (int v = 0; v < nv; ++v) trow[v + 1] += trow[v];
        std::vector<int> cur(trow.begin(), trow.end() - 1);
        for (int u = 0; u < nv; ++u)
            for (int p = g->row[u]; p < g->row[u + 1]; ++p) tcol[cur[g->col[p]]++] = u;
        MG_HIP(hipMalloc(&d_trow, sizeof(int) * ((size_t)nv + 1)));
        MG_HIP(hipMalloc(&d_tcol, sizeof(int) * tcol.size()));
        MG_HIP(hipMemcpy(d_trow, trow.data(), sizeof(int) * ((size_t)nv + 1), hipMemcpyHostToDevice));
        MG_HIP(hipMemcpy(d_tcol, tcol.data(), sizeof(int) * tcol.size(), hipMemcpyHostToDevice));
        hr = trow.data();
        crow = d_trow;
        ccol = d_tcol;
    }
    // degree round-robin numbering (host: one sort of nv in-degrees)
    std::vector<int> order(nv);
    for (int v = 0; v < nv; ++v) order[v] = v;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return (hr[a + 1] - hr[a]) > (hr[b + 1] - hr[b]); });
    e->pr_perm_h.assign(nv, 0);
    for (int r = 0; r < nv; ++r) e->pr_perm_h[order[r]] = (r % np) * P + r / np;
    int *perm = nullptr, *irow = nullptr, *icol = nullptr;
    long long *irow64 = nullptr;
    MG_HIP(hipMalloc(&perm, sizeof(int) * std::max(1, nv)));
    MG_HIP(hipMemcpy(perm, e->pr_perm_h.data(), sizeof(int) * nv, hipMemcpyHostToDevice));
    MG_HIP(hipMalloc(&e->deg, sizeof(int) * (nvi + 64)));
    MG_HIP(hipMemsetAsync(e->deg, 0, sizeof(int) * (nvi + 64), s));
    // out-degrees (the divisor of every message) and internal row lengths
    hipLaunchKernelGGL(k_pb_ideg, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, s, nv, e->row, perm, e->deg);
    if (crow != e->row) {
        MG_HIP(hipMalloc(&ilen, sizeof(int) * (nvi + 64)));
        MG_HIP(hipMemsetAsync(ilen, 0, sizeof(int) * (nvi + 64), s));
        hipLaunchKernelGGL(k_pb_ideg, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, s, nv, crow, perm, ilen);
    }
    MG_HIP(hipMalloc(&irow64, sizeof(long long) * (nvi + 1)));
    MG_TRY(mg_scan64(s, nullptr, ilen ? ilen : e->deg, nvi, irow64));
    MG_HIP(hipMalloc(&irow, sizeof(int) * (nvi + 1)));
    MG_HIP(hipMalloc(&icol, sizeof(int) * std::max<long long>(1, e->ne)));
    hipLaunchKernelGGL(k_pb_narrow, dim3(grid_for(nvi + 1, 256, 8192)), dim3(256), 0, s, nvi, irow64, irow);
    hipLaunchKernelGGL(k_pb_relabel, dim3(grid_for(nv, 4, 8192)), dim3(256), 0, s, nv, crow, ccol, perm, irow64, icol);
    std::vector<int> hrow(nvi + 1);
    MG_HIP(hipMemcpyAsync(hrow.data(), irow, sizeof(int) * (nvi + 1), hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    hipFree(irow64);
    hipFree(perm);
    if (ilen) hipFree(ilen);
    if (d_trow) hipFree(d_trow);
    if (d_tcol) hipFree(d_tcol);
    int rc = mg_pb_layout_build(e, hrow, irow, icol);
    hipFree(irow);
    hipFree(icol);
    return rc;
}

// the tile streams over an internal CSR (irow/icol on the device, hrow on the host)
int mg_pb_layout_build(mg_engine *e, const std::vector<int> &hrow, const int *irow, const int *icol) {
    mg_pb_layout &L = e->pb;
    hipStream_t s = e->stream;
    const int P = L.P;
    const int np = L.np;
    const int nvi = (int)e->pr_nvi;
    // build items: row ranges of one partition with at most kBuildEdges edges (or one row)
    std::vector<BItem> items;
    for (int d = 0; d < np; ++d) {
        const int xe = std::min(nvi, (d + 1) * P);
        int x = d * P;
        while (x < xe) {
            int y = x + 1;
            while (y < xe && hrow[y + 1] - hrow[x] <= kBuildEdges) ++y;
            items.push_back(BItem{d, x, y, 0});
            x = y;
        }
    }
    const long long nt = (long long)np * np;
    BItem *d_items = nullptr;
    int *count = nullptr, *cursor = nullptr;
    long long *pb = nullptr, *pa = nullptr, *boff = nullptr, *aoff = nullptr;
    auto cleanup = [&]() {
        hipFree(d_items); hipFree(count); hipFree(cursor); hipFree(pb); hipFree(pa); hipFree(boff); hipFree(aoff);
    };
    if (hipMalloc(&d_items, sizeof(BItem) * std::max<size_t>(1, items.size())) != hipSuccess ||
        hipMalloc(&count, sizeof(int) * nt) != hipSuccess || hipMalloc(&cursor, sizeof(int) * nt) != hipSuccess ||
        hipMalloc(&pb, sizeof(long long) * nt) != hipSuccess || hipMalloc(&pa, sizeof(long long) * nt) != hipSuccess ||
        hipMalloc(&boff, sizeof(long long) * (nt + 1)) != hipSuccess || hipMalloc(&aoff, sizeof(long long) * (nt + 1)) != hipSuccess) {
        cleanup();
        return mg_fail(MG_ERR_NOMEM, "partitioned layout: tile tables do not fit");
    }
    MG_HIP(hipMemcpy(d_items, items.data(), sizeof(BItem) * items.size(), hipMemcpyHostToDevice));
    MG_HIP(hipMemsetAsync(count, 0, sizeof(int) * nt, s));
    MG_HIP(hipMemsetAsync(cursor, 0, sizeof(int) * nt, s));
    const int nitems = (int)items.size();
    const size_t lds_count = sizeof(int) * np;
    const size_t lds_scatter = sizeof(int) * (np + P + 1);
    // dynamic LDS above the 64 KiB default needs an explicit per-kernel limit
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_count, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(lds_count, 1024)));
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(lds_scatter, 1024)));
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_scatter_msg, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(float) * P)));
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_gather, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(double) * P)));
    if (lds_scatter > 160 * 1024) { cleanup(); return mg_fail(MG_ERR_LIMIT, "partitioned layout: %d partitions exceed LDS", np); }
    hipLaunchKernelGGL(k_pb_count, dim3(grid_for(nitems, 1, 4096)), dim3(kBT), lds_count, s, icol, irow, d_items, nitems,
                       L.P, np, count);
    hipLaunchKernelGGL(k_pb_pad, dim3(grid_for(nt, 256, 8192)), dim3(256), 0, s, count, np, pb, pa);
    int rc = mg_scan64(s, pb, nullptr, nt, boff);
    if (rc == MG_OK) rc = mg_scan64(s, pa, nullptr, nt, aoff);
    if (rc != MG_OK) { cleanup(); return rc; }
    long long slots = 0;
    MG_HIP(hipMemcpy(&slots, boff + nt, sizeof(long long), hipMemcpyDeviceToHost));
    if (slots >= (1ll << 31)) { cleanup(); return mg_fail(MG_ERR_LIMIT, "partitioned layout: %lld slots exceed int32", slots); }
    L.slots = slots;
    const long long sl = slots > 0 ? slots : kG;
    if (hipMalloc(&L.a_src, sizeof(uint16_t) * sl) != hipSuccess || hipMalloc(&L.b_dst, sizeof(uint16_t) * sl) != hipSuccess ||
        hipMalloc(&L.amap, sizeof(int) * (sl / kG + 1)) != hipSuccess || hipMalloc(&L.msg, sizeof(float) * sl) != hipSuccess) {
        cleanup();
        return mg_fail(MG_ERR_NOMEM, "partitioned layout: %lld slots do not fit", slots);
    }
    hipLaunchKernelGGL(k_fill_u16, dim3(grid_for(sl, 256, 8192)), dim3(256), 0, s, L.a_src, sl);
    hipLaunchKernelGGL(k_fill_u16, dim3(grid_for(sl, 256, 8192)), dim3(256), 0, s, L.b_dst, sl);
    MG_HIP(hipMemsetAsync(L.amap, 0, sizeof(int) * (sl / kG + 1), s));
    MG_HIP(hipMemsetAsync(L.msg, 0, sizeof(float) * sl, s));
    hipLaunchKernelGGL(k_pb_scatter, dim3(grid_for(nitems, 1, 4096)), dim3(kBT), lds_scatter, s, icol, irow, d_items,
                       nitems, L.P, np, cursor, boff, aoff, L.a_src, L.b_dst, L.amap);
    MG_HIP(hipGetLastError());
    // partition starts of both streams -> work units of about equal slot counts
    std::vector<long long> hb(nt + 1), ha(nt + 1);
    MG_HIP(hipMemcpy(hb.data(), boff, sizeof(long long) * (nt + 1), hipMemcpyDeviceToHost));
    MG_HIP(hipMemcpy(ha.data(), aoff, sizeof(long long) * (nt + 1), hipMemcpyDeviceToHost));
    MG_HIP(hipStreamSynchronize(s));
    cleanup();
    // unit bounds stay multiples of the 16-slot group (vector loads of 8 slots)
    const long long unit = ((std::max<long long>(32768, 2 * (slots / std::max(1, np)) + 1) + kG - 1) / kG) * kG;
    std::vector<int4> ua, ub;
    std::vector<int> split;
    for (int p = 0; p < np; ++p) {
        const long long a0 = ha[(long long)p * np], a1 = ha[(long long)(p + 1) * np];
        for (long long x = a0; x < a1; x += unit) ua.push_back(make_int4(p, (int)x, (int)std::min(a1, x + unit), 0));
        const long long b0 = hb[(long long)p * np], b1 = hb[(long long)(p + 1) * np];
        const int nb = (int)((b1 - b0 + unit - 1) / unit);
        if (nb <= 1) ub.push_back(make_int4(p, (int)b0, (int)b1, 0));
        else {
            split.push_back(p);
            for (long long x = b0; x < b1; x += unit) ub.push_back(make_int4(p, (int)x, (int)std::min(b1, x + unit), 1));
        }
    }
    L.nA = (int)ua.size();
    L.nB = (int)ub.size();
    L.nApply = (int)split.size();
    MG_HIP(hipMalloc(&L.unitsA, sizeof(int4) * std::max<size_t>(1, ua.size())));
    MG_HIP(hipMalloc(&L.unitsB, sizeof(int4) * std::max<size_t>(1, ub.size())));
    MG_HIP(hipMalloc(&L.split_parts, sizeof(int) * std::max<size_t>(1, split.size())));
    if (!ua.empty()) MG_HIP(hipMemcpy(L.unitsA, ua.data(), sizeof(int4) * ua.size(), hipMemcpyHostToDevice));
    if (!ub.empty()) MG_HIP(hipMemcpy(L.unitsB, ub.data(), sizeof(int4) * ub.size(), hipMemcpyHostToDevice));
    if (!split.empty()) MG_HIP(hipMemcpy(L.split_parts, split.data(), sizeof(int) * split.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pb_zero_f64, dim3(grid_for((long long)nvi + 64, 256, 4096)), dim3(256), 0, s, e->psum, (long long)nvi + 64);
    MG_HIP(hipStreamSynchronize(s));
    return MG_OK;
}

// one dense iteration over the partitioned layout (without the control kernels)
int mg_pb_iteration(mg_engine *e) {
    mg_pb_layout &L = e->pb;
    hipStream_t s = e->stream;
    const int variant = e->algo == MG_PR_T ? 0 : 1;
    const int P = L.P;
    const int G = e->opt.grid_blocks > 0 ? e->opt.grid_blocks : 2048;
    int slot = -1;
    // algorithmic bytes of one iteration (pull formulation): 8E + 16V
    const double alg = 8.0 * (double)e->ne + 16.0 * e->nv;
    if (e->prof) mg_prof_begin(e, "k_pb_scatter_msg", 0.5 * alg, &slot);
    hipLaunchKernelGGL(k_pb_scatter_msg, dim3(grid_for(L.nA, 1, G)), dim3(kBT), sizeof(float) * P, s, e->ctl, L.unitsA, L.nA,
                       L.P, e->nv, L.np, e->rank, e->deg, L.a_src, L.amap, L.msg);
    if (e->prof) mg_prof_end(e, slot);
    if (e->prof) mg_prof_begin(e, "k_pb_gather", 0.5 * alg, &slot);
    hipLaunchKernelGGL(k_pb_gather, dim3(grid_for(L.nB, 1, G)), dim3(kBT), sizeof(double) * P, s, e->ctl, L.unitsB, L.nB,
                       L.P, e->nv, L.np, variant, e->opt.pr_damping, e->opt.pr_tolerance, L.b_dst, L.msg, e->deg, e->rank,
                       e->vfc, e->psum);
    if (e->prof) mg_prof_end(e, slot);
    if (L.nApply > 0)
        hipLaunchKernelGGL(k_pb_apply_split, dim3(grid_for((long long)L.nApply * P, kBT, G)), dim3(kBT), 0, s, e->ctl,
                           L.split_parts, L.nApply, L.P, e->nv, L.np, variant, e->opt.pr_damping, e->opt.pr_tolerance, e->deg,
                           e->rank, e->vfc, e->psum);
    MG_HIP(hipGetLastError());
    return MG_OK;
}
