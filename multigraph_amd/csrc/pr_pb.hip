// PageRank over the partitioned layout: the role of gen_structure.cu (layout build)
// and of _phase1 / _phase2 in processi.cu (one dense iteration), re-designed for gfx950.
//
// Layout.  Vertices are cut into np partitions of P ids.  An edge of row x with
// column y carries rank mass from y (source) to x (destination); it belongs to tile
// (s, d) = (y / P, x / P).  Inside a tile the edges are sorted by source, and every
// run of edges with one source shares one message: a message slot stands for a
// (source vertex, destination partition) pair.  Streams, each tile padded to groups
// of 8 slots:
//   a_src (message slots, source-major: s, then d): local source id (16 bit);
//   amap[g]: for the g-th group of a_src slots, the position of their messages in the
//            destination-major message array msg;
//   b_dst (edge slots, destination-major: d, then s): local destination id, with bit
//            15 set where a new message (source) starts;
//   gbase[g]: message index of the first edge slot of the g-th group;
//   b_w (SSSP only): per edge slot, its weight (8 or 16 bit).
//
// One iteration = two streaming passes, no random HBM access and no global atomics:
//   scatter: per source partition, stage c = rank / degree of its P vertices in LDS,
//            stream a_src and write msg[amap position] = c[local source];
//   gather:  per destination partition, every wave walks 512-slot segments of its
//            edge stream, stages the segment's messages (at most 520) in its own LDS
//            slice and adds each edge's message into P FP64 LDS accumulators (native
//            LDS FP64 atomic add), then the workgroup applies the PR_T / PR_D rule to
//            the partition's vertices, writes the ranks and (fused scatter) the
//            partition's messages for the next iteration.
// Work is cut into units of about equal slot counts; a destination partition whose
// stream is cut into several units adds its partial sums into an FP64 vector instead
// and is applied by a separate pass.  Traffic per iteration, with M message slots and
// S edge slots: 6.5 B * M in the scatter, 4 B * M + 2.5 B * S in the gather, plus
// 20 B per vertex.  On RMAT graphs M is a half to three quarters of S.
#include <algorithm>
#include <type_traits>
#include <vector>
#include "engine.hpp"
#include "hook_calls.hpp"

using namespace mg;

namespace {

constexpr int kG = 8;           // slots per group of either stream (one 16-byte vector of ids)
constexpr int kBT = 512;        // threads per workgroup of the layout and phase kernels
constexpr int kBuildEdges = 1 << 16;


int grid_for(long long items, int per_block, int cap) {
    long long g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

__global__ void k_fill_u16(uint16_t *p, long long n, uint16_t v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, long long n, uint32_t v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
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

// padded edge slots per tile (destination-major)
__global__ void k_pb_pad(const int *__restrict__ count, long long nt, long long *__restrict__ pb) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (long long)gridDim.x * blockDim.x)
        pb[t] = ((count[t] + kG - 1) / kG) * kG;
}

// padded message slots per tile, in destination-major (pb) and source-major (pa) order
__global__ void k_pb_pad_pairs(const int *__restrict__ npair, int np, long long *__restrict__ pb, long long *__restrict__ pa) {
    const long long n = (long long)np * np;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long c = ((npair[t] + kG - 1) / kG) * kG;
        const long long d = t / np, s = t % np;
        pb[t] = c;
        pa[s * np + d] = c;
    }
}

// pass 2: place every edge of an item as key (local source << 16 | local destination)
// in its tile; ranks inside a tile come from a per-tile cursor reserved once per
// (item, tile) and LDS counters inside the item
__global__ __launch_bounds__(kBT) void k_pb_place(const int *__restrict__ col, const int *__restrict__ row,
                                                  const BItem *__restrict__ items, int nitems, int P, int np,
                                                  int *__restrict__ cursor, const long long *__restrict__ boff,
                                                  uint32_t *__restrict__ keys) {
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
            keys[boff[(long long)b.d * np + s] + k] = ((uint32_t)(y - s * P) << 16) | (uint32_t)(x - b.d * P);
        }
        __syncthreads();
    }
}

// pass 3: sort every tile's keys (bitonic network in LDS) and count its messages
// (distinct sources); tiles above kSortMax keys are left to the host (npair = -1)
constexpr int kSortMax = 8192;
__global__ __launch_bounds__(kBT) void k_pb_tile_sort(long long nt, const int *__restrict__ count, const long long *__restrict__ boff,
                                                      uint32_t *__restrict__ keys, int *__restrict__ npair) {
    __shared__ uint32_t s_k[kSortMax];
    __shared__ int s_tmp[kBT / MG_WAVE];
    for (long long t = blockIdx.x; t < nt; t += gridDim.x) {
        const int n = count[t];
        if (n == 0 || n > kSortMax) {
            if (threadIdx.x == 0) npair[t] = n == 0 ? 0 : -1;
            continue;
        }
        int n2 = MG_WAVE;
        while (n2 < n) n2 <<= 1;
        uint32_t *k = keys + boff[t];
        for (int i = threadIdx.x; i < n2; i += kBT) s_k[i] = i < n ? k[i] : 0xFFFFFFFFu;
        __syncthreads();
        for (int size = 2; size <= n2; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = threadIdx.x; i < n2; i += kBT) {
                    const int j = i ^ stride;
                    if (j > i) {
                        const uint32_t a = s_k[i], b = s_k[j];
                        if ((a > b) == ((i & size) == 0)) { s_k[i] = b; s_k[j] = a; }
                    }
                }
                __syncthreads();
            }
        int c = 0;
        for (int i = threadIdx.x; i < n; i += kBT) {
            c += (i == 0 || (s_k[i] >> 16) != (s_k[i - 1] >> 16)) ? 1 : 0;
            k[i] = s_k[i];
        }
        c = block_sum<kBT / MG_WAVE>(c, s_tmp);
        if (threadIdx.x == 0) npair[t] = c;
    }
}

// SSSP: the weight of every edge slot, in the destination-major order of b_dst.  A
// tile's sorted keys give each slot's local source and destination; the weight is
// found in the engine's CSR (rows sorted by column) through the inverse numbering.
template <class W>
__global__ __launch_bounds__(kBT) void k_pb_weights(long long nt, int np, int P, const int *__restrict__ count,
                                                    const long long *__restrict__ boff, const uint32_t *__restrict__ keys,
                                                    const int *__restrict__ inv, const int *__restrict__ row,
                                                    const int *__restrict__ col, const int *__restrict__ w,
                                                    W *__restrict__ b_w) {
    for (long long t = blockIdx.x; t < nt; t += gridDim.x) {
        const int n = count[t];
        if (n == 0) continue;
        const long long d = t / np, s = t % np, eb = boff[t];
        for (int j = threadIdx.x; j < n; j += kBT) {
            const uint32_t key = keys[eb + j];
            const int src = inv[s * P + (int)(key >> 16)], dst = inv[d * P + (int)(key & 0xFFFFu)];
            // edge positions reach 2^31 - 1: the midpoint is taken without adding them
            int lo = row[dst], hi = row[dst + 1] - 1;
            while (lo < hi) { const int mid = lo + ((hi - lo) >> 1); if (col[mid] < src) lo = mid + 1; else hi = mid; }
            b_w[eb + j] = (W)w[lo];
        }
    }
}

// pass 4: per tile, number the messages (a block scan of the new-source flags) and
// write b_dst, gbase, a_src and amap
__global__ __launch_bounds__(kBT) void k_pb_emit(long long nt, int np, const int *__restrict__ count,
                                                 const long long *__restrict__ boff, const long long *__restrict__ poffB,
                                                 const long long *__restrict__ poffA, const uint32_t *__restrict__ keys,
                                                 uint16_t *__restrict__ b_dst, int *__restrict__ gbase,
                                                 uint16_t *__restrict__ a_src, int *__restrict__ amap) {
    __shared__ int s_tmp[kBT / MG_WAVE + 1];
    for (long long t = blockIdx.x; t < nt; t += gridDim.x) {
        const int n = count[t];
        if (n == 0) continue;
        const long long d = t / np, s = t % np;
        const long long eb = boff[t], pb0 = poffB[t], pa0 = poffA[s * np + d];
        int carry = 0;  // messages started before this chunk of the tile
        for (int c0 = 0; c0 < n; c0 += kBT) {
            const int j = c0 + threadIdx.x;
            uint32_t key = 0;
            int flag = 0;
            if (j < n) {
                key = keys[eb + j];
                flag = (j == 0 || (keys[eb + j - 1] >> 16) != (key >> 16)) ? 1 : 0;
            }
            int tot;
            const int pre = block_excl_scan<kBT / MG_WAVE>(flag, s_tmp, &tot);
            const int pi = carry + pre + flag - 1;  // message of slot j
            if (j < n) {
                b_dst[eb + j] = (uint16_t)((key & 0x7FFFu) | ((uint32_t)flag << 15));
                if (flag) {
                    a_src[pa0 + pi] = (uint16_t)(key >> 16);
                    if ((pi % kG) == 0) amap[(pa0 + pi) / kG] = (int)(pb0 + pi);
                }
                if (((eb + j) % kG) == 0) gbase[(eb + j) / kG] = (int)(pb0 + pi);
            }
            carry += tot;
            __syncthreads();
        }
    }
}

// ---- one iteration ---------------------------------------------------------------
// unit = {partition, first slot, end slot, split flag}
// partition p holds local ids [0, part_count(p)) under the round-robin numbering
__device__ __forceinline__ int part_count(int p, int nreal, int np) { return nreal / np + (p < nreal % np ? 1 : 0); }

// Messages of one source partition: c (rank / degree of its vertices) is staged in
// LDS, the partition's source-major message slots [a0, a1) are streamed, and every
// group of 8 is written into its destination-major run.  8 slots (one group) per
// thread and step: one 16-byte id load, two 16-byte streaming stores; padding writes
// 0.  The next group's loads are issued before this group is written (two register
// sets, no copies); loads past the range repeat its last group.
template <int NT, class T = float>
__device__ __forceinline__ void scatter_part(int a0, int a1, const T *__restrict__ s_c, const uint16_t *__restrict__ a_src,
                                             const int *__restrict__ amap, T *__restrict__ msg) {
    const int step = NT * kG;
    auto ld = [&](int i, uint4 &id, int &base) {
        const int c = min(i, a1 - kG);
        id = ld_stream_u4(&a_src[c]);
        base = amap[c / kG];
    };
    auto emit = [&](int i, const uint4 &id, int base) {
        if (i >= a1) return;
        const uint32_t w[4] = {id.x, id.y, id.z, id.w};
        T m[8];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t a = w[k] & 0xFFFFu, b = w[k] >> 16;
            m[2 * k] = a != 0xFFFFu ? s_c[a] : T(0);
            m[2 * k + 1] = b != 0xFFFFu ? s_c[b] : T(0);
        }
        // streaming stores: the messages are read once, by the next gather
        st_stream_f4(&msg[base], m[0], m[1], m[2], m[3]);
        st_stream_f4(&msg[base + 4], m[4], m[5], m[6], m[7]);
    };
    int i = a0 + (int)threadIdx.x * kG;
    if (i < a1) {
        uint4 ia, ib;
        int ba, bb;
        ld(i, ia, ba);
        for (;;) {
            ld(i + step, ib, bb);
            emit(i, ia, ba);
            i += step;
            if (i >= a1) break;
            ld(i + step, ia, ba);
            emit(i, ib, bb);
            i += step;
            if (i >= a1) break;
        }
    }
}

// SSSP with 16-bit messages: as scatter_part, from int distances staged in LDS,
// each narrowed to 16 bits (65535 stands for "unreached or above 65534": it relaxes
// nothing), 8 messages per 16-byte streaming store (message runs start at multiples
// of 8 slots)
__device__ __forceinline__ uint32_t narrow16(int d) { return (uint32_t)min((unsigned)d, 0xFFFFu); }
template <int NT>
__device__ __forceinline__ void scatter_part16(int a0, int a1, const int *__restrict__ s_d, const uint16_t *__restrict__ a_src,
                                               const int *__restrict__ amap, uint16_t *__restrict__ msg) {
    const int step = NT * kG;
    auto ld = [&](int i, uint4 &id, int &base) {
        const int c = min(i, a1 - kG);
        id = ld_stream_u4(&a_src[c]);
        base = amap[c / kG];
    };
    auto emit = [&](int i, const uint4 &id, int base) {
        if (i >= a1) return;
        const uint32_t w[4] = {id.x, id.y, id.z, id.w};
        int o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t a = w[k] & 0xFFFFu, b = w[k] >> 16;
            const uint32_t ma = a != 0xFFFFu ? narrow16(s_d[a]) : 0xFFFFu;
            const uint32_t mb = b != 0xFFFFu ? narrow16(s_d[b]) : 0xFFFFu;
            o[k] = (int)(ma | (mb << 16));
        }
        st_stream_f4((int *)&msg[base], o[0], o[1], o[2], o[3]);
    };
    int i = a0 + (int)threadIdx.x * kG;
    if (i < a1) {
        uint4 ia, ib;
        int ba, bb;
        ld(i, ia, ba);
        for (;;) {
            ld(i + step, ib, bb);
            emit(i, ia, ba);
            i += step;
            if (i >= a1) break;
            ld(i + step, ia, ba);
            emit(i, ib, bb);
            i += step;
            if (i >= a1) break;
        }
    }
}

// message buffer an iteration reads (its number's parity) and the one it writes
__device__ __forceinline__ const float *msg_in(const Ctl *ctl, const float *msg, long long pairs) {
    return msg + (long long)(ctl->iter & 1) * pairs;
}
__device__ __forceinline__ float *msg_out(const Ctl *ctl, float *msg, long long pairs) {
    return msg + (long long)((ctl->iter + 1) & 1) * pairs;
}

// messages of partitions whose destination side was split over several gather units
// (their ranks are final only after k_pb_apply_split); unit = {partition, a0, a1, 0}
template <class H>
__global__ __launch_bounds__(kBT) void k_pb_scatter_msg(Ctl *ctl, const int4 *__restrict__ units, int nunits, int P, int nv, int np,
                                                        const float *__restrict__ rank, const int *__restrict__ deg,
                                                        const uint16_t *__restrict__ a_src, const int *__restrict__ amap,
                                                        float *__restrict__ msg, long long pairs) {
    if (ctl->mode == MODE_DONE) return;
    extern __shared__ float s_c[];
    float *out = msg_out(ctl, msg, pairs);
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int4 un = units[u];
        const int v0 = un.x * P;
        const int pv = part_count(un.x, nv, np);
        for (int k = threadIdx.x; k < pv; k += kBT) {
            float c;
            H::get_partial_result(&c, rank[v0 + k], deg[v0 + k]);
            s_c[k] = c;
        }
        __syncthreads();
        scatter_part<kBT>(un.y, un.z, s_c, a_src, amap, out);
        __syncthreads();
    }
}

// apply the PR rule at v from its FP64 sum; returns (changed, |f1 - old|)
template <class H>
__device__ __forceinline__ int pr_apply(int v, double sum, int variant, float damping, float tol,
                                        const int *__restrict__ deg, float *__restrict__ rank, int *__restrict__ active,
                                        double *res) {
    float f1;
    const float old = rank[v];
    const int ch = H::update_condition(&rank[v], &active[v], sum, variant, damping, tol, &f1);
    *res += fabs((double)f1 - (double)old);
    return ch;
}

// Gather work: every wavefront walks segments of kSeg groups (512 edge slots) of its
// unit on its own.  A segment's messages are one contiguous range of at most
// 512 + 8 entries (pairs never outnumber edge slots inside a tile, and only the first
// tile of a segment can start mid-tile); the wave stages them in its own LDS slice and
// adds them into the block's FP64 accumulators.  The next segment's ids and messages
// are loaded into registers before the current one is processed, and no block-wide
// barrier is taken between segments.
constexpr int kBG = 1024;                 // threads per gather workgroup
constexpr int kSeg = MG_WAVE;             // groups per wave segment: one per lane
constexpr int kStageM = (kSeg * kG + 16 + MG_WAVE - 1) / MG_WAVE;  // message loads per lane
constexpr int kStageW = kStageM * MG_WAVE;                            // staged messages per wave
constexpr int kStage = (kBG / MG_WAVE) * kStageW;           // staged messages per workgroup

// A segment's loads are split in two stages whose results are only used one
// iteration later, so the in-order wait counter never stalls on loads just issued.
struct SegHdr {
    uint4 id;     // this lane's 8 edge slots (lanes past the segment repeat its last group)
    int ga, gb;   // gbase at the segment's first group and one past its last
    int gl;       // gbase at this lane's group
    int pad;      // this lane is past the segment: its slots count as padding
};
struct Seg {
    uint4 id;
    int base;     // message index of the lane's first slot, relative to the staged range
    int n;        // staged messages
    float m[kStageM];
};

// unconditional loads only (no register is written after a load is issued)
__device__ __forceinline__ void load_hdr(SegHdr &h, int gs, int ge, const uint16_t *__restrict__ b_dst,
                                         const int *__restrict__ gbase) {
    const int g = min(gs + lane(), ge - 1);
    h.pad = gs + lane() >= ge;
    h.ga = gbase[gs];
    h.gb = gbase[ge];
    h.id = ld_stream_u4(&b_dst[(long long)g * kG]);
    h.gl = gbase[g];
}

// loads every staged message (indices past the range repeat the last one: no branches)
__device__ __forceinline__ void load_msg(Seg &sg, const SegHdr &h, const float *__restrict__ msg, long long npairs) {
    const long long q0 = h.ga;
    const long long q1 = min(npairs, (long long)h.gb + 1);
    sg.n = (int)(q1 - q0);
    const uint32_t pm = h.pad ? 0x7FFF7FFFu : 0u;
    sg.id = make_uint4(h.id.x | pm, h.id.y | pm, h.id.z | pm, h.id.w | pm);
    sg.base = h.pad ? 0 : h.gl - h.ga;
#pragma unroll
    for (int k = 0; k < kStageM; ++k) sg.m[k] = ld_stream(&msg[q0 + min(lane() + k * MG_WAVE, sg.n - 1)]);
}

constexpr int kPer = 16;  // vertices per gather thread: P <= kPer * kBG

// Hooks: the LDS accumulators start at initialize's identity, every edge adds its
// message with accumulate (LDS FP64 atomic add), the partition's two hubs sum in
// registers with accumulate_nonAtomics, the workgroup applies update_condition and
// writes the next messages from get_partial_result.
template <class H>
__global__ __launch_bounds__(kBG) void k_pb_gather(Ctl *ctl, const int4 *__restrict__ units, int nunits, int P, int nv, int np,
                                                   int variant, float damping, float tol,
                                                   const uint16_t *__restrict__ b_dst, const int *__restrict__ gbase,
                                                   float *__restrict__ msgbuf, long long npairs,
                                                   const int *__restrict__ apart, const uint16_t *__restrict__ a_src,
                                                   const int *__restrict__ amap,
                                                   const int *__restrict__ deg, float *__restrict__ rank,
                                                   int *__restrict__ active, double *__restrict__ psum) {
    if (ctl->mode == MODE_DONE) return;
    const float *msg = msg_in(ctl, msgbuf, npairs);
    float *out = msg_out(ctl, msgbuf, npairs);
    extern __shared__ double s_acc[];  // P + 1 FP64 accumulators, then kStageW staged messages per wave
    const int wid = threadIdx.x / MG_WAVE;
    float *s_msg = (float *)(s_acc + P + 1) + wid * kStageW;
    __shared__ double s_tmp[kBG / MG_WAVE];
    double res = 0.0;
    int ch = 0;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int4 un = units[u];
        const int v0 = un.x * P;
        const int pv = part_count(un.x, nv, np);
        for (int k = threadIdx.x; k <= P; k += kBG) H::initialize(&s_acc[k]);
        __syncthreads();
        const int g0 = un.y / kG, g1 = un.z / kG;
        const int nseg = (g1 - g0 + kSeg - 1) / kSeg;
        // three-stage pipeline over this wave's segments s, s + W, s + 2W: headers two
        // ahead, messages one ahead, so each dependent load has a segment's time to land
        constexpr int W = kBG / MG_WAVE;
        auto seg_lo = [&](int si) { return g0 + si * kSeg; };
        auto seg_hi = [&](int si) { return min(g1, g0 + (si + 1) * kSeg); };
        // Rotation over register sets, unrolled so no in-flight value is ever copied;
        // loads past the last segment repeat it (no exec-mask branches).
        auto hdr = [&](SegHdr &h, int si) {
            si = min(si, nseg - 1);
            load_hdr(h, seg_lo(si), seg_hi(si), b_dst, gbase);
        };
        double hub, hub1;
        H::initialize(&hub);
        H::initialize(&hub1);
        auto process = [&](const Seg &sg) {
            wave_sync();  // the previous segment's staged messages are consumed
#pragma unroll
            for (int k = 0; k < kStageM; ++k) s_msg[lane() + k * MG_WAVE] = sg.m[k];
            wave_sync();
            const uint32_t w[4] = {sg.id.x, sg.id.y, sg.id.z, sg.id.w};
            int pi = sg.base;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t v = (k & 1) ? (w[k >> 1] >> 16) : (w[k >> 1] & 0xFFFFu);
                if (k > 0) pi += (int)(v >> 15);
                const uint32_t dst = v & 0x7FFFu;
                const double m = (double)s_msg[pi];
                // local ids 0 and 1 are the partition's two highest-degree vertices: id 0
                // takes ~10% of the edges, and lanes of one instruction adding to one LDS
                // address serialise, so their sums stay in registers (0.4 % faster with
                // id 1 too, 2.5 % slower with ids 0..3); padding (0x7FFF) adds nothing
                if (dst == 0) H::accumulate_nonAtomics(&hub, m);
                else if (dst == 1) H::accumulate_nonAtomics(&hub1, m);
                else if (dst != 0x7FFFu) H::accumulate(&s_acc[dst], m);
            }
        };
        int sgi = wid;
        if (sgi < nseg) {
            // three message buffers: while one segment is processed, the messages of
            // the next two and the header of the third are in flight
            SegHdr h;
            Seg a, b, c;
            hdr(h, sgi);
            load_msg(a, h, msg, npairs);
            hdr(h, sgi + W);
            load_msg(b, h, msg, npairs);
            hdr(h, sgi + 2 * W);
            for (;;) {
                load_msg(c, h, msg, npairs);
                hdr(h, sgi + 3 * W);
                process(a);
                sgi += W;
                if (sgi >= nseg) break;
                load_msg(a, h, msg, npairs);
                hdr(h, sgi + 3 * W);
                process(b);
                sgi += W;
                if (sgi >= nseg) break;
                load_msg(b, h, msg, npairs);
                hdr(h, sgi + 3 * W);
                process(c);
                sgi += W;
                if (sgi >= nseg) break;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {  // the lanes' hub sums merged
            H::accumulate_nonAtomics(&hub, __shfl_xor(hub, o));
            H::accumulate_nonAtomics(&hub1, __shfl_xor(hub1, o));
        }
        double zero;
        H::initialize(&zero);
        if (lane() == 0 && hub != zero) H::accumulate(&s_acc[0], hub);
        if (lane() == 0 && hub1 != zero) H::accumulate(&s_acc[1], hub1);
        __syncthreads();
        if (un.w) {  // partition split over several units: add partials, applied later
            for (int k = threadIdx.x; k < pv; k += kBG) {
                const double a = s_acc[k];
                if (a != zero) H::accumulate(&psum[v0 + k], a);
            }
        } else {
            // apply, then write the partition's messages for the next iteration: the
            // new c = rank / degree is staged in LDS over the consumed accumulators
            float cl[kPer];
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int k = threadIdx.x + j * kBG;
                cl[j] = 0.f;
                if (k < pv) {
                    ch |= pr_apply<H>(v0 + k, s_acc[k], variant, damping, tol, deg, rank, active, &res);
                    H::get_partial_result(&cl[j], rank[v0 + k], deg[v0 + k]);
                }
            }
            __syncthreads();
            float *s_c = (float *)s_acc;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int k = threadIdx.x + j * kBG;
                if (k < pv) s_c[k] = cl[j];
            }
            __syncthreads();
            scatter_part<kBG>(apart[un.x], apart[un.x + 1], s_c, a_src, amap, out);
        }
        __syncthreads();
    }
    res = block_sum<kBG / MG_WAVE>(res, s_tmp);
    const int any = __syncthreads_or(ch);
    if (threadIdx.x == 0) {
        if (res != 0.0) atomicAdd(&ctl->resid, res);
        if (any) atomicOr(&ctl->changed, 1);
    }
}

// vertices of split partitions: apply from the FP64 partial sums and clear them
template <class H>
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
        ch |= pr_apply<H>(v, psum[v], variant, damping, tol, deg, rank, active, &res);
        H::initialize(&psum[v]);
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

// ---- SSSP sweeps over the layout ---------------------------------------------------
// The min-plus form of the gather: a message is its source's distance (one per
// (source, destination partition) pair), every edge slot adds its own weight, and a
// destination keeps the minimum in an LDS accumulator (native LDS atomicMin).  This is
// one dense all-edge relaxation (process.cu:819-934) with no random HBM access per
// edge: ids, weights and messages are streamed, distances are touched once per
// vertex.  Distances stay in engine ids (dist[inv[internal id]]); a vertex whose
// distance drops gets the level's stamp, as in every SSSP level.

// messages from the current distances: every source partition before a pull sweep
// that no pull sweep prepared (when = 0), or the partitions whose destination side was
// split, after the gather (when = 1); unit = {partition, a0, a1, 0}
// (force: rebuild every message whatever the last sweep prepared, after the width
// changed).  MT = uint16_t: 16-bit messages; a finite distance above 65534 is sent as
// 65535 (relaxes nothing), so its vertex is stamped with the level (it stays in the
// frontier) and ctl->sat tells the host to continue with 32-bit messages.
template <class MT>
__global__ __launch_bounds__(kBT) void k_ps_scatter(Ctl *ctl, int when, int force, const int4 *__restrict__ units, int nunits,
                                                    int P, int nv, int np, const int *__restrict__ inv,
                                                    const int *__restrict__ dist, int *__restrict__ vfc,
                                                    const uint16_t *__restrict__ a_src, const int *__restrict__ amap,
                                                    MT *__restrict__ msgbuf, long long npairs) {
    if (ctl->mode != MODE_BU) return;
    const int iter = ctl->iter;
    if (when == 0 && !force && ctl->msg_iter == iter) return;
    MT *out = msgbuf;
    extern __shared__ int s_dist[];
    bool sat = false;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int4 un = units[u];
        const int v0 = un.x * P;
        const int pv = part_count(un.x, nv, np);
        for (int k = threadIdx.x; k < pv; k += kBT) {
            const int v = inv[v0 + k];
            const int d = dist[v];
            s_dist[k] = d;
            if (sizeof(MT) == 2 && d != (int)MG_INF && d >= 0xFFFF) {
                vfc[v] = iter;
                sat = true;
            }
        }
        __syncthreads();
        if constexpr (sizeof(MT) == 2) scatter_part16<kBT>(un.y, un.z, s_dist, a_src, amap, out);
        else scatter_part<kBT, int>(un.y, un.z, s_dist, a_src, amap, out);
        __syncthreads();
    }
    if (sat) ctl->sat = 1;
}

// a segment of the SSSP gather: the PageRank segment plus the lane's 8 slot weights
struct SegW {
    uint4 id, wt;
    int base, n;
    int m[kStageM];
};

template <class MT>
__device__ __forceinline__ void load_msg_w(SegW &sg, const SegHdr &h, const uint4 &wt, const MT *__restrict__ msg,
                                           long long npairs) {
    const long long q0 = h.ga;
    const long long q1 = min(npairs, (long long)h.gb + 1);
    sg.n = (int)(q1 - q0);
    const uint32_t pm = h.pad ? 0x7FFF7FFFu : 0u;
    sg.id = make_uint4(h.id.x | pm, h.id.y | pm, h.id.z | pm, h.id.w | pm);
    sg.wt = wt;
    sg.base = h.pad ? 0 : h.gl - h.ga;
#pragma unroll
    for (int k = 0; k < kStageM; ++k) {
        if constexpr (sizeof(MT) == 2) {
            const uint32_t m = msg[q0 + min(lane() + k * MG_WAVE, sg.n - 1)];
            sg.m[k] = m == 0xFFFFu ? (int)MG_INF : (int)m;
        } else {
            sg.m[k] = ld_stream(&msg[q0 + min(lane() + k * MG_WAVE, sg.n - 1)]);
        }
    }
}

// WT: uint8_t weights when every weight fits a byte (a slot's 8 weights are one 8-byte
// load), else uint16_t (one 16-byte load); MT: the message width (see k_ps_scatter)
// Hooks: LDS minima start at initialize's identity; an edge's candidate is
// get_partial_result(message, weight), combined by accumulate into the LDS slot (LDS
// atomicMin) or by accumulate_nonAtomics into the hub's register; a vertex applies
// with update_condition, or fused_update_condition when its partition is split over
// several workgroups.
template <class H, class WT, class MT>
__global__ __launch_bounds__(kBG) void k_ps_gather(Ctl *ctl, const int4 *__restrict__ units, int nunits, int P, int nv, int np,
                                                   const uint16_t *__restrict__ b_dst, const WT *__restrict__ b_w,
                                                   const int *__restrict__ gbase, MT *__restrict__ msgbuf, long long npairs,
                                                   const int *__restrict__ apart, const uint16_t *__restrict__ a_src,
                                                   const int *__restrict__ amap, const int *__restrict__ inv,
                                                   int *__restrict__ dist, int *__restrict__ vfc) {
    if (ctl->mode != MODE_BU) return;
    const int iter = ctl->iter;
    // One message buffer: a partition's new distances are written over the messages
    // this sweep reads, so partitions gathered later already relax from them.  Any mix
    // of old and new messages is an upper bound of the true distances, so the result
    // is exact whatever the order; only the number of sweeps depends on it.
    const MT *msg = msgbuf;
    MT *out = msgbuf;
    bool sat = false;
    extern __shared__ int s_min[];  // P + 1 minima, then kStageW staged messages per wave
    const int wid = threadIdx.x / MG_WAVE;
    int *s_msg = s_min + P + 1 + wid * kStageW;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int4 un = units[u];
        const int v0 = un.x * P;
        const int pv = part_count(un.x, nv, np);
        for (int k = threadIdx.x; k <= P; k += kBG) H::initialize(&s_min[k]);
        __syncthreads();
        const int g0 = un.y / kG, g1 = un.z / kG;
        const int nseg = (g1 - g0 + kSeg - 1) / kSeg;
        constexpr int W = kBG / MG_WAVE;
        auto hdr = [&](SegHdr &h, uint4 &wt, int si) {
            si = min(si, nseg - 1);
            const int gs = g0 + si * kSeg, ge = min(g1, g0 + (si + 1) * kSeg);
            load_hdr(h, gs, ge, b_dst, gbase);
            const long long wg = (long long)min(gs + lane(), ge - 1) * kG;
            if (sizeof(WT) == 1) {
                const uint2 x = ld_stream_u2(&b_w[wg]);
                wt = make_uint4(x.x, x.y, 0u, 0u);
            } else {
                wt = ld_stream_u4(&b_w[wg]);
            }
        };
        int hub;
        H::initialize(&hub);
        auto process = [&](const SegW &sg) {
            wave_sync();
#pragma unroll
            for (int k = 0; k < kStageM; ++k) s_msg[lane() + k * MG_WAVE] = sg.m[k];
            wave_sync();
            const uint32_t w[4] = {sg.id.x, sg.id.y, sg.id.z, sg.id.w};
            const uint32_t ww[4] = {sg.wt.x, sg.wt.y, sg.wt.z, sg.wt.w};
            int pi = sg.base;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t v = (k & 1) ? (w[k >> 1] >> 16) : (w[k >> 1] & 0xFFFFu);
                const int wt = sizeof(WT) == 1 ? (int)((ww[k >> 2] >> (8 * (k & 3))) & 0xFFu)
                                              : (int)((k & 1) ? (ww[k >> 1] >> 16) : (ww[k >> 1] & 0xFFFFu));
                if (k > 0) pi += (int)(v >> 15);
                const uint32_t dst = v & 0x7FFFu;
                const int m = s_msg[pi];
                // local id 0 is the partition's highest-degree vertex: lanes of one
                // instruction updating one LDS address serialise, so its minimum stays in
                // a register; padding (0x7FFF) and unreached sources relax nothing
                if (dst == 0x7FFFu || m == (int)MG_INF) continue;
                int cand;
                H::get_partial_result(&cand, m, wt);
                if (dst == 0) H::accumulate_nonAtomics(&hub, cand, iter);
                else H::accumulate(&s_min[dst], cand, iter);
            }
        };
        int sgi = wid;
        if (sgi < nseg) {
            SegHdr h;
            uint4 wt;
            SegW a, b, c;
            hdr(h, wt, sgi);
            load_msg_w(a, h, wt, msg, npairs);
            hdr(h, wt, sgi + W);
            load_msg_w(b, h, wt, msg, npairs);
            hdr(h, wt, sgi + 2 * W);
            for (;;) {
                load_msg_w(c, h, wt, msg, npairs);
                hdr(h, wt, sgi + 3 * W);
                process(a);
                sgi += W;
                if (sgi >= nseg) break;
                load_msg_w(a, h, wt, msg, npairs);
                hdr(h, wt, sgi + 3 * W);
                process(b);
                sgi += W;
                if (sgi >= nseg) break;
                load_msg_w(b, h, wt, msg, npairs);
                hdr(h, wt, sgi + 3 * W);
                process(c);
                sgi += W;
                if (sgi >= nseg) break;
            }
        }
        Acc<H> ident;
        acc_init(ident);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) H::accumulate_nonAtomics(&hub, __shfl_xor(hub, o), iter);
        if (lane() == 0 && hub != ident.v) H::accumulate(&s_min[0], hub, iter);
        __syncthreads();
        if (un.w) {  // partition split over several units: each lowers what it found
            for (int k = threadIdx.x; k < pv; k += kBG) {
                Acc<H> a = ident;
                a.v = s_min[k];
                if (a.v == ident.v) continue;
                const int v = inv[v0 + k];
                acc_apply_atomic<H>(&dist[v], nullptr, &vfc[v], a, iter);
            }
        } else {
            // apply, then write the partition's messages for the next sweep: its
            // distances are staged in LDS over the consumed minima
            int cl[kPer];
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int k = threadIdx.x + j * kBG;
                cl[j] = MG_INF;
                if (k < pv) {
                    const int v = inv[v0 + k];
                    Acc<H> a = ident;
                    a.v = s_min[k];
                    int d = dist[v];
                    if (acc_apply<H>(d, &dist[v], nullptr, &vfc[v], a, iter)) d = a.v;
                    if (sizeof(MT) == 2 && d != (int)MG_INF && d >= 0xFFFF) {  // cannot be sent: see k_ps_scatter
                        vfc[v] = iter;
                        sat = true;
                    }
                    cl[j] = d;
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int k = threadIdx.x + j * kBG;
                if (k < pv) s_min[k] = cl[j];
            }
            __syncthreads();
            if constexpr (sizeof(MT) == 2) scatter_part16<kBG>(apart[un.x], apart[un.x + 1], s_min, a_src, amap, out);
            else scatter_part<kBG, int>(apart[un.x], apart[un.x + 1], s_min, a_src, amap, out);
        }
        __syncthreads();
    }
    if (sat) ctl->sat = 1;
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->msg_iter = iter + 1;
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
    // balanced and np can be any count.  Larger partitions share more messages (one per
    // source vertex and destination partition), so np is the smallest multiple of the
    // CU count whose partitions fit the gather's LDS (FP64 accumulators plus staged
    // messages).  part_log2 > 0 forces P = 2^part_log2 instead.
    int cus = 256;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    const long long pmax = (160 * 1024 - 4 * kStage - 2048) / 8 - 1;
    int np, P;
    if (e->opt.part_log2 > 0) {
        P = (int)std::min<long long>(1ll << std::min(e->opt.part_log2, 20), pmax);
        np = nv > 0 ? (nv + P - 1) / P : 1;
    } else if (nv < cus) {
        np = 1;
        P = std::max(1, nv);
    } else {
        const long long k = ((long long)nv + (long long)cus * pmax - 1) / ((long long)cus * pmax);
        np = (int)((long long)cus * std::max(1ll, k));
        P = (nv + np - 1) / np;
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
    if (e->algo == MG_SSSP) {  // SSSP keeps distances in engine ids: internal id -> compacted id
        std::vector<int> inv((size_t)nvi, 0);
        for (int v = 0; v < nv; ++v) inv[e->pr_perm_h[v]] = v;
        MG_HIP(hipMalloc(&L.inv, sizeof(int) * std::max<long long>(1, nvi)));
        MG_HIP(hipMemcpy(L.inv, inv.data(), sizeof(int) * nvi, hipMemcpyHostToDevice));
    }
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
    int *count = nullptr, *cursor = nullptr, *npair = nullptr;
    long long *pb = nullptr, *pa = nullptr, *boff = nullptr, *aoff = nullptr, *poffB = nullptr;
    uint32_t *keys = nullptr;
    auto cleanup = [&]() {
        hipFree(d_items); hipFree(count); hipFree(cursor); hipFree(npair); hipFree(pb); hipFree(pa); hipFree(boff);
        hipFree(aoff); hipFree(poffB); hipFree(keys);
        d_items = nullptr; count = cursor = npair = nullptr; pb = pa = boff = aoff = poffB = nullptr; keys = nullptr;
    };
    if (hipMalloc(&d_items, sizeof(BItem) * std::max<size_t>(1, items.size())) != hipSuccess ||
        hipMalloc(&count, sizeof(int) * nt) != hipSuccess || hipMalloc(&cursor, sizeof(int) * nt) != hipSuccess ||
        hipMalloc(&npair, sizeof(int) * nt) != hipSuccess ||
        hipMalloc(&pb, sizeof(long long) * nt) != hipSuccess || hipMalloc(&pa, sizeof(long long) * nt) != hipSuccess ||
        hipMalloc(&boff, sizeof(long long) * (nt + 1)) != hipSuccess || hipMalloc(&aoff, sizeof(long long) * (nt + 1)) != hipSuccess ||
        hipMalloc(&poffB, sizeof(long long) * (nt + 1)) != hipSuccess) {
        cleanup();
        return mg_fail(MG_ERR_NOMEM, "partitioned layout: tile tables do not fit");
    }
    MG_HIP(hipMemcpy(d_items, items.data(), sizeof(BItem) * items.size(), hipMemcpyHostToDevice));
    MG_HIP(hipMemsetAsync(count, 0, sizeof(int) * nt, s));
    MG_HIP(hipMemsetAsync(cursor, 0, sizeof(int) * nt, s));
    const int nitems = (int)items.size();
    const size_t lds_count = sizeof(int) * np;
    const size_t lds_place = sizeof(int) * (np + P + 1);
    const size_t lds_gather = sizeof(double) * (P + 1) + sizeof(float) * kStage;
    if (lds_place > 160 * 1024 || lds_gather > 160 * 1024 - 1024) {
        cleanup();
        return mg_fail(MG_ERR_LIMIT, "partitioned layout: %d partitions of %d ids exceed LDS", np, P);
    }
    // dynamic LDS above the 64 KiB default needs an explicit per-kernel limit
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_count, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(lds_count, 1024)));
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_place, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(lds_place, 1024)));
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_scatter_msg<MG_PR_HOOKS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(float) * P)));
    MG_HIP(hipFuncSetAttribute((const void *)k_pb_gather<MG_PR_HOOKS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_gather));
    // edges per tile -> padded edge slots -> tile offsets of the destination-major stream
    hipLaunchKernelGGL(k_pb_count, dim3(grid_for(nitems, 1, 4096)), dim3(kBT), lds_count, s, icol, irow, d_items, nitems,
                       L.P, np, count);
    hipLaunchKernelGGL(k_pb_pad, dim3(grid_for(nt, 256, 8192)), dim3(256), 0, s, count, nt, pb);
    int rc = mg_scan64(s, pb, nullptr, nt, boff);
    if (rc != MG_OK) { cleanup(); return rc; }
    long long slots = 0;
    MG_HIP(hipMemcpy(&slots, boff + nt, sizeof(long long), hipMemcpyDeviceToHost));
    // slot indices are int32 in the units and kernels; a margin of 2^24 keeps every
    // segment bound (slot + a few hundred) inside the type
    if (slots > (1ll << 31) - (1ll << 24)) {
        cleanup();
        return mg_fail(MG_ERR_LIMIT, "partitioned layout: %lld slots exceed the int32 range", slots);
    }
    L.slots = slots;
    const long long sl = slots > 0 ? slots : kG;
    if (hipMalloc(&keys, sizeof(uint32_t) * sl) != hipSuccess) { cleanup(); return mg_fail(MG_ERR_NOMEM, "partitioned layout: keys"); }
    hipLaunchKernelGGL(k_pb_place, dim3(grid_for(nitems, 1, 4096)), dim3(kBT), lds_place, s, icol, irow, d_items, nitems,
                       L.P, np, cursor, boff, keys);
    hipLaunchKernelGGL(k_pb_tile_sort, dim3(grid_for(nt, 1, 8192)), dim3(kBT), 0, s, nt, count, boff, keys, npair);
    MG_HIP(hipGetLastError());
    // tiles above the LDS sort size (only graphs with very few partitions): sort on the host
    {
        std::vector<int> hn(nt), hc(nt);
        std::vector<long long> hb(nt + 1);
        MG_HIP(hipMemcpyAsync(hn.data(), npair, sizeof(int) * nt, hipMemcpyDeviceToHost, s));
        MG_HIP(hipMemcpyAsync(hc.data(), count, sizeof(int) * nt, hipMemcpyDeviceToHost, s));
        MG_HIP(hipMemcpyAsync(hb.data(), boff, sizeof(long long) * (nt + 1), hipMemcpyDeviceToHost, s));
        MG_HIP(hipStreamSynchronize(s));
        bool fixed = false;
        std::vector<uint32_t> tk;
        for (long long t = 0; t < nt; ++t) {
            if (hn[t] >= 0) continue;
            tk.resize(hc[t]);
            MG_HIP(hipMemcpy(tk.data(), keys + hb[t], sizeof(uint32_t) * hc[t], hipMemcpyDeviceToHost));
            std::sort(tk.begin(), tk.end());
            int c = 0;
            for (int i = 0; i < hc[t]; ++i) c += (i == 0 || (tk[i] >> 16) != (tk[i - 1] >> 16)) ? 1 : 0;
            MG_HIP(hipMemcpy(keys + hb[t], tk.data(), sizeof(uint32_t) * hc[t], hipMemcpyHostToDevice));
            hn[t] = c;
            fixed = true;
        }
        if (fixed) MG_HIP(hipMemcpy(npair, hn.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
    }
    if (e->algo == MG_SSSP) {  // per edge slot weights, from the sorted keys (padding: 0)
        const size_t wb = L.w8 ? 1 : 2;
        if (hipMalloc(&L.b_w, wb * sl + 16) != hipSuccess) { cleanup(); return mg_fail(MG_ERR_NOMEM, "partitioned layout: weights"); }
        MG_HIP(hipMemsetAsync(L.b_w, 0, wb * sl + 16, s));
        if (L.w8)
            hipLaunchKernelGGL(k_pb_weights<uint8_t>, dim3(grid_for(nt, 1, 8192)), dim3(kBT), 0, s, nt, np, P, count, boff, keys, L.inv,
                               e->row, e->col, e->w, (uint8_t *)L.b_w);
        else
            hipLaunchKernelGGL(k_pb_weights<uint16_t>, dim3(grid_for(nt, 1, 8192)), dim3(kBT), 0, s, nt, np, P, count, boff, keys,
                               L.inv, e->row, e->col, e->w, (uint16_t *)L.b_w);
        MG_HIP(hipGetLastError());
    }
    // messages per tile -> padded message slots in both orders -> offsets
    hipLaunchKernelGGL(k_pb_pad_pairs, dim3(grid_for(nt, 256, 8192)), dim3(256), 0, s, npair, np, pb, pa);
    rc = mg_scan64(s, pb, nullptr, nt, poffB);
    if (rc == MG_OK) rc = mg_scan64(s, pa, nullptr, nt, aoff);
    if (rc != MG_OK) { cleanup(); return rc; }
    long long pairs = 0;
    MG_HIP(hipMemcpy(&pairs, poffB + nt, sizeof(long long), hipMemcpyDeviceToHost));
    L.pairs = pairs;
    const long long pl = pairs > 0 ? pairs : kG;
    const long long ngroups = sl / kG;
    if (hipMalloc(&L.a_src, sizeof(uint16_t) * pl) != hipSuccess || hipMalloc(&L.amap, sizeof(int) * (pl / kG + 1)) != hipSuccess ||
        hipMalloc(&L.msg, sizeof(float) * 2 * pl) != hipSuccess || hipMalloc(&L.b_dst, sizeof(uint16_t) * sl) != hipSuccess ||
        hipMalloc(&L.gbase, sizeof(int) * (ngroups + 1)) != hipSuccess) {
        cleanup();
        return mg_fail(MG_ERR_NOMEM, "partitioned layout: %lld edge slots / %lld messages do not fit", slots, pairs);
    }
    hipLaunchKernelGGL(k_fill_u16, dim3(grid_for(pl, 256, 8192)), dim3(256), 0, s, L.a_src, pl, (uint16_t)0xFFFF);
    hipLaunchKernelGGL(k_fill_u16, dim3(grid_for(sl, 256, 8192)), dim3(256), 0, s, L.b_dst, sl, (uint16_t)0x7FFF);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(pl / kG + 1, 256, 8192)), dim3(256), 0, s, (uint32_t *)L.amap, pl / kG + 1, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2 * pl, 256, 8192)), dim3(256), 0, s, (uint32_t *)L.msg, 2 * pl, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ngroups + 1, 256, 8192)), dim3(256), 0, s, (uint32_t *)L.gbase, ngroups + 1,
                       (uint32_t)pairs);
    hipLaunchKernelGGL(k_pb_emit, dim3(grid_for(nt, 1, 8192)), dim3(kBT), 0, s, nt, np, count, boff, poffB, aoff, keys,
                       L.b_dst, L.gbase, L.a_src, L.amap);
    MG_HIP(hipGetLastError());
    // partition starts of both streams -> work units of about equal slot counts
    std::vector<long long> hb(nt + 1), ha(nt + 1);
    MG_HIP(hipMemcpy(hb.data(), boff, sizeof(long long) * (nt + 1), hipMemcpyDeviceToHost));
    MG_HIP(hipMemcpy(ha.data(), aoff, sizeof(long long) * (nt + 1), hipMemcpyDeviceToHost));
    MG_HIP(hipStreamSynchronize(s));
    cleanup();
    // unit bounds stay multiples of the 16-slot group (vector loads of 8 slots)
    const long long unit = ((std::max<long long>(32768, 2 * (slots / std::max(1, np)) + 1) + kG - 1) / kG) * kG;
    // gather units cut the destination-major edge stream; a partition cut into several
    // units is applied, and its messages written, by separate passes (units A)
    std::vector<int4> ua, ub;
    std::vector<int> split, apart(np + 1);
    for (int p = 0; p <= np; ++p) apart[p] = (int)ha[(long long)p * np];
    for (int p = 0; p < np; ++p) {
        const long long b0 = hb[(long long)p * np], b1 = hb[(long long)(p + 1) * np];
        const int nb = (int)((b1 - b0 + unit - 1) / unit);
        if (nb <= 1) ub.push_back(make_int4(p, (int)b0, (int)b1, 0));
        else {
            split.push_back(p);
            for (long long x = b0; x < b1; x += unit) ub.push_back(make_int4(p, (int)x, (int)std::min(b1, x + unit), 1));
            const long long a0 = apart[p], a1 = apart[p + 1];
            for (long long x = a0; x < a1; x += unit) ua.push_back(make_int4(p, (int)x, (int)std::min(a1, x + unit), 0));
        }
    }
    L.nA = (int)ua.size();
    L.nB = (int)ub.size();
    L.nApply = (int)split.size();
    if (e->algo == MG_SSSP) {  // every source partition, in units: messages rebuilt from distances
        std::vector<int4> us;
        for (int p = 0; p < np; ++p)
            for (long long x = apart[p]; x < apart[p + 1]; x += unit) us.push_back(make_int4(p, (int)x, (int)std::min<long long>(apart[p + 1], x + unit), 0));
        L.nS = (int)us.size();
        MG_HIP(hipMalloc(&L.unitsS, sizeof(int4) * std::max<size_t>(1, us.size())));
        if (!us.empty()) MG_HIP(hipMemcpy(L.unitsS, us.data(), sizeof(int4) * us.size(), hipMemcpyHostToDevice));
        const int lds_sc = (int)(sizeof(int) * P);
        MG_HIP(hipFuncSetAttribute((const void *)k_ps_scatter<int>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_sc));
        MG_HIP(hipFuncSetAttribute((const void *)k_ps_scatter<uint16_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_sc));
        const int lds_ps = (int)(sizeof(int) * (P + 1) + sizeof(int) * kStage);
        MG_HIP(hipFuncSetAttribute((const void *)k_ps_gather<MG_SSSP_HOOKS, uint8_t, int>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_ps));
        MG_HIP(hipFuncSetAttribute((const void *)k_ps_gather<MG_SSSP_HOOKS, uint16_t, int>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_ps));
        MG_HIP(hipFuncSetAttribute((const void *)k_ps_gather<MG_SSSP_HOOKS, uint8_t, uint16_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_ps));
        MG_HIP(hipFuncSetAttribute((const void *)k_ps_gather<MG_SSSP_HOOKS, uint16_t, uint16_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_ps));
    }
    MG_HIP(hipMalloc(&L.unitsA, sizeof(int4) * std::max<size_t>(1, ua.size())));
    MG_HIP(hipMalloc(&L.unitsB, sizeof(int4) * std::max<size_t>(1, ub.size())));
    MG_HIP(hipMalloc(&L.split_parts, sizeof(int) * std::max<size_t>(1, split.size())));
    MG_HIP(hipMalloc(&L.apart, sizeof(int) * (np + 1)));
    MG_HIP(hipMemcpy(L.apart, apart.data(), sizeof(int) * (np + 1), hipMemcpyHostToDevice));
    if (!ua.empty()) MG_HIP(hipMemcpy(L.unitsA, ua.data(), sizeof(int4) * ua.size(), hipMemcpyHostToDevice));
    if (!ub.empty()) MG_HIP(hipMemcpy(L.unitsB, ub.data(), sizeof(int4) * ub.size(), hipMemcpyHostToDevice));
    if (!split.empty()) MG_HIP(hipMemcpy(L.split_parts, split.data(), sizeof(int) * split.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pb_zero_f64, dim3(grid_for((long long)nvi + 64, 256, 4096)), dim3(256), 0, s, e->psum, (long long)nvi + 64);
    MG_HIP(hipStreamSynchronize(s));
    return MG_OK;
}

// one dense iteration over the partitioned layout (without the control kernel): the
// gather of every unsplit partition also writes its messages for the next iteration
int mg_pb_iteration(mg_engine *e) {
    mg_pb_layout &L = e->pb;
    hipStream_t s = e->stream;
    const int variant = e->algo == MG_PR_T ? 0 : 1;
    const int P = L.P;
    const int G = e->opt.grid_blocks > 0 ? e->opt.grid_blocks : 2048;
    int slot = -1;
    // algorithmic bytes of one iteration (pull formulation): 8E + 16V
    const double alg = 8.0 * (double)e->ne + 16.0 * e->nv;
    if (e->prof) mg_prof_begin(e, "k_pb_gather", alg, &slot);
    hipLaunchKernelGGL(k_pb_gather<MG_PR_HOOKS>, dim3(grid_for(L.nB, 1, G)), dim3(kBG), sizeof(double) * (P + 1) + sizeof(float) * kStage, s,
                       e->ctl, L.unitsB, L.nB, L.P, e->nv, L.np, variant, e->opt.pr_damping, e->opt.pr_tolerance, L.b_dst,
                       L.gbase, L.msg, L.pairs, L.apart, L.a_src, L.amap, e->deg, e->rank, e->vfc, e->psum);
    if (e->prof) mg_prof_end(e, slot);
    if (L.nApply > 0) {
        hipLaunchKernelGGL(k_pb_apply_split<MG_PR_HOOKS>, dim3(grid_for((long long)L.nApply * P, kBT, G)), dim3(kBT), 0, s, e->ctl,
                           L.split_parts, L.nApply, L.P, e->nv, L.np, variant, e->opt.pr_damping, e->opt.pr_tolerance, e->deg,
                           e->rank, e->vfc, e->psum);
        hipLaunchKernelGGL(k_pb_scatter_msg<MG_PR_HOOKS>, dim3(grid_for(L.nA, 1, G)), dim3(kBT), sizeof(float) * P, s, e->ctl, L.unitsA,
                           L.nA, L.P, e->nv, L.np, e->rank, e->deg, L.a_src, L.amap, L.msg, L.pairs);
    }
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// one SSSP pull sweep over the layout, gated on the control block (a push level runs
// none of it): messages from distances unless the last sweep wrote them, the min-plus
// gather (unsplit partitions also write their next messages), then the messages of
// split partitions
int mg_ps_sweep(mg_engine *e) {
    mg_pb_layout &L = e->pb;
    hipStream_t s = e->stream;
    const int P = L.P;
    const int G = e->opt.grid_blocks > 0 ? e->opt.grid_blocks : 2048;
    const int force = e->pb_force ? 1 : 0;
    e->pb_force = false;
    int slot = -1;
    // one full relaxation pass: 12E + 8V (+4)
    const double alg = 12.0 * (double)e->ne + 8.0 * e->nv + 4.0;
    const size_t lds = sizeof(int) * (P + 1) + sizeof(int) * kStage;
    auto sweep = [&](auto *msg) {
        using MT = std::remove_pointer_t<decltype(msg)>;
        if (L.nS > 0)
            hipLaunchKernelGGL(k_ps_scatter<MT>, dim3(grid_for(L.nS, 1, G)), dim3(kBT), sizeof(int) * P, s, e->ctl, 0, force, L.unitsS,
                               L.nS, P, e->nv, L.np, L.inv, e->level, e->vfc, L.a_src, L.amap, msg, L.pairs);
        if (e->prof) mg_prof_begin(e, "k_ps_gather", alg, &slot);
        if (L.w8)
            hipLaunchKernelGGL((k_ps_gather<MG_SSSP_HOOKS, uint8_t, MT>), dim3(grid_for(L.nB, 1, G)), dim3(kBG), lds, s, e->ctl, L.unitsB, L.nB, P,
                               e->nv, L.np, L.b_dst, (const uint8_t *)L.b_w, L.gbase, msg, L.pairs, L.apart, L.a_src, L.amap, L.inv,
                               e->level, e->vfc);
        else
            hipLaunchKernelGGL((k_ps_gather<MG_SSSP_HOOKS, uint16_t, MT>), dim3(grid_for(L.nB, 1, G)), dim3(kBG), lds, s, e->ctl, L.unitsB, L.nB, P,
                               e->nv, L.np, L.b_dst, (const uint16_t *)L.b_w, L.gbase, msg, L.pairs, L.apart, L.a_src, L.amap, L.inv,
                               e->level, e->vfc);
        if (e->prof) mg_prof_end(e, slot);
        if (L.nA > 0)
            hipLaunchKernelGGL(k_ps_scatter<MT>, dim3(grid_for(L.nA, 1, G)), dim3(kBT), sizeof(int) * P, s, e->ctl, 1, 0, L.unitsA,
                               L.nA, P, e->nv, L.np, L.inv, e->level, e->vfc, L.a_src, L.amap, msg, L.pairs);
    };
    // 16-bit messages live in the second half of the message allocation (PageRank's
    // second buffer; pairs is a multiple of 8, so it starts 32-byte aligned)
    if (e->pb16) sweep((uint16_t *)((float *)L.msg + std::max<long long>(L.pairs, kG)));
    else sweep((int *)L.msg);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// before the first iteration: its messages (rank 0) are all zero
int mg_pb_reset(mg_engine *e) {
    mg_pb_layout &L = e->pb;
    const long long n = 2 * std::max<long long>(L.pairs, kG);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, e->stream, (uint32_t *)L.msg, n, 0u);
    MG_HIP(hipGetLastError());
    return MG_OK;
}
