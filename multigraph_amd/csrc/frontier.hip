// Frontier engine for BFS, SSSP and the BC forward pass (the role of process.cu) and
// the BC backward pass (the role of process2.cu), hand-written for gfx950.
//
// Per level the engine runs
//   * a push level (sparse queue): expand + contract.  Small frontiers (at most
//     16K edges) go through one kernel, one workgroup per 256 frontier vertices with a
//     block-wide exclusive scan of their degrees.  Larger frontiers are edge-balanced
//     across the whole grid: a tiled scan of the frontier's degrees, the owner of the
//     first edge of every 2048-edge chunk, then one workgroup per chunk, so a hub's
//     edges spread over many workgroups and no workgroup owns more than 2048 edges;
//     winners are appended with one atomic per wave (ballot + mbcnt);
//   * or a pull level (dense bitmap): one wave per 64 vertices, the frontier bitmap
//     word of a vertex's neighbour is tested per edge, and the wave's 64 results are
//     the next frontier word itself (ballot), so no atomics build the bitmap;
//   * then a one-thread control kernel that chooses push or pull for the next level
//     from frontier degree sums (deterministic, replacing the reference's wall-clock
//     sampling, process.cu:1161-1255).
// All kernels are grid-stride and read their sizes from the device control block,
// so a run of levels can be captured in a hipGraph and replayed with no host sync.
#include <chrono>
#include <mutex>

#include "engine.hpp"
#include "hook_calls.hpp"
#include "lb.hpp"

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
// A source of more than kSmallMF edges is pushed by the edge-balanced kernels; its
// degree scan is trivial (one entry at offset 0, every chunk owned by it), so the seed
// writes the scan's tables and the first level skips the scan kernels.
__global__ __launch_bounds__(256) void k_seed(Ctl *ctl, int *state, int *q, double *sigma, int src, int deg_src, long long ne,
                                              int out_begin, int *eoff, int *tsum, int *cstart) {
    const Ctl c = seed_ctl(deg_src, ne, out_begin);
    if (c.lb)
        for (int k = threadIdx.x; k <= c.lb_chunks; k += blockDim.x) cstart[k] = 0;
    if (threadIdx.x != 0) return;
    state[src] = 0;
    if (sigma) sigma[src] = 1.0;
    q[0] = src;
    if (c.lb) { eoff[0] = 0; tsum[0] = 0; }
    *ctl = c;
}

// mirror: the host's pinned copy of the control block (eager levels).  It gets the new
// block, then, after a system-scope fence, its sequence number: the host spins on that
// number instead of copying the block back and blocking on the stream.
// A pull level leaves its frontier's size and degree sum as one partial per workgroup
// (nparts of them; 0 after a push level): the control kernel's workgroup totals them
// first, so a pull level needs no reduction kernel of its own.
__global__ __launch_bounds__(256) void k_ctl(Ctl *ctl, Params p, int *lvl_off, long long *lvl_mf, Ctl *mirror, int seq,
                                             const int *__restrict__ part_n, const long long *__restrict__ part_mf,
                                             int nparts) {
    if (nparts > 0 && ctl->mode == MODE_BU) {
        __shared__ long long s_tmp[256 / MG_WAVE];
        long long n = 0, m = 0;
        for (int i = threadIdx.x; i < nparts; i += blockDim.x) { n += part_n[i]; m += part_mf[i]; }
        n = block_sum<256 / MG_WAVE>(n, s_tmp);
        m = block_sum<256 / MG_WAVE>(m, s_tmp);
        if (threadIdx.x == 0) { ctl->out_n = (int)n; ctl->out_mf = m; }
    }
    if (threadIdx.x != 0) return;
    Ctl c = *ctl;
    if (c.mode != MODE_DONE) {
        if (lvl_off) {
            lvl_off[c.iter] = c.out_begin;
            lvl_off[c.iter + 1] = c.out_begin + c.out_n;
            lvl_mf[c.iter] = c.out_mf;  // degree sum of the level (the BC backward pass picks its direction by it)
        }
        c.edges += c.mf;
        advance(c, p);
        *ctl = c;
    }
    if (mirror) {
        c.seq = seq - 1;
        *mirror = c;
        __threadfence_system();
        __hip_atomic_store(&mirror->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Appends of one step of a whole workgroup (every thread of it calls this, in step):
// the winners' slots in the next queue come from ONE counter atomic per workgroup and
// step, their degree sum from one more.  Atomics per wave on the shared counter
// serialised a hub's push level (RMAT-22, 163 K winners: 59 -> 20 us).
template <int B, bool kTrackDeg, class OT>
__device__ __forceinline__ void block_append(Ctl *ctl, int *q, int out_begin, const OT *row, bool win, int u,
                                             int *s_wc, long long *s_wd) {
    constexpr int NW = B / MG_WAVE;
    const uint64_t m = ballot(win);
    const int wv = threadIdx.x / MG_WAVE;
    long long d = 0;
    if (kTrackDeg) d = wave_sum(win ? (long long)(row[u + 1] - row[u]) : 0ll);
    if (lane() == 0) {
        s_wc[wv] = __builtin_popcountll(m);
        s_wd[wv] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        long long dt = 0;
        for (int i = 0; i < NW; ++i) {
            const int c = s_wc[i];
            s_wc[i] = tot;
            tot += c;
            dt += s_wd[i];
        }
        s_wc[NW] = tot ? atomicAdd(&ctl->out_n, tot) : 0;
        if (kTrackDeg && dt) atomicAdd((unsigned long long *)&ctl->out_mf, (unsigned long long)dt);
    }
    __syncthreads();
    if (win) q[out_begin + s_wc[NW] + s_wc[wv] + mbcnt(m)] = u;
    __syncthreads();  // s_wc is refilled by the next step
}

// Push level: expand the frontier's edges, contract winners into the next queue.
template <class H, bool kTrackDeg, class OT = int>
__global__ __launch_bounds__(kBlock) void k_push(Ctl *ctl, const Params P, const OT *__restrict__ row,
                                                 const int *__restrict__ col, const int *__restrict__ wgt,
                                                 int *__restrict__ state, int *__restrict__ vfc,
                                                 double *__restrict__ sigma, int *__restrict__ q,
                                                 HugeChunk *__restrict__ chunks) {
    if (ctl->mode != MODE_TD || ctl->lb) return;
    const int iter = ctl->iter, in_begin = ctl->in_begin, n = ctl->in_n, out_begin = ctl->out_begin;
    __shared__ int s_off[kBlock];
    __shared__ OT s_rb[kBlock];
    __shared__ int s_val[kBlock];
    __shared__ double s_val2[kBlock];
    __shared__ int s_tmp[kBlock / MG_WAVE + 1];
    __shared__ int s_wc[kBlock / MG_WAVE + 1];
    __shared__ long long s_wd[kBlock / MG_WAVE];
    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const int i = base + threadIdx.x;
        int v = -1, deg = 0, val = 0;
        OT rb = 0;
        double val2 = 0.0;
        if (i < n) {
            v = q[in_begin + i];
            rb = row[v];
            deg = (int)(row[v + 1] - rb);
            val = state[v];
            if (sigma) val2 = sigma[v];
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
                const OT p = s_rb[lo] + (e - s_off[lo]);
                u = ld_stream(&col[p]);
                win = relax<H>(s_val[lo], s_val2[lo], u, p, wgt, state, vfc, sigma, iter);
            }
            block_append<kBlock, kTrackDeg>(ctl, q, out_begin, row, win != 0, u, s_wc, s_wd);
        }
        __syncthreads();
    }
}

// ---- edge-balanced push (frontiers above kSmallMF edges) --------------------------
// 1. degrees of the frontier, exclusive scan inside tiles of kTile entries
template <class OT = int>
__global__ __launch_bounds__(kBlock) void k_lb_deg(Ctl *ctl, const OT *__restrict__ row, const int *__restrict__ q,
                                                   int *__restrict__ eoff, int *__restrict__ tsum) {
    if (ctl->mode != MODE_TD || !ctl->lb) return;
    const int n = ctl->in_n, in_begin = ctl->in_begin;
    __shared__ int s_tmp[kBlock / MG_WAVE + 1];
    constexpr int kPer = kTile / kBlock;
    for (int t = blockIdx.x; t * kTile < n; t += gridDim.x) {
        const int i0 = t * kTile + threadIdx.x * kPer;
        int d[kPer];
        int local = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = i0 + k;
            d[k] = 0;
            if (i < n) { const int v = q[in_begin + i]; d[k] = (int)(row[v + 1] - row[v]); }
            local += d[k];
        }
        int total;
        int pre = block_excl_scan<kBlock / MG_WAVE>(local, s_tmp, &total);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (i0 + k < n) eoff[i0 + k] = pre;
            pre += d[k];
        }
        if (threadIdx.x == 0) tsum[t] = total;
        __syncthreads();
    }
}

// 2. one workgroup scans the tile sums in place and sizes the chunk grid
__global__ __launch_bounds__(1024) void k_lb_tscan(Ctl *ctl, int *__restrict__ tsum) {
    if (ctl->mode != MODE_TD || !ctl->lb) return;
    const int T = (ctl->in_n + kTile - 1) / kTile;
    __shared__ int s_tmp[1024 / MG_WAVE + 1];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int b = 0; b < T; b += 1024) {
        const int i = b + threadIdx.x;
        const int v = i < T ? tsum[i] : 0;
        int total;
        const int pre = block_excl_scan<1024 / MG_WAVE>(v, s_tmp, &total);
        const int carry = s_carry;
        if (i < T) tsum[i] = carry + pre;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int epb = lb_epb(s_carry);
        ctl->lb_total = s_carry;
        ctl->lb_chunks = (s_carry + epb - 1) / epb;
    }
}

// global exclusive edge offset of frontier entry i
__device__ __forceinline__ int lb_off(const int *eoff, const int *toff, int i) { return eoff[i] + toff[i / kTile]; }

// largest frontier entry i with lb_off(i) <= e (the owner of virtual edge e)
__device__ int lb_owner(const int *eoff, const int *toff, int n, int e) {
    const int T = (n + kTile - 1) / kTile;
    int lo = 0, hi = T - 1;
    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (toff[mid] <= e) lo = mid; else hi = mid - 1; }
    const int t = lo, base = toff[t];
    int a = t * kTile, b = min(n, (t + 1) * kTile) - 1;
    while (a < b) { const int mid = (a + b + 1) >> 1; if (base + eoff[mid] <= e) a = mid; else b = mid - 1; }
    return a;
}

// 3. owner of the first edge of every chunk (cstart[chunks] = owner of the last edge)
__global__ void k_lb_part(Ctl *ctl, const int *__restrict__ eoff, const int *__restrict__ toff, int *__restrict__ cstart) {
    if (ctl->mode != MODE_TD || !ctl->lb) return;
    const int n = ctl->in_n, nch = ctl->lb_chunks, total = ctl->lb_total, epb = lb_epb(total);
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= nch; c += gridDim.x * blockDim.x) {
        const int e = c < nch ? c * epb : total - 1;
        cstart[c] = lb_owner(eoff, toff, n, e);
    }
}

// 4. the push: each workgroup takes kEPB consecutive edges of the frontier's edge list
template <class H, int B, class OT = int>
__global__ __launch_bounds__(B) void k_push_lb(Ctl *ctl, const OT *__restrict__ row, const int *__restrict__ col,
                                                    const int *__restrict__ wgt, int *__restrict__ state,
                                                    int *__restrict__ vfc, double *__restrict__ sigma,
                                                    int *__restrict__ q, const int *__restrict__ eoff,
                                                    const int *__restrict__ toff, const int *__restrict__ cstart) {
    if (ctl->mode != MODE_TD || !ctl->lb) return;
    const int iter = ctl->iter, in_begin = ctl->in_begin, n = ctl->in_n, out_begin = ctl->out_begin;
    const int nch = ctl->lb_chunks, total = ctl->lb_total, epb = lb_epb(total);
    constexpr int kCap = kEPB + 1;
    __shared__ int s_off[kCap];
    __shared__ OT s_rb[kCap];
    __shared__ int s_v[kCap];
    __shared__ int s_wc[B / MG_WAVE + 1];
    __shared__ long long s_wd[B / MG_WAVE];
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const int i0 = cstart[c];
        const int i1 = cstart[c + 1];
        const int cnt = i1 - i0 + 1;
        const bool staged = cnt <= kCap;
        if (staged) {
            for (int k = threadIdx.x; k < cnt; k += B) {
                const int i = i0 + k;
                const int v = q[in_begin + i];
                s_off[k] = lb_off(eoff, toff, i);
                s_rb[k] = row[v];
                s_v[k] = v;
            }
        }
        __syncthreads();
        const int e0 = c * epb, e1 = min(total, e0 + epb);
        for (int eb = e0; eb < e1; eb += B) {
            const int e = eb + threadIdx.x;
            int win = 0, u = 0;
            if (e < e1) {
                int k, off, v;
                OT rb;
                if (staged) {
                    int lo = 0, hi = cnt - 1;
                    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (s_off[mid] <= e) lo = mid; else hi = mid - 1; }
                    k = lo; off = s_off[k]; rb = s_rb[k]; v = s_v[k];
                } else {  // only reachable with zero-degree frontier entries
                    const int i = lb_owner(eoff, toff, n, e);
                    v = q[in_begin + i]; off = lb_off(eoff, toff, i); rb = row[v];
                }
                const OT p = rb + (e - off);
                u = ld_stream(&col[p]);
                win = relax<H>(state[v], sigma ? sigma[v] : 0.0, u, p, wgt, state, vfc, sigma, iter);
            }
            block_append<B, true>(ctl, q, out_begin, row, win != 0, u, s_wc, s_wd);
        }
        __syncthreads();
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

// Unvisited vertices above pull_split edges, cut into static kPullChunk-edge chunks
// (PullChunk {v, p0, p1, k}, k = the vertex's slot): a wave per chunk scans it and
// merges its accumulator into the vertex's slot (hlev, hsum) with accumulate (FP64
// path-count sums: exact in any order); k_pull_huge_fin then applies the slots to
// level, sigma, bitmaps and counts and resets them to the identity.  Without this
// split, one lane walks a 160 K-edge hub alone (BC scans every parent).  The chunks are
// walked by the last workgroups of the k_pull launch (waves wave, wave + nwaves, ...).
template <class H, class OT>
__device__ __forceinline__ void pull_huge_chunks(int iter, const uint64_t *fr, int nch, const PullChunkT<OT> *__restrict__ ch,
                                                 const int *__restrict__ col, const double *__restrict__ sigma,
                                                 const uint64_t *vis, int *__restrict__ hlev, double *__restrict__ hsum,
                                                 int wave, int nwaves) {
    for (int c = wave; c < nch; c += nwaves) {
        const PullChunkT<OT> k = ch[c];
        if (bm_test(vis, k.v)) continue;
        Acc<H> a;
        acc_init(a);
        bool found = false;
        for (OT p0 = k.p0; p0 < k.p1; p0 += MG_WAVE) {
            const OT p = p0 + lane();
            bool f = false;
            if (p < k.p1) {
                const int u = ld_stream(&col[p]);
                f = bm_test(fr, u);
                if (f) acc_edge<H>(a, iter - 1, H::kV2 ? sigma[u] : 0.0, 0, iter);
            }
            const bool any = ballot(f) != 0;
            found = found || any;
            if (H::kFirstWins && any) break;
        }
        if (!found) continue;
        a = acc_wave<H>(a, iter);
        if (lane() == 0) acc_atomic<H>(&hlev[k.k], H::kV2 ? &hsum[k.k] : nullptr, a, iter);
    }
}

// Pull level for BFS and the BC forward pass: every unvisited vertex accumulates the
// messages of its parents in the current frontier (a parent u of the frontier has
// level iter - 1; get_partial_result gives its message, accumulate_nonAtomics combines
// it into the vertex's register accumulator) and applies the result with
// update_condition.  BFS stops at the first parent found (its hooks say the first
// message is final); BC scans every parent, summing their path counts.  row / col are
// the in-edges; orow, the out-edge rows of a directed graph (null when symmetric),
// gives the degree sum.  The new frontier is the bitmap itself; its size and degree sum
// leave as one partial per workgroup (no global atomics), summed by k_ctl.
#ifndef MG_PULL_STEP
#define MG_PULL_STEP 4
#endif
constexpr int kPullStep = MG_PULL_STEP;  // BC pull: edges a lane loads together per step
#ifndef MG_PULL_FLAT
#define MG_PULL_FLAT 1            // BC pull: a word's rows as one flat edge range (0: lane / wave tiers)
#endif

template <class H, int B, class OT = int>
__global__ __launch_bounds__(B) void k_pull(Ctl *ctl, int nv, int cut, int lane_max, const OT *__restrict__ row,
                                                 const int *__restrict__ col, const OT *__restrict__ orow,
                                                 int *__restrict__ level,
                                                 double *__restrict__ sigma, uint64_t *__restrict__ vis,
                                                 uint64_t *__restrict__ bm0, uint64_t *__restrict__ bm1,
                                                 int *__restrict__ part_n, long long *__restrict__ part_mf,
                                                 int hblocks, int hnch, const PullChunkT<OT> *__restrict__ hch,
                                                 int *__restrict__ hlev, double *__restrict__ hsum) {
    if (ctl->mode != MODE_BU) return;
    constexpr bool kSum = !H::kFirstWins;  // every parent contributes
    const int iter = ctl->iter;
    const int plev = iter - 1;  // level of the frontier's vertices
    const uint64_t *fr = ctl->bm_cur ? bm1 : bm0;
    uint64_t *nx = ctl->bm_cur ? bm0 : bm1;
    // the last hblocks workgroups take the chunks of the rows above `cut` edges (in the
    // same launch: those rows are never walked by the others, and the words the others
    // write keep those vertices' bits)
    const int mblocks = gridDim.x - hblocks;
    if ((int)blockIdx.x >= mblocks) {
        pull_huge_chunks<H, OT>(iter, fr, hnch, hch, col, sigma, vis, hlev, hsum, ((blockIdx.x - mblocks) * B + threadIdx.x) >> 6,
                                (hblocks * B) >> 6);
        return;
    }
    const int words = (nv + 63) >> 6;
    const int wave = (blockIdx.x * B + threadIdx.x) >> 6;
    const int nwaves = (mblocks * B) >> 6;
#if MG_PULL_FLAT
    // per wave: the flat walk's row prefixes and starts, and the rows' accumulators
    __shared__ int s_incl[kSum ? B : 1];
    __shared__ OT s_base[kSum ? B : 1];
    __shared__ typename H::V1T s_lv[kSum ? B : 1];
    __shared__ double s_ls[kSum ? B : 1];
#endif
    int cnt = 0;
    long long mf = 0;
    for (int w = wave; w < words; w += nwaves) {
        const uint64_t vw = vis[w];
        const int v = (w << 6) + lane();
        Acc<H> a;
        acc_init(a);
        OT b = 0, e = 0;
        int deg = 0;
        const bool open = vw != ~0ull && !((vw >> lane()) & 1ull);
        if (open) {
            b = row[v];
            e = row[v + 1];
            deg = (int)(e - b);
            if (e - b > cut) e = b;  // left to the chunk workgroups
            if (!kSum) {
                for (OT p = b; p < e; ++p)
                    if (bm_test(fr, col[p])) { acc_edge<H>(a, plev, 0.0, 0, iter); break; }
            }
        }
        if (kSum) {
#if MG_PULL_FLAT
            // The word's rows (those of at most `cut` edges; longer ones are chunked)
            // walked as one flat edge range: step by step, lane l takes kPullStep
            // consecutive edges of it.  Every step is full whatever the row lengths, and
            // the rows of consecutive vertices are consecutive in col, so a step reads one
            // contiguous run.  A lane sums its edges' path counts per row in registers and
            // adds each row's sum into that row's LDS slot (accumulate) when its edges move
            // to the next row; path counts are integers, exact in FP64 in any order.
            const int wl = threadIdx.x & ~(MG_WAVE - 1);
            const int hd = (int)(e - b);  // 0 for a closed or cut row
            int incl = hd;                // inclusive prefix of the row lengths over the lanes
#pragma unroll
            for (int o = 1; o < MG_WAVE; o <<= 1) {
                const int t = __shfl_up(incl, o);
                if (lane() >= o) incl += t;
            }
            const int T = __shfl(incl, MG_WAVE - 1);
            {
                Acc<H> id;
                acc_init(id);
                wave_sync();  // the previous word's slots have been read
                s_incl[threadIdx.x] = incl;
                s_base[threadIdx.x] = b - (OT)(incl - hd);  // col index of flat edge x of this row: s_base + x
                s_lv[threadIdx.x] = id.v;
                s_ls[threadIdx.x] = id.s;
                wave_sync();
            }
            for (int x0 = 0; x0 < T; x0 += MG_WAVE * kPullStep) {
                const int xb = x0 + lane() * kPullStep;
                int r = 0;  // row of flat edge xb: the number of rows ending at or before it
                if (xb < T) {
#pragma unroll
                    for (int d = MG_WAVE / 2; d > 0; d >>= 1)
                        if (s_incl[wl + r + d - 1] <= xb) r += d;
                }
                int u[kPullStep], rj[kPullStep];
                bool f[kPullStep];
                double sv[kPullStep];
#pragma unroll
                for (int k = 0; k < kPullStep; ++k) {
                    const int x = xb + k;
                    f[k] = x < T;
                    if (f[k])
                        while (s_incl[wl + r] <= x) ++r;  // ends at r <= 63: incl[63] = T > x
                    rj[k] = r;
                    u[k] = f[k] ? col[s_base[wl + r] + x] : 0;
                }
#pragma unroll
                for (int k = 0; k < kPullStep; ++k) f[k] = f[k] && bm_test(fr, u[k]);
#pragma unroll
                for (int k = 0; k < kPullStep; ++k) sv[k] = f[k] ? sigma[u[k]] : 0.0;
                Acc<H> cur;
                acc_init(cur);
                int cr = rj[0];
#pragma unroll
                for (int k = 0; k < kPullStep; ++k) {
                    if (rj[k] != cr) {
                        if (!acc_is_identity<H>(cur)) acc_atomic<H>(&s_lv[wl + cr], &s_ls[wl + cr], cur, iter);
                        acc_init(cur);
                        cr = rj[k];
                    }
                    if (f[k]) acc_edge<H>(cur, plev, sv[k], 0, iter);
                }
                if (!acc_is_identity<H>(cur)) acc_atomic<H>(&s_lv[wl + cr], &s_ls[wl + cr], cur, iter);
            }
            wave_sync();
            a.v = s_lv[threadIdx.x];
            a.s = s_ls[threadIdx.x];
#else
            // lanes walk rows of at most lane_max edges alone, stepping together for the
            // longest of them; longer rows are walked by the whole wave, one at a time
            // (lanes over edges, then the lanes' accumulators merged).  Path counts are
            // integers, exact in FP64 whatever the order of the additions.
            const bool alone = e - b <= lane_max;
            int steps = alone ? (int)(e - b) : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) steps = max(steps, __shfl_xor(steps, o));
            // kPullStep edges per step, their column, bitmap and path-count loads issued
            // together (three dependent loads per edge otherwise serialise)
            for (int k = 0; k < steps; k += kPullStep) {
                int u[kPullStep];
                bool f[kPullStep];
                double sv[kPullStep];
#pragma unroll
                for (int j = 0; j < kPullStep; ++j) {
                    f[j] = alone && b + k + j < e;
                    u[j] = f[j] ? col[b + k + j] : 0;  // caching load: the row's next step reads the same line
                }
#pragma unroll
                for (int j = 0; j < kPullStep; ++j) f[j] = f[j] && bm_test(fr, u[j]);
#pragma unroll
                for (int j = 0; j < kPullStep; ++j) sv[j] = f[j] ? sigma[u[j]] : 0.0;
#pragma unroll
                for (int j = 0; j < kPullStep; ++j)
                    if (f[j]) acc_edge<H>(a, plev, sv[j], 0, iter);
            }
            uint64_t heavy = ballot(!alone);
            while (heavy) {
                const int l = __builtin_ctzll(heavy);
                heavy &= heavy - 1;
                const OT q0 = __shfl(b, l), q1 = __shfl(e, l);
                Acc<H> h;
                acc_init(h);
                for (OT p0 = q0; p0 < q1; p0 += 4 * MG_WAVE) {  // four rounds of loads in flight
                    int u4[4];
                    bool f4[4];
                    double s4[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const OT p = p0 + j * MG_WAVE + lane();
                        f4[j] = p < q1;
                        u4[j] = f4[j] ? ld_stream(&col[p]) : 0;
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) f4[j] = f4[j] && bm_test(fr, u4[j]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) s4[j] = f4[j] ? sigma[u4[j]] : 0.0;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (f4[j]) acc_edge<H>(h, plev, s4[j], 0, iter);
                }
                h = acc_wave<H>(h, iter);
                if (lane() == l) a = h;
            }
#endif
        }
        // an unvisited vertex's level is the identity: its register copy is the old value
        Acc<H> none;
        acc_init(none);
        const bool found = open && acc_apply<H>(none.v, &level[v], sigma ? &sigma[v] : nullptr, nullptr, a, iter);
        const uint64_t m = ballot(found);
        if (lane() == 0) { nx[w] = m; if (m) vis[w] = vw | m; }
        if (found) {
            cnt += 1;
            mf += orow ? orow[v + 1] - orow[v] : deg;  // out-degree: the next level's push edges
        }
    }
    __shared__ long long s_tmp[B / MG_WAVE];
    const long long tn = block_sum<B / MG_WAVE>((long long)cnt, s_tmp);
    const long long tm = block_sum<B / MG_WAVE>(mf, s_tmp);
    if (threadIdx.x == 0) { part_n[blockIdx.x] = (int)tn; part_mf[blockIdx.x] = tm; }
}

// the chunked rows' slots applied (after the k_pull launch whose chunk workgroups filled them)
template <class H, class OT = int>
__global__ __launch_bounds__(kBlock) void k_pull_huge_fin(Ctl *ctl, int nh, const int *__restrict__ hv,
                                                          const OT *__restrict__ row, int *__restrict__ level,
                                                          double *__restrict__ sigma, uint64_t *__restrict__ vis,
                                                          uint64_t *__restrict__ bm0, uint64_t *__restrict__ bm1,
                                                          int *__restrict__ hlev, double *__restrict__ hsum,
                                                          int *__restrict__ part_n, long long *__restrict__ part_mf) {
    if (ctl->mode != MODE_BU) return;
    const int iter = ctl->iter;
    uint64_t *nx = ctl->bm_cur ? bm0 : bm1;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nh; k += gridDim.x * kBlock) {
        Acc<H> a;
        a.v = hlev[k];
        a.s = H::kV2 ? hsum[k] : 0.0;
        if (acc_is_identity<H>(a)) continue;
        const int v = hv[k];
        Acc<H> none;
        acc_init(none);
        hlev[k] = none.v;  // the slot back to the identity for the next level
        if (H::kV2) hsum[k] = none.s;
        if (bm_test(vis, v) || !acc_apply<H>(none.v, &level[v], H::kV2 ? &sigma[v] : nullptr, nullptr, a, iter)) continue;
        atomicOr((unsigned long long *)&nx[v >> 6], 1ull << (v & 63));
        atomicOr((unsigned long long *)&vis[v >> 6], 1ull << (v & 63));
        atomicAdd(&part_n[0], 1);
        atomicAdd((unsigned long long *)&part_mf[0], (unsigned long long)(row[v + 1] - row[v]));
    }
}

// slots of the split pull vertices at the identity (once, at build)
template <class H>
__global__ void k_pull_slots_init(int nh, int *__restrict__ hlev, double *__restrict__ hsum) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nh; k += gridDim.x * blockDim.x) {
        Acc<H> a;
        acc_init(a);
        hlev[k] = a.v;
        hsum[k] = a.s;
    }
}

// Queue of the frontier bitmap bm[qsrc] (after a pull level), in vertex order, at
// q[in_begin]: popcounts per workgroup range, one-workgroup scan, ordered write.
__global__ __launch_bounds__(kBlock) void k_bmq_count(Ctl *ctl, int words, const uint64_t *__restrict__ bm0,
                                                      const uint64_t *__restrict__ bm1, int *__restrict__ part) {
    if (!ctl->need_queue) return;
    const uint64_t *bm = ctl->qsrc ? bm1 : bm0;
    const int per = (words + gridDim.x - 1) / gridDim.x;
    const int w0 = blockIdx.x * per, w1 = min(words, w0 + per);
    long long c = 0;
    for (int w = w0 + threadIdx.x; w < w1; w += kBlock) c += __builtin_popcountll(bm[w]);
    __shared__ long long s_tmp[kBlock / MG_WAVE];
    c = block_sum<kBlock / MG_WAVE>(c, s_tmp);
    if (threadIdx.x == 0) part[blockIdx.x] = (int)c;
}

__global__ __launch_bounds__(1024) void k_bmq_scan(Ctl *ctl, int *__restrict__ part, int nparts) {
    if (!ctl->need_queue) return;
    __shared__ int s_tmp[1024 / MG_WAVE + 1];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int b = 0; b < nparts; b += 1024) {
        const int i = b + threadIdx.x;
        const int v = i < nparts ? part[i] : 0;
        int total;
        const int pre = block_excl_scan<1024 / MG_WAVE>(v, s_tmp, &total);
        const int carry = s_carry;
        if (i < nparts) part[i] = carry + pre;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_bmq_write(Ctl *ctl, int words, const uint64_t *__restrict__ bm0,
                                                      const uint64_t *__restrict__ bm1, const int *__restrict__ part,
                                                      int *__restrict__ q) {
    if (!ctl->need_queue) return;
    const uint64_t *bm = ctl->qsrc ? bm1 : bm0;
    const int per = (words + gridDim.x - 1) / gridDim.x;
    const int w0 = blockIdx.x * per, w1 = min(words, w0 + per);
    int base = ctl->in_begin + part[blockIdx.x];
    __shared__ int s_tmp[kBlock / MG_WAVE + 1];
    for (int wb = w0; wb < w1; wb += kBlock) {
        const int w = wb + threadIdx.x;
        uint64_t m = w < w1 ? bm[w] : 0ull;
        int total;
        int pos = base + block_excl_scan<kBlock / MG_WAVE>((int)__builtin_popcountll(m), s_tmp, &total);
        while (m) {
            const int bit = __builtin_ctzll(m);
            q[pos++] = (w << 6) + bit;
            m &= m - 1;
        }
        base += total;
        __syncthreads();
    }
}

// -------------------------------------------------------------------- SSSP pull
// Dense SSSP level without global atomics: every vertex takes the minimum of
// dist[u] + w over its reached neighbours u, edge-balanced over the whole CSR in
// chunks of kEPB edges.  This is the reference's dense all-edge relaxation
// (process.cu:653-934) without its message buffers: on a multi-die GPU device-scope
// atomics resolve beyond the die's L2, so a push over a large frontier pays one such
// atomic per improving edge.
//
// A lane takes kPE consecutive edges of the chunk: their ids and weights arrive as
// 16-byte loads and all kPE source distances are gathered before any is used.  The
// chunk's row offsets are staged in LDS; a lane finds the owner of its first edge by
// binary search and keeps a running minimum per owner in a register, so it issues one
// LDS atomicMin per owner run, not per edge (consecutive edges mostly share an owner).
// Then an owner whose row lies inside the chunk is written with a plain store, the
// (at most two) owners that straddle a chunk boundary with atomicMin.  Every reached
// source relaxes its edges, read live (an update made earlier in this sweep
// propagates at once): a frontier probe per edge costs more random traffic than the
// redundant relaxations it would skip.
constexpr int kPE = kEPB / kBlock;  // consecutive edges per lane (8)

// Hooks: a lane's running minimum is its register accumulator (initialize,
// get_partial_result per edge, accumulate_nonAtomics); an owner run is flushed into the
// owner's LDS slot with accumulate (LDS atomicMin); an owner whose row lies inside the
// chunk is applied with update_condition, one that straddles a chunk boundary (another
// workgroup may update it too) with fused_update_condition.  Unreached sources
// (distance MG_INF) send nothing.
template <class H, class OT = int>
__global__ __launch_bounds__(kBlock) void k_pull_sssp(Ctl *ctl, int nch, long long ne, const OT *__restrict__ row,
                                                      const int *__restrict__ col, const int *__restrict__ wgt,
                                                      const int *__restrict__ cstart, int *__restrict__ dist,
                                                      int *__restrict__ vfc) {
    if (ctl->mode != MODE_BU) return;
    const int iter = ctl->iter;
    __shared__ OT s_row[kEPB + 2];
    __shared__ int s_best[kEPB + 1];
    Acc<H> ident;
    acc_init(ident);
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const OT e0 = (OT)c * kEPB;
        const OT e1 = (OT)min((long long)e0 + kEPB, ne);
        const int u0 = cstart[c], u1 = cstart[c + 1];
        const int nown = u1 - u0 + 1;
        const OT eb = e0 + threadIdx.x * kPE;
        int cu[kPE], wv[kPE], du[kPE];
        if (eb + kPE <= e1) {
            const uint4 c0 = ld_stream_u4(&col[eb]), c1 = ld_stream_u4(&col[eb + 4]);
            const uint4 w0 = ld_stream_u4(&wgt[eb]), w1 = ld_stream_u4(&wgt[eb + 4]);
            cu[0] = c0.x; cu[1] = c0.y; cu[2] = c0.z; cu[3] = c0.w; cu[4] = c1.x; cu[5] = c1.y; cu[6] = c1.z; cu[7] = c1.w;
            wv[0] = w0.x; wv[1] = w0.y; wv[2] = w0.z; wv[3] = w0.w; wv[4] = w1.x; wv[5] = w1.y; wv[6] = w1.z; wv[7] = w1.w;
        } else {
#pragma unroll
            for (int j = 0; j < kPE; ++j) {
                cu[j] = eb + j < e1 ? col[eb + j] : 0;
                wv[j] = eb + j < e1 ? wgt[eb + j] : 0;
            }
        }
#pragma unroll
        for (int j = 0; j < kPE; ++j) du[j] = eb + j < e1 ? dist[cu[j]] : (int)MG_INF;
        if (nown > kEPB + 1) {  // empty rows inside the chunk (not after compaction): owners by search in global row
            for (int j = 0; j < kPE; ++j) {
                const OT e = eb + j;
                if (e >= e1 || du[j] == (int)MG_INF) continue;
                int lo = u0, hi = u1;
                while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (row[mid] <= e) lo = mid; else hi = mid - 1; }
                Acc<H> a = ident;
                acc_edge<H>(a, du[j], 0.0, wv[j], iter);
                acc_apply_atomic<H>(&dist[lo], nullptr, &vfc[lo], a, iter);
            }
            continue;  // uniform: nown is the same for the whole workgroup
        }
        for (int k = threadIdx.x; k <= nown; k += kBlock) s_row[k] = row[u0 + k];
        for (int k = threadIdx.x; k < nown; k += kBlock) s_best[k] = ident.v;
        __syncthreads();
        if (eb < e1) {
            int lo = 0, hi = nown - 1;  // owner of edge eb: the last k with s_row[k] <= eb
            while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (s_row[mid] <= eb) lo = mid; else hi = mid - 1; }
            int k = lo;
            OT next = s_row[lo + 1];
            Acc<H> a = ident;
#pragma unroll
            for (int j = 0; j < kPE; ++j) {
                const OT e = eb + j;
                if (e < e1) {
                    while (e >= next) {
                        if (a.v != ident.v) acc_atomic<H>(&s_best[k], nullptr, a, iter);
                        a = ident;
                        ++k;
                        next = s_row[k + 1];
                    }
                    if (du[j] != (int)MG_INF) acc_edge<H>(a, du[j], 0.0, wv[j], iter);
                }
            }
            if (a.v != ident.v) acc_atomic<H>(&s_best[k], nullptr, a, iter);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < nown; k += kBlock) {
            Acc<H> a = ident;
            a.v = s_best[k];
            if (a.v == ident.v) continue;
            const int o = u0 + k;
            if (s_row[k] < e0 || s_row[k + 1] > e1) acc_apply_atomic<H>(&dist[o], nullptr, &vfc[o], a, iter);
            else acc_apply<H>(dist[o], &dist[o], nullptr, &vfc[o], a, iter);
        }
        __syncthreads();
    }
}

// Frontier of stamp `iter + offset` from the stamps: a bitmap word per wave
// (ballot) and per-workgroup partial counts (size, degree sum) for k_ctl.
// gate 0: always; 1: only on a pull level; 2: only right after a switch to pull
template <class OT = int>
__global__ __launch_bounds__(kBlock) void k_sssp_scan(Ctl *ctl, int nv, int offset, int gate, const int *__restrict__ vfc,
                                                      const OT *__restrict__ row, uint64_t *__restrict__ bm,
                                                      int *__restrict__ part_n, long long *__restrict__ part_mf) {
    if (gate >= 1 && ctl->mode != MODE_BU) return;
    if (gate == 2 && !ctl->convert) return;
    const int stamp = ctl->iter + offset;
    const int words = (nv + 63) >> 6;
    const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * kBlock) >> 6;
    int cnt = 0;
    long long mf = 0;
    for (int w = wave; w < words; w += nwaves) {
        const int v = (w << 6) + lane();
        const bool hit = v < nv && vfc[v] == stamp;
        const uint64_t m = ballot(hit);
        if (lane() == 0) bm[w] = m;
        if (hit) { cnt += 1; mf += row[v + 1] - row[v]; }
    }
    __shared__ long long s_tmp[kBlock / MG_WAVE];
    const long long tn = block_sum<kBlock / MG_WAVE>((long long)cnt, s_tmp);
    const long long tm = block_sum<kBlock / MG_WAVE>(mf, s_tmp);
    if (threadIdx.x == 0 && part_n) { part_n[blockIdx.x] = (int)tn; part_mf[blockIdx.x] = tm; }
}

// -------------------------------------------------------------------- BC backward
// Levels as bytes for the backward pass (255 = level 255 or deeper, or unreached):
// every edge of the pass probes its neighbour's level, and the byte array (nv bytes)
// stays in each die's L2 where the int array would not.
__global__ void k_lev8(int nv, const int *__restrict__ level, uint8_t *__restrict__ lev8) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        const int l = level[v];
        lev8[v] = (uint8_t)(l >= 0 && l < 255 ? l : 255);
    }
}
// is w one level deeper than the level being processed (l + 1)?
__device__ __forceinline__ bool bc_next(int l, int w, const uint8_t *__restrict__ lev8, const int *__restrict__ level) {
    return l + 1 < 255 ? lev8[w] == (uint8_t)(l + 1) : level[w] == l + 1;
}

// delta[v] = sum over neighbours w one level deeper of sigma[v]/sigma[w] * (1 + delta[w])
// (process2.cu:5-123, validate.cu:182-189), for the vertices of one level.  Degree
// tiers (the engine's bin edges, 32 / 2048 / 1024 by default): a lane per vertex of at
// most lane_max edges, a wave per vertex below split, and the whole grid (chunks of
// `chunk` edges, FP64 partials) from split up.  Sums are
// Neumaier-compensated FP32 over terms split into two floats, so each term keeps ~48 bits.
// The per-edge rule, its compensated sum and the vertex update are BCHooks' backward
// hooks (back_term, back_accumulate, back_update): c[w] = (1 + delta[w]) / sigma[w] is
// stored (in the hooks' CT, FP32 by default) once per vertex when its delta is final,
// so a successor edge costs one 4-byte gather and a multiply, no division.
template <class H>
__device__ __forceinline__ void bc_term(KSum &acc, double sv, float cw) {
    double t;
    H::back_term(&t, sv, (double)cw);
    H::back_accumulate(&acc, t);
}

#ifndef MG_BACK_STEP
#define MG_BACK_STEP 4
#endif
constexpr int kBackStep = MG_BACK_STEP;  // BC backward: edges (or rounds of edges) a lane loads together

// Tiers: a lane per vertex of at most lane_max edges (here), a wave per vertex below
// split (k_bc_back_wave, over a list this kernel appends to), chunks of `chunk` edges
// from split up (k_bc_back_huge).  The wave tier runs as its own kernel so that every listed vertex
// has a wave of its own: walking them from a per-workgroup list left a few thousand
// waves on the whole GPU for a level of hub neighbours.
template <class H, int B>
__global__ __launch_bounds__(B) void k_bc_back(int l, const int *__restrict__ lvl_off, const int *__restrict__ q,
                                                    const int *__restrict__ row, const int *__restrict__ col,
                                                    const int *__restrict__ level, const uint8_t *__restrict__ lev8,
                                                    const double *__restrict__ sigma,
                                                    float *__restrict__ delta, float *__restrict__ cw,
                                                    HugeChunk *__restrict__ chunks, int *__restrict__ nchunks,
                                                    int *__restrict__ wl, int *__restrict__ wn, int lane_max, int split,
                                                    int chunk) {
    const int b = lvl_off[l], n = lvl_off[l + 1] - b;
    __shared__ int s_wc[B / MG_WAVE + 1];
    for (int base = blockIdx.x * B; base < n; base += gridDim.x * B) {
        const int i = base + threadIdx.x;
        int v = 0, rb = 0, re = 0;
        if (i < n) {
            v = q[b + i];
            rb = row[v];
            re = row[v + 1];
        }
        const int deg = re - rb;
        const bool mid = i < n && deg > lane_max && deg < split;
        const int pos = block_append_slot<B>(wn, mid, s_wc);
        if (mid) wl[pos] = v;
        if (i < n && deg <= lane_max) {
            const double sv = sigma[v];
            KSum acc;
            // kBackStep edges per step: their column, level and c loads are in flight together
            for (int p = rb; p < re; p += kBackStep) {
                int w4[kBackStep];
                bool f[kBackStep];
                float c4[kBackStep];
#pragma unroll
                for (int j = 0; j < kBackStep; ++j) { f[j] = p + j < re; w4[j] = f[j] ? col[p + j] : 0; }
#pragma unroll
                for (int j = 0; j < kBackStep; ++j) f[j] = f[j] && bc_next(l, w4[j], lev8, level);
#pragma unroll
                for (int j = 0; j < kBackStep; ++j) c4[j] = f[j] ? cw[w4[j]] : 0.f;
#pragma unroll
                for (int j = 0; j < kBackStep; ++j)
                    if (f[j]) bc_term<H>(acc, sv, c4[j]);
            }
            H::back_update(&delta[v], &cw[v], acc.value(), sv);
        } else if (i < n && deg >= split) {
            const int nch = (deg + chunk - 1) / chunk;
            const int c0 = atomicAdd(nchunks, nch);
            for (int j = 0; j < nch; ++j) chunks[c0 + j] = HugeChunk{v, j};
        }
    }
}

// a wave per listed vertex (32 to 2047 edges): lanes over edges, kBackStep rounds of
// loads in flight, the lanes' compensated sums merged
template <class H, int B>
__global__ __launch_bounds__(B) void k_bc_back_wave(int l, const int *__restrict__ wl, const int *__restrict__ wn,
                                                         const int *__restrict__ row, const int *__restrict__ col,
                                                         const int *__restrict__ level, const uint8_t *__restrict__ lev8,
                                                         const double *__restrict__ sigma, float *__restrict__ delta,
                                                         float *__restrict__ cw) {
    const int n = *wn;
    const int wave = (blockIdx.x * B + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * B) >> 6;
    for (int k = wave; k < n; k += nwaves) {
        const int v = wl[k];
        const int rb = row[v], re = row[v + 1];
        const double sv = sigma[v];
        KSum acc;
        for (int p0 = rb; p0 < re; p0 += kBackStep * MG_WAVE) {
            int w4[kBackStep];
            bool f[kBackStep];
            float c4[kBackStep];
#pragma unroll
            for (int j = 0; j < kBackStep; ++j) {
                const int p = p0 + j * MG_WAVE + lane();
                f[j] = p < re;
                w4[j] = f[j] ? ld_stream(&col[p]) : 0;
            }
#pragma unroll
            for (int j = 0; j < kBackStep; ++j) f[j] = f[j] && bc_next(l, w4[j], lev8, level);
#pragma unroll
            for (int j = 0; j < kBackStep; ++j) c4[j] = f[j] ? cw[w4[j]] : 0.f;
#pragma unroll
            for (int j = 0; j < kBackStep; ++j)
                if (f[j]) bc_term<H>(acc, sv, c4[j]);
        }
        acc = wave_ksum(acc);
        if (lane() == 0) H::back_update(&delta[v], &cw[v], acc.value(), sv);
    }
}

template <class H, int B>
__global__ __launch_bounds__(B) void k_bc_back_huge(int l, const int *__restrict__ row, const int *__restrict__ col,
                                                         const int *__restrict__ level, const uint8_t *__restrict__ lev8,
                                                         const double *__restrict__ sigma,
                                                         const float *__restrict__ cw, const HugeChunk *__restrict__ chunks,
                                                         const int *__restrict__ nchunks, double *__restrict__ dacc, int chunk) {
    const int nch = *nchunks;
    __shared__ float s_s[B / MG_WAVE], s_c[B / MG_WAVE];
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const HugeChunk hc = chunks[c];
        const int v = hc.v;
        const int b = row[v] + hc.j * chunk, e = min(row[v + 1], b + chunk);
        const double sv = sigma[v];
        KSum acc;
        for (int p0 = b; p0 < e; p0 += kBackStep * B) {
            int w4[kBackStep];
            bool f[kBackStep];
            float c4[kBackStep];
#pragma unroll
            for (int j = 0; j < kBackStep; ++j) {
                const int p = p0 + j * B + threadIdx.x;
                f[j] = p < e;
                w4[j] = f[j] ? ld_stream(&col[p]) : 0;
            }
#pragma unroll
            for (int j = 0; j < kBackStep; ++j) f[j] = f[j] && bc_next(l, w4[j], lev8, level);
#pragma unroll
            for (int j = 0; j < kBackStep; ++j) c4[j] = f[j] ? cw[w4[j]] : 0.f;
#pragma unroll
            for (int j = 0; j < kBackStep; ++j)
                if (f[j]) bc_term<H>(acc, sv, c4[j]);
        }
        acc = wave_ksum(acc);
        if (lane() == 0) { s_s[threadIdx.x >> 6] = acc.s; s_c[threadIdx.x >> 6] = acc.c; }
        __syncthreads();
        if (threadIdx.x == 0) {
            KSum t;
            for (int k = 0; k < B / MG_WAVE; ++k) { t.add(s_s[k]); t.add(s_c[k]); }
            atomicAdd(&dacc[v], (double)t.s + (double)t.c);
        }
        __syncthreads();
    }
}

template <class H>
__global__ void k_bc_back_fin(const HugeChunk *__restrict__ chunks, int *__restrict__ nchunks, double *__restrict__ dacc,
                              const double *__restrict__ sigma, float *__restrict__ delta, float *__restrict__ cw) {
    const int nch = *nchunks;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nch; c += gridDim.x * blockDim.x) {
        if (chunks[c].j != 0) continue;  // the first chunk of each vertex finalizes it
        const int v = chunks[c].v;
        H::back_update(&delta[v], &cw[v], (float)dacc[v], sigma[v]);
        dacc[v] = 0.0;
    }
}

// Backward level l by push: the vertices w of level l + 1 add c[w] into an FP64
// accumulator of each neighbour v one level up (l); k_bc_push_fin then sets
// delta[v] = sigma[v] * sum.  Chosen over the pull when level l + 1 has far fewer edges
// than level l (from a hub, the widest level sits above a thin one).  Tiers as in the
// pull: a lane per vertex below 32 edges, a wave below 2048, kHugeChunk-edge chunks
// above (k_bc_push_huge).
template <class H>
__global__ __launch_bounds__(kBlock) void k_bc_push(int l, const int *__restrict__ lvl_off, const int *__restrict__ q,
                                                    const int *__restrict__ row, const int *__restrict__ col,
                                                    const int *__restrict__ level, const uint8_t *__restrict__ lev8,
                                                    const float *__restrict__ cw, double *__restrict__ dacc,
                                                    HugeChunk *__restrict__ chunks, int *__restrict__ nchunks) {
    const int b = lvl_off[l + 1], n = lvl_off[l + 2] - b;
    __shared__ int s_wv[kBlock];
    __shared__ int s_wn;
    const int up = l;  // the level the contributions go to
    auto is_up = [&](int v) { return up < 255 ? lev8[v] == (uint8_t)up : level[v] == up; };
    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        if (threadIdx.x == 0) s_wn = 0;
        __syncthreads();
        const int i = base + threadIdx.x;
        if (i < n) {
            const int w = q[b + i];
            const int rb = row[w], re = row[w + 1], deg = re - rb;
            if (deg < 32) {
                const double c = cw[w];
                for (int p = rb; p < re; ++p) {
                    const int v = col[p];
                    if (is_up(v)) atomicAdd(&dacc[v], c);
                }
            } else if (deg < 2048) {
                s_wv[atomicAdd(&s_wn, 1)] = w;
            } else {
                const int nch = (deg + kHugeChunk - 1) / kHugeChunk;
                const int c0 = atomicAdd(nchunks, nch);
                for (int j = 0; j < nch; ++j) chunks[c0 + j] = HugeChunk{w, j};
            }
        }
        __syncthreads();
        const int wn = s_wn;
        for (int k = threadIdx.x >> 6; k < wn; k += kBlock / MG_WAVE) {
            const int w = s_wv[k];
            const double c = cw[w];
            for (int p = row[w] + lane(); p < row[w + 1]; p += MG_WAVE) {
                const int v = ld_stream(&col[p]);
                if (is_up(v)) atomicAdd(&dacc[v], c);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_bc_push_huge(int l, const int *__restrict__ row, const int *__restrict__ col,
                                                         const int *__restrict__ level, const uint8_t *__restrict__ lev8,
                                                         const float *__restrict__ cw, const HugeChunk *__restrict__ chunks,
                                                         const int *__restrict__ nchunks, double *__restrict__ dacc) {
    const int nch = *nchunks;
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const HugeChunk hc = chunks[c];
        const int w = hc.v;
        const int b = row[w] + hc.j * kHugeChunk, e = min(row[w + 1], b + kHugeChunk);
        const double cv = cw[w];
        for (int p = b + threadIdx.x; p < e; p += kBlock) {
            const int v = ld_stream(&col[p]);
            if (l < 255 ? lev8[v] == (uint8_t)l : level[v] == l) atomicAdd(&dacc[v], cv);
        }
    }
}

// level l after a push: delta = sigma * sum, its c for the level above, sums cleared
template <class H>
__global__ void k_bc_push_fin(int l, const int *__restrict__ lvl_off, const int *__restrict__ q, const double *__restrict__ sigma,
                              double *__restrict__ dacc, float *__restrict__ delta, float *__restrict__ cw) {
    const int b = lvl_off[l], n = lvl_off[l + 1] - b;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int v = q[b + i];
        const double sv = sigma[v];
        double t;  // the successors' c summed by the push: delta = sigma[v] * sum
        H::back_term(&t, sv, dacc[v]);
        H::back_update(&delta[v], &cw[v], (float)t, sv);
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

// the launch statements with kB bound to the engine's workgroup size for the pull,
// edge-balanced push and BC backward kernels (mg_options.block_threads)
#define MG_WITH_BLOCK(e, ...)                                                   \
    do {                                                                        \
        switch ((e)->opt.block_threads) {                                       \
        case 512: { constexpr int kB = 512; __VA_ARGS__; } break;               \
        case 1024: { constexpr int kB = 1024; __VA_ARGS__; } break;             \
        default: { constexpr int kB = kBlock; __VA_ARGS__; } break;             \
        }                                                                       \
    } while (0)

// the launch statements with OT bound to the engine's row-offset type and rowp / irowp
// to its out- / in-edge rows: int, or long long on a graph past 2^31 - 1 edges (or
// with mg_options.wide_offsets), which is symmetric, so its rows serve both ways
#define MG_WITH_OFFSETS(e, ...)                                                               \
    do {                                                                                      \
        if ((e)->wide) {                                                                      \
            using OT = long long;                                                             \
            const OT *rowp = (e)->row64, *irowp = (e)->row64;                                 \
            (void)rowp; (void)irowp;                                                          \
            __VA_ARGS__;                                                                      \
        } else {                                                                              \
            using OT = int;                                                                   \
            const OT *rowp = (e)->row, *irowp = (e)->in_row();                                \
            (void)rowp; (void)irowp;                                                          \
            __VA_ARGS__;                                                                      \
        }                                                                                     \
    } while (0)

int mg_frontier_build(mg_engine *e, const mg_graph *g) {
    const int nv = e->nv;
    const long long ne = e->ne;
    const long long padv = (long long)nv + 64;
    MG_HIP(hipMalloc(&e->level, sizeof(int) * padv));
    e->qcap = (e->algo == MG_SSSP) ? 2 * (long long)nv + 2 : (long long)nv + 2;
    MG_HIP(hipMalloc(&e->q, sizeof(int) * e->qcap));
    // chunks of one level: the backward's (ne / back_chunk + ne / back_split at most) or
    // the backward push's (kHugeChunk-edge chunks of vertices of 2048+ edges)
    e->chunk_cap = ne / e->opt.back_chunk + ne / e->opt.back_split + 2 * (ne / 2048) + nv / 1024 + 64;
    MG_HIP(hipMalloc(&e->chunks, sizeof(HugeChunk) * e->chunk_cap));
    MG_HIP(hipMalloc(&e->eoff, sizeof(int) * e->qcap));
    MG_HIP(hipMalloc(&e->tsum, sizeof(int) * (e->qcap / kTile + 2)));
    MG_HIP(hipMalloc(&e->cstart, sizeof(int) * (lb_max_chunks(ne) + 4)));
    MG_HIP(hipMalloc(&e->part_n, sizeof(int) * (max_grid(e) + 8)));
    MG_HIP(hipMalloc(&e->part_mf, sizeof(long long) * (max_grid(e) + 8)));
    MG_HIP(hipMalloc(&e->ctl, sizeof(Ctl)));
    MG_HIP(hipMalloc(&e->aux, sizeof(int) * 16));
    // coherent: k_ctl writes the host's copy of the control block directly
    MG_HIP(hipHostMalloc(&e->h_ctl, sizeof(Ctl), hipHostMallocCoherent));
    e->hdeg.resize((size_t)nv);
    for (int v = 0; v < nv; ++v) e->hdeg[(size_t)v] = (int32_t)(g->off((int64_t)v + 1) - g->off(v));
    // a directed graph's pull levels read in-edges: transpose the CSR once (counting
    // sort by destination; scanning sources in order leaves every row sorted by source)
    std::vector<int32_t> hrow_t, hcol_t;
    // host row offsets of the pull's in-edges, as 64-bit values (a wide graph is
    // symmetric: its rows serve both ways)
    auto irow = [&](int v) -> long long { return hrow_t.empty() ? g->off(v) : (long long)hrow_t[(size_t)v]; };
    if (!g->symmetric && nv > 0) {
        hrow_t.assign((size_t)nv + 1, 0);
        hcol_t.resize((size_t)ne);
        std::vector<int32_t> hw_t(e->algo == MG_SSSP ? (size_t)ne : 0);
        for (long long p = 0; p < ne; ++p) ++hrow_t[(size_t)g->col[p] + 1];
        for (int v = 0; v < nv; ++v) hrow_t[v + 1] += hrow_t[v];
        std::vector<int32_t> cur(hrow_t.begin(), hrow_t.end() - 1);
        for (int u = 0; u < nv; ++u)
            for (int p = g->row[u]; p < g->row[u + 1]; ++p) {
                const int32_t q = cur[g->col[p]]++;
                hcol_t[q] = u;
                if (!hw_t.empty()) hw_t[q] = g->w[p];
            }
        MG_HIP(hipMalloc(&e->trow, sizeof(int) * ((long long)nv + 1)));
        MG_HIP(hipMalloc(&e->tcol, sizeof(int) * (ne > 0 ? ne : 1)));
        MG_HIP(hipMemcpy(e->trow, hrow_t.data(), sizeof(int) * ((long long)nv + 1), hipMemcpyHostToDevice));
        if (ne > 0) MG_HIP(hipMemcpy(e->tcol, hcol_t.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
        if (!hw_t.empty()) {
            MG_HIP(hipMalloc(&e->tw, sizeof(int) * ne));
            MG_HIP(hipMemcpy(e->tw, hw_t.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
        }
    }
    if (e->algo == MG_SSSP) {
        MG_HIP(hipMalloc(&e->vfc, sizeof(int) * padv));
        const long long words = (nv + 63) / 64 + 1;
        MG_HIP(hipMalloc(&e->bm[0], sizeof(uint64_t) * words));
        e->pnch = (int)((ne + kEPB - 1) / kEPB);
        MG_HIP(hipMalloc(&e->pcstart, sizeof(int) * (e->pnch + 2)));
        if (e->pnch > 0)
            MG_WITH_OFFSETS(e, hipLaunchKernelGGL((k_chunk_part<OT>), dim3(grid_for(e->pnch + 1, 256, 4096)), dim3(256), 0, e->stream,
                                                  nv, ne, irowp, e->pnch, e->pcstart));
        // dense sweeps over the partitioned layout (symmetric graphs, weights in 16 bits),
        // else over the CSR (always on an engine with 64-bit offsets)
        bool w16 = g->symmetric && e->opt.pull_layout != 1 && nv >= 2 && !e->wide;
        int wmax = 0;
        for (long long p = 0; w16 && p < ne; ++p) {
            w16 = g->w[p] >= 0 && g->w[p] <= 0xFFFF;
            wmax = std::max(wmax, (int)g->w[p]);
        }
        if (w16) {
            e->pb.w8 = wmax <= 0xFF;
            // a layout past 2^31 - 1 edge slots (padded tiles of a graph near int32
            // offsets, e.g. RMAT-26) is refused by its builder: the sweeps then run over
            // the CSR, as for weights above 16 bits
            const int rc = mg_pb_build(e, g);
            if (rc == MG_OK) {
                e->sssp_pb = true;
                e->pb16_ok = e->opt.pull_layout == 0;  // 2: 32-bit messages throughout
            } else if (rc != MG_ERR_LIMIT) {
                return rc;
            }
        }
    } else {
        const long long words = (nv + 63) / 64 + 1;
        MG_HIP(hipMalloc(&e->bm[0], sizeof(uint64_t) * words));
        MG_HIP(hipMalloc(&e->bm[1], sizeof(uint64_t) * words));
        MG_HIP(hipMalloc(&e->vis, sizeof(uint64_t) * words));
        // pull chunks of the vertices above pull_split edges (static), with the engine's
        // offset type
        std::vector<PullChunkT<long long>> pc;
        std::vector<int> hv;
        for (int v = 0; v < nv; ++v) {
            const long long b = irow(v), d = irow(v + 1) - b;
            if (d <= e->opt.pull_split) continue;
            for (long long p = b; p < b + d; p += kPullChunk)
                pc.push_back(PullChunkT<long long>{v, p, std::min(b + d, p + kPullChunk), (int)hv.size()});
            hv.push_back(v);
        }
        e->pull_nch = (int)pc.size();
        e->pull_nh = (int)hv.size();
        if (!pc.empty()) {
            MG_HIP(hipMalloc(&e->pull_hv, sizeof(int) * hv.size()));
            MG_HIP(hipMalloc(&e->pull_found, sizeof(int) * hv.size()));
            MG_HIP(hipMalloc(&e->pull_sum, sizeof(double) * hv.size()));
            if (e->wide) {
                MG_HIP(hipMalloc(&e->pull_ch, sizeof(PullChunkT<long long>) * pc.size()));
                MG_HIP(hipMemcpy(e->pull_ch, pc.data(), sizeof(PullChunkT<long long>) * pc.size(), hipMemcpyHostToDevice));
            } else {
                std::vector<PullChunk> pc32(pc.size());
                for (size_t k = 0; k < pc.size(); ++k) pc32[k] = PullChunk{pc[k].v, (int)pc[k].p0, (int)pc[k].p1, pc[k].k};
                MG_HIP(hipMalloc(&e->pull_ch, sizeof(PullChunk) * pc.size()));
                MG_HIP(hipMemcpy(e->pull_ch, pc32.data(), sizeof(PullChunk) * pc.size(), hipMemcpyHostToDevice));
            }
            MG_HIP(hipMemcpy(e->pull_hv, hv.data(), sizeof(int) * hv.size(), hipMemcpyHostToDevice));
            if (e->algo == MG_BC)
                hipLaunchKernelGGL(k_pull_slots_init<MG_BC_HOOKS>, dim3(grid_for((long long)hv.size(), 256, 64)), dim3(256), 0,
                                   e->stream, (int)hv.size(), e->pull_found, e->pull_sum);
            else
                hipLaunchKernelGGL(k_pull_slots_init<MG_BFS_HOOKS>, dim3(grid_for((long long)hv.size(), 256, 64)), dim3(256), 0,
                                   e->stream, (int)hv.size(), e->pull_found, e->pull_sum);
        }
    }
    if (e->algo == MG_BC) {
        MG_HIP(hipMalloc(&e->sigma, sizeof(double) * padv));
        MG_HIP(hipMalloc(&e->delta, sizeof(float) * padv));
        MG_HIP(hipMalloc(&e->dacc, sizeof(double) * padv));
        MG_HIP(hipMalloc(&e->lev8, padv));
        MG_HIP(hipMalloc(&e->bc_cw, sizeof(float) * padv));
        MG_HIP(hipMalloc(&e->lvl_off, sizeof(int) * ((long long)nv + 4)));
        MG_HIP(hipMalloc(&e->lvl_mf, sizeof(long long) * ((long long)nv + 4)));
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(padv, 256, 4096)), dim3(256), 0, e->stream, e->dacc, padv, 0.0);
    }
    Params &P = e->P;
    P.nv = nv;
    P.ne = ne;
    P.log_queue = (e->algo == MG_SSSP) ? 0 : 1;
    // pull levels read in-edges: the CSR itself when symmetric, else its transpose
    P.forced = e->opt.mode;
    P.sssp = e->algo == MG_SSSP ? 1 : 0;
    P.need_log = e->algo == MG_BC ? 1 : 0;
    P.huge = e->opt.huge_degree;
    P.alpha = e->opt.alpha;
    P.beta = e->opt.beta;
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
    if (nv > 0 && !e->hdeg.empty()) {
        deg_src = e->hdeg[(size_t)source];
    } else if (nv > 0) {
        if (e->wide) {
            long long rr[2];
            MG_HIP(hipMemcpy(rr, e->row64 + source, sizeof(long long) * 2, hipMemcpyDeviceToHost));
            deg_src = (int)(rr[1] - rr[0]);
        } else {
            int rr[2];
            MG_HIP(hipMemcpy(rr, e->row + source, sizeof(int) * 2, hipMemcpyDeviceToHost));
            deg_src = rr[1] - rr[0];
        }
    }
    const int out_begin = (e->P.log_queue) ? 1 : nv;
    if (nv > 0)
        hipLaunchKernelGGL(k_seed, dim3(1), dim3(256), 0, s, e->ctl, e->level, e->q, e->sigma, source, deg_src, e->ne, out_begin, e->eoff,
                           e->tsum, e->cstart);
    MG_HIP(hipGetLastError());
    e->source = source;
    e->iter = 1;
    e->done = (nv == 0);
    e->levels = 1;
    e->pb16 = e->pb16_ok;
    e->pb_force = false;
    if (nv > 0) *e->h_ctl = seed_ctl(deg_src, e->ne, out_begin);  // what k_seed writes on the device
    return MG_OK;
}

static int launch_after_ctl(mg_engine *e);

// launch the kernels of one level; `known` = host knows the control block (eager mode)
static int launch_level(mg_engine *e, bool known) {
    const Ctl &c = *e->h_ctl;
    hipStream_t s = e->stream;
    const int G = max_grid(e);
    const int nv = e->nv;
    const bool td = !known || c.mode == MODE_TD;
    bool bu = !known || c.mode == MODE_BU;
    int slot = -1;
    if (bu && e->algo != MG_SSSP && (!known || c.convert)) {
        const int gb = grid_for((nv + 63) / 64, kBlock / MG_WAVE, G);
        if (e->prof) mg_prof_begin(e, "k_level2bm", 4.0 * nv + nv / 4.0, &slot);
        hipLaunchKernelGGL(k_level2bm, dim3(gb), dim3(kBlock), 0, s, e->ctl, nv, e->level, e->vis, e->bm[0], e->bm[1]);
        if (e->prof) mg_prof_end(e, slot);
    }
    if (td) {
        // algorithmic bytes: queue entry, row pair and own state per frontier vertex;
        // column and neighbour state per edge (+ weight for SSSP)
        const double per_edge = e->algo == MG_SSSP ? 12.0 : 8.0;
        const double bytes = known ? (16.0 * c.in_n + per_edge * (double)c.mf) : 0.0;
        if (!known || !c.lb) {
            const int gt = known ? grid_for(c.in_n, kBlock, G) : G;
            if (e->prof) mg_prof_begin(e, "k_push", bytes, &slot);
            MG_WITH_OFFSETS(e, switch (e->algo) {
                case MG_BFS:
                    hipLaunchKernelGGL((k_push<MG_BFS_HOOKS, true, OT>), dim3(gt), dim3(kBlock), 0, s, e->ctl, e->P, rowp, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
                    break;
                case MG_SSSP:
                    hipLaunchKernelGGL((k_push<MG_SSSP_HOOKS, true, OT>), dim3(gt), dim3(kBlock), 0, s, e->ctl, e->P, rowp, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
                    break;
                default:
                    hipLaunchKernelGGL((k_push<MG_BC_HOOKS, true, OT>), dim3(gt), dim3(kBlock), 0, s, e->ctl, e->P, rowp, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->chunks);
                    break;
            });
            if (e->prof) mg_prof_end(e, slot);
        }
        if (!known || c.lb) {
            const long long maxch = known ? (c.mf + lb_epb(c.mf) - 1) / lb_epb(c.mf) : lb_max_chunks(e->ne);
            const int gd = known ? grid_for(c.in_n, kTile, G) : G;
            if (!(known && c.iter == 1)) {  // the source level's scan tables come from k_seed
                if (e->prof) mg_prof_begin(e, "k_lb_scan", 12.0 * c.in_n, &slot);
                MG_WITH_OFFSETS(e, hipLaunchKernelGGL((k_lb_deg<OT>), dim3(gd), dim3(kBlock), 0, s, e->ctl, rowp, e->q, e->eoff, e->tsum));
                hipLaunchKernelGGL(k_lb_tscan, dim3(1), dim3(1024), 0, s, e->ctl, e->tsum);
                hipLaunchKernelGGL(k_lb_part, dim3(grid_for(maxch + 1, 256, G)), dim3(256), 0, s, e->ctl, e->eoff, e->tsum, e->cstart);
                if (e->prof) mg_prof_end(e, slot);
            }
            const int gp = known ? grid_for(maxch, 1, G) : G;
            if (e->prof) mg_prof_begin(e, "k_push_lb", bytes, &slot);
            if (e->wide) {  // BFS / SSSP only
                if (e->algo == MG_BFS)
                    hipLaunchKernelGGL((k_push_lb<MG_BFS_HOOKS, kBlock, long long>), dim3(gp), dim3(kBlock), 0, s, e->ctl, e->row64, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->eoff, e->tsum, e->cstart);
                else
                    hipLaunchKernelGGL((k_push_lb<MG_SSSP_HOOKS, kBlock, long long>), dim3(gp), dim3(kBlock), 0, s, e->ctl, e->row64, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->eoff, e->tsum, e->cstart);
            } else {
                MG_WITH_BLOCK(e, switch (e->algo) {
                    case MG_BFS:
                        hipLaunchKernelGGL((k_push_lb<MG_BFS_HOOKS, kB>), dim3(gp), dim3(kB), 0, s, e->ctl, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->eoff, e->tsum, e->cstart);
                        break;
                    case MG_SSSP:
                        hipLaunchKernelGGL((k_push_lb<MG_SSSP_HOOKS, kB>), dim3(gp), dim3(kB), 0, s, e->ctl, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->eoff, e->tsum, e->cstart);
                        break;
                    default:
                        hipLaunchKernelGGL((k_push_lb<MG_BC_HOOKS, kB>), dim3(gp), dim3(kB), 0, s, e->ctl, e->row, e->col, e->w, e->level, e->vfc, e->sigma, e->q, e->eoff, e->tsum, e->cstart);
                        break;
                });
            }
            if (e->prof) mg_prof_end(e, slot);
        }
    }
    const int words = (nv + 63) / 64;
    const int gsc = grid_for(words, kBlock / MG_WAVE, G);  // pull / scan grid = number of partials
    int red_parts = 0;  // partials of a pull level, totalled by the control kernel
    if (bu && e->algo == MG_SSSP) {
        // one full relaxation pass: row offsets, col + weight + the source's distance per
        // edge, the vertex's distance write: 12E + 8V (+4)
        const double bytes = 12.0 * (double)e->ne + 4.0 * (nv + 1) + 4.0 * nv;
        if (e->sssp_pb) {
            MG_TRY(mg_ps_sweep(e));
        } else {
        if (e->prof) mg_prof_begin(e, "k_pull_sssp", bytes, &slot);
        MG_WITH_OFFSETS(e, hipLaunchKernelGGL((k_pull_sssp<MG_SSSP_HOOKS, OT>), dim3(grid_for(e->pnch, 1, G)), dim3(kBlock), 0, s, e->ctl,
                                              e->pnch, e->ne, irowp, e->in_col(), e->in_w(), e->pcstart, e->level, e->vfc));
        if (e->prof) mg_prof_end(e, slot);
        }
        MG_WITH_OFFSETS(e, hipLaunchKernelGGL((k_sssp_scan<OT>), dim3(gsc), dim3(kBlock), 0, s, e->ctl, nv, 0, 1, e->vfc, rowp, e->bm[0],
                                              e->part_n, e->part_mf));
        red_parts = gsc;
        bu = false;
    }
    if (bu) {
        const double bytes = known ? (8.0 * nv / 64.0 * 3 + 4.0 * (double)c.mu + 8.0 * nv) : 0.0;
        if (e->prof) mg_prof_begin(e, "k_pull", bytes, &slot);
        const int *orow = e->trow ? e->row : nullptr;
        // rows above pull_split edges go to the chunk workgroups; an engine without
        // chunks (none was that long, or no table was built) walks every row here
        const int cut = e->pull_ch ? e->opt.pull_split : 0x7FFFFFFF;
        const int lanes = e->opt.lane_edges;
        int gpl = gsc;  // k_pull's row workgroups = its partials; the chunk workgroups follow
        const int hn = e->pull_nch;
        if (e->wide) {  // BFS only: a wide graph is symmetric (in-edge rows = out-edge rows)
            const int hb = hn > 0 ? grid_for(hn, kBlock / MG_WAVE, G) : 0;
            hipLaunchKernelGGL((k_pull<MG_BFS_HOOKS, kBlock, long long>), dim3(gpl + hb), dim3(kBlock), 0, s, e->ctl, nv, cut, lanes,
                               e->row64, e->col, (const long long *)nullptr, e->level, e->sigma, e->vis, e->bm[0], e->bm[1], e->part_n,
                               e->part_mf, hb, hn, (const PullChunkT<long long> *)e->pull_ch, e->pull_found, e->pull_sum);
        } else {
            const PullChunk *pch = (const PullChunk *)e->pull_ch;
            MG_WITH_BLOCK(e, gpl = grid_for(words, kB / MG_WAVE, G);
                const int hb = hn > 0 ? grid_for(hn, kB / MG_WAVE, G) : 0;
                if (e->algo == MG_BC)
                    hipLaunchKernelGGL((k_pull<MG_BC_HOOKS, kB>), dim3(gpl + hb), dim3(kB), 0, s, e->ctl, nv, cut, lanes, e->in_row(), e->in_col(), orow, e->level, e->sigma, e->vis, e->bm[0], e->bm[1], e->part_n, e->part_mf, hb, hn, pch, e->pull_found, e->pull_sum);
                else
                    hipLaunchKernelGGL((k_pull<MG_BFS_HOOKS, kB>), dim3(gpl + hb), dim3(kB), 0, s, e->ctl, nv, cut, lanes, e->in_row(), e->in_col(), orow, e->level, e->sigma, e->vis, e->bm[0], e->bm[1], e->part_n, e->part_mf, hb, hn, pch, e->pull_found, e->pull_sum));
        }
        if (hn > 0) {  // the chunked rows' slots, applied once every chunk is in
            const int gf = grid_for(e->pull_nh, kBlock, 64);
            MG_WITH_OFFSETS(e,
                if (e->algo == MG_BC)
                    hipLaunchKernelGGL((k_pull_huge_fin<MG_BC_HOOKS, OT>), dim3(gf), dim3(kBlock), 0, s, e->ctl, e->pull_nh, e->pull_hv, rowp,
                                       e->level, e->sigma, e->vis, e->bm[0], e->bm[1], e->pull_found, e->pull_sum, e->part_n, e->part_mf);
                else
                    hipLaunchKernelGGL((k_pull_huge_fin<MG_BFS_HOOKS, OT>), dim3(gf), dim3(kBlock), 0, s, e->ctl, e->pull_nh, e->pull_hv, rowp,
                                       e->level, e->sigma, e->vis, e->bm[0], e->bm[1], e->pull_found, e->pull_sum, e->part_n, e->part_mf));
        }
        if (e->prof) mg_prof_end(e, slot);
        red_parts = gpl;
    }
    const bool mirror = known && e->mirror;
    hipLaunchKernelGGL(k_ctl, dim3(1), dim3(256), 0, s, e->ctl, e->P, e->lvl_off, e->lvl_mf, mirror ? e->h_ctl : (Ctl *)nullptr,
                       mirror ? ++e->ctl_seq : 0, (const int *)e->part_n, (const long long *)e->part_mf, red_parts);
    if (!known) MG_TRY(launch_after_ctl(e));
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// after the control kernel: the queue of a pull level's frontier (when the next level
// pushes, or for the BC level log), and for SSSP the stamp bitmap a pull level reads
static int launch_after_ctl(mg_engine *e) {
    hipStream_t s = e->stream;
    const int G = max_grid(e);
    const int nv = e->nv;
    const int words = (nv + 63) / 64;
    const int gq = grid_for(words, 32, G);
    hipLaunchKernelGGL(k_bmq_count, dim3(gq), dim3(kBlock), 0, s, e->ctl, words, e->bm[0], e->bm[1], e->part_n);
    hipLaunchKernelGGL(k_bmq_scan, dim3(1), dim3(1024), 0, s, e->ctl, e->part_n, gq);
    hipLaunchKernelGGL(k_bmq_write, dim3(gq), dim3(kBlock), 0, s, e->ctl, words, e->bm[0], e->bm[1], e->part_n, e->q);
    if (e->algo == MG_SSSP)
        MG_WITH_OFFSETS(e, hipLaunchKernelGGL((k_sssp_scan<OT>), dim3(grid_for(words, kBlock / MG_WAVE, G)), dim3(kBlock), 0, s, e->ctl, nv,
                                              -1, 2, e->vfc, rowp, e->bm[0], (int *)nullptr, (long long *)nullptr));
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// wait for the control block of level `seq` in the pinned mirror (written by k_ctl);
// after ~2 ms of spinning, block on the stream instead, which also reports a failed
// kernel, and copy the block back if the mirror still lags
static int wait_ctl(mg_engine *e, int seq) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned k = 0;; ++k) {
        if (__atomic_load_n(&e->h_ctl->seq, __ATOMIC_ACQUIRE) == seq) return MG_OK;
        if ((k & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
    }
    MG_HIP(hipStreamSynchronize(e->stream));
    if (__atomic_load_n(&e->h_ctl->seq, __ATOMIC_ACQUIRE) == seq) return MG_OK;
    MG_HIP(hipMemcpy(e->h_ctl, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    return MG_OK;
}

int mg_frontier_level(mg_engine *e) {
    if (e->done) return MG_OK;
    MG_TRY(launch_level(e, true));
    MG_TRY(wait_ctl(e, e->ctl_seq));
    const Ctl &c = *e->h_ctl;
    if (e->pb16 && c.sat) {  // a distance above 65534 met a 16-bit message: widen them
        e->pb16 = false;
        e->pb_force = true;
    }
    if (c.mode != MODE_DONE && (c.need_queue || (e->algo == MG_SSSP && c.mode == MODE_BU && c.convert)))
        MG_TRY(launch_after_ctl(e));
    e->iter = c.iter;
    e->levels = c.levels;
    e->edges_traversed = c.edges;
    if (c.mode == MODE_DONE) e->done = true;
    return MG_OK;
}

int mg_frontier_capture(mg_engine *e, int levels, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    MG_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = MG_OK;
    // captured SSSP sweeps use 32-bit messages: no host step between them could widen
    // 16-bit ones
    const bool narrow = e->pb16;
    e->pb16 = false;
    for (int i = 0; i < levels && rc == MG_OK; ++i) rc = launch_level(e, false);
    e->pb16 = narrow;
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
    if (e->nv == 0 || L < 1) return MG_OK;  // nothing was reached
    hipStream_t s = e->stream;
    const int G = max_grid(e);
    std::vector<int> off(L + 2, 0);
    std::vector<long long> mf(L + 1, 0);
    // level offsets and degree sums live on the device (written by the control kernel);
    // level 0 is {source}
    MG_HIP(hipMemcpyAsync(off.data(), e->lvl_off, sizeof(int) * (L + 1), hipMemcpyDeviceToHost, s));
    if (L > 1) MG_HIP(hipMemcpyAsync(mf.data() + 1, e->lvl_mf + 1, sizeof(long long) * (L - 1), hipMemcpyDeviceToHost, s));
    int rs[2] = {0, 0};
    MG_HIP(hipMemcpyAsync(rs, e->row + e->source, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    mf[0] = rs[1] - rs[0];
    off[0] = 0;
    off[1] = 1;
    off[L + 1] = off[L];  // the (empty) level below the deepest
    MG_HIP(hipMemcpyAsync(e->lvl_off, off.data(), sizeof(int) * (L + 2), hipMemcpyHostToDevice, s));
    int *nch = e->aux;  // split-vertex chunk counter
    hipLaunchKernelGGL(k_lev8, dim3(grid_for(e->nv, 256, 4096)), dim3(256), 0, s, e->nv, e->level, e->lev8);
    for (int l = L - 1; l >= 0; --l) {
        const int n = off[l + 1] - off[l];
        if (n <= 0) continue;
        int slot = -1;
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, nch);
        // push from level l + 1 when its edges are fewer than an eighth of level l's: a
        // global FP64 atomic per contribution costs several probes of the pull (on
        // RMAT-22 from sampled sources, a factor 2 here chose the push on levels where
        // the pull took half the time; factors 8 and 16 measure alike).  The deepest
        // level has no successors and is finished by the push's final pass alone.  A
        // push walks the in-edges of level l + 1 (the transposed CSR of a directed
        // graph, whose degree sums here are then out-degrees: an estimate).
        const long long mf_next = l + 1 < L ? mf[l + 1] : 0;
        if (mf_next == 0 || 8 * mf_next < mf[l]) {
            const int n1 = off[l + 2] - off[l + 1];
            if (e->prof) mg_prof_begin(e, "k_bc_push", 0.0, &slot);
            if (n1 > 0) {
                hipLaunchKernelGGL(k_bc_push<MG_BC_HOOKS>, dim3(grid_for(n1, kBlock, G)), dim3(kBlock), 0, s, l, e->lvl_off, e->q, e->in_row(),
                                   e->in_col(), e->level, e->lev8, e->bc_cw, e->dacc, e->chunks, nch);
                hipLaunchKernelGGL(k_bc_push_huge, dim3(G), dim3(kBlock), 0, s, l, e->in_row(), e->in_col(), e->level, e->lev8, e->bc_cw,
                                   e->chunks, nch, e->dacc);
            }
            hipLaunchKernelGGL(k_bc_push_fin<MG_BC_HOOKS>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, l, e->lvl_off, e->q, e->sigma, e->dacc,
                               e->delta, e->bc_cw);
            if (e->prof) mg_prof_end(e, slot);
            continue;
        }
        int *wn = e->aux + 1;  // wave-tier list length (the list reuses the degree-scan buffer)
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, wn);
        if (e->prof) mg_prof_begin(e, "k_bc_back", 0.0, &slot);
        const int lanes = e->opt.lane_edges, split = e->opt.back_split, chunk = e->opt.back_chunk;
        MG_WITH_BLOCK(e,
            hipLaunchKernelGGL((k_bc_back<MG_BC_HOOKS, kB>), dim3(grid_for(n, kB, G)), dim3(kB), 0, s, l, e->lvl_off, e->q, e->row, e->col,
                               e->level, e->lev8, e->sigma, e->delta, e->bc_cw, e->chunks, nch, e->eoff, wn, lanes, split, chunk);
            hipLaunchKernelGGL((k_bc_back_wave<MG_BC_HOOKS, kB>), dim3(grid_for(n, kB / MG_WAVE, 16384)), dim3(kB), 0, s, l, e->eoff, wn,
                               e->row, e->col, e->level, e->lev8, e->sigma, e->delta, e->bc_cw));
        if (e->prof) mg_prof_end(e, slot);
        MG_WITH_BLOCK(e,
            hipLaunchKernelGGL((k_bc_back_huge<MG_BC_HOOKS, kB>), dim3(G), dim3(kB), 0, s, l, e->row, e->col, e->level, e->lev8, e->sigma, e->bc_cw,
                               e->chunks, nch, e->dacc, chunk));
        hipLaunchKernelGGL(k_bc_back_fin<MG_BC_HOOKS>, dim3(64), dim3(256), 0, s, e->chunks, nch, e->dacc, e->sigma, e->delta, e->bc_cw);
    }
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// ------------------------------------------------------------ allocation-free BFS
// mg_bfs_noalloc: the BFS of mg_build + mg_run for callers that cannot allocate.  The
// CSR and the level array are the caller's device buffers; everything else the level
// kernels need (queue, degree-scan tables, bitmaps, partial counters, control block)
// lives in one fixed device array, carved as below.  The array's address is a host
// stand-in until hipGetSymbolAddress resolves it on the current device, once per device.
namespace {
constexpr int kNaMaxNv = 1 << 22;                 // vertices
constexpr long long kNaMaxNe = 0x7FFFFFFFLL;      // edges (int32 row offsets)
constexpr int kNaGrid = 2048;                     // workgroups of the grid-stride kernels
constexpr size_t na_al(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kNaQ = na_al(sizeof(int) * ((size_t)kNaMaxNv + 2));                 // queue / level log
constexpr size_t kNaEoff = kNaQ;                                                      // tile-local degree offsets
constexpr size_t kNaTsum = na_al(sizeof(int) * ((size_t)(kNaMaxNv + 2) / kTile + 2)); // tile sums
constexpr size_t kNaCst = na_al(sizeof(int) * ((size_t)lb_max_chunks(kNaMaxNe) + 4)); // chunk owners
constexpr size_t kNaBm = na_al(sizeof(uint64_t) * ((size_t)kNaMaxNv / 64 + 1));      // one bitmap
constexpr size_t kNaPn = na_al(sizeof(int) * (kNaGrid + 8));
constexpr size_t kNaPm = na_al(sizeof(long long) * (kNaGrid + 8));
constexpr size_t kNaCtl = na_al(sizeof(Ctl));
constexpr size_t kNaBytes = kNaQ + kNaEoff + kNaTsum + kNaCst + 3 * kNaBm + kNaPn + kNaPm + kNaCtl;
constexpr int kNaMaxDev = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

static std::mutex g_na_mu;
static void *g_na_ptr[kNaMaxDev];

// device address of `scratch` on the current device (resolved once per device)
static int na_scratch(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return mg_fail(MG_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(err));
        g_na_ptr[dev] = p;
    }
    *out = (unsigned char *)g_na_ptr[dev];
    return MG_OK;
}

extern "C" int mg_bfs_noalloc(int nv, int64_t ne, const int32_t *row, const int32_t *col, int source, int mode,
                              int32_t *level, void *stream) {
    if (nv < 0 || ne < 0) return mg_fail(MG_ERR_ARG, "negative size (nv %d, ne %lld)", nv, (long long)ne);
    if (nv > kNaMaxNv || ne > kNaMaxNe)
        return mg_fail(MG_ERR_LIMIT, "mg_bfs_noalloc takes at most %d vertices and %lld edges (got %d, %lld): its scratch "
                       "is a fixed %zu-byte device array", kNaMaxNv, kNaMaxNe, nv, (long long)ne, kNaBytes);
    if (nv == 0) return MG_OK;
    if (!row || !level || (ne > 0 && !col)) return mg_fail(MG_ERR_ARG, "null device pointer");
    if (source < 0 || source >= nv) return mg_fail(MG_ERR_ARG, "source %d outside [0,%d)", source, nv);
    if (mode != MG_MODE_HYBRID && mode != MG_MODE_SPARSE && mode != MG_MODE_DENSE)
        return mg_fail(MG_ERR_ARG, "unknown mode %d", mode);
    int dev = 0;
    MG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDev) return mg_fail(MG_ERR_LIMIT, "device %d beyond %d", dev, kNaMaxDev);
    unsigned char *base = nullptr;
    MG_TRY(na_scratch(dev, &base));

    // an engine on the stack whose buffers are the caller's and the scratch array's
    mg_engine eng;
    mg_engine *e = &eng;
    Ctl hc{};
    e->algo = MG_BFS;
    e->nv = nv;
    e->ne = ne;
    e->stream = (hipStream_t)stream;
    e->row = const_cast<int *>(row);
    e->col = const_cast<int *>(col);
    e->level = level;
    unsigned char *p = base;
    e->q = (int *)p;             p += kNaQ;
    e->eoff = (int *)p;          p += kNaEoff;
    e->tsum = (int *)p;          p += kNaTsum;
    e->cstart = (int *)p;        p += kNaCst;
    e->bm[0] = (uint64_t *)p;    p += kNaBm;
    e->bm[1] = (uint64_t *)p;    p += kNaBm;
    e->vis = (uint64_t *)p;      p += kNaBm;
    e->part_n = (int *)p;        p += kNaPn;
    e->part_mf = (long long *)p; p += kNaPm;
    e->ctl = (Ctl *)p;
    e->qcap = (long long)nv + 2;
    e->h_ctl = &hc;
    e->mirror = false;
    e->opt.mode = mode;
    e->opt.alpha = kAlpha;
    e->opt.beta = kBeta;
    e->opt.grid_blocks = kNaGrid;
    e->opt.lane_edges = 32;
    e->opt.pull_split = kPullHuge;
    e->opt.block_threads = kBlock;
    Params &P = e->P;
    P.nv = nv;
    P.ne = ne;
    P.log_queue = 1;
    P.forced = mode;
    P.sssp = 0;
    P.need_log = 0;
    P.alpha = e->opt.alpha;
    P.beta = e->opt.beta;

    // reset (resetValue): every level unreached, the source at level 0, queue {source}
    hipStream_t s = e->stream;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, level, (long long)nv, (int)MG_INF);
    int rr[2] = {0, 0};
    MG_HIP(hipMemcpyAsync(rr, row + source, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    hipLaunchKernelGGL(k_seed, dim3(1), dim3(256), 0, s, e->ctl, level, e->q, (double *)nullptr, source, rr[1] - rr[0], ne, 1, e->eoff,
                       e->tsum, e->cstart);
    MG_HIP(hipGetLastError());
    MG_HIP(hipMemcpyAsync(&hc, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    // levels: the control block comes back by a copy into host memory after each level
    for (;;) {
        MG_TRY(launch_level(e, true));
        MG_HIP(hipMemcpyAsync(&hc, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        MG_HIP(hipStreamSynchronize(s));
        if (hc.mode == MODE_DONE) break;
        if (hc.need_queue) MG_TRY(launch_after_ctl(e));
    }
    MG_HIP(hipStreamSynchronize(s));
    return MG_OK;
}
