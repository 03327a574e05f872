// Batched multi-source BFS / BC / SSSP: up to 64 sources run together, one bit per
// source packed per vertex (visited word, frontier words) and one slot per source per
// vertex for the outputs, stored vertex-major ([v][64]) so that the 64 lanes of a
// wavefront are the 64 sources of one vertex and every per-vertex row is one
// contiguous 256 B (int/float) or 512 B (double) access.
//
//   BFS / BC forward, per level:
//     push  (few frontier vertices): a wave per frontier vertex u walks its edges;
//           bits = F[u] & ~visited[v] are or-ed into the next frontier word of v
//           (atomicOr), BC adds sigma[u][s] into sigma[v][s] for every bit s;
//     pull  (many):  every vertex not yet reached by all sources or-s the frontier
//           words of its neighbours (BFS stops as soon as the missing bits are
//           covered); BC, lanes = sources, sums sigma of the parents of each new bit;
//     the new frontier (vertex, word) is compacted into a level log (count, scan,
//     ordered write: no contended counters).
//   BC backward, per level d from the deepest: the words of level d+1 are scattered
//     into a dense array, and a wave per logged (v, word) of level d sums, lane s,
//     sigma[v][s] / sigma[w][s] * (1 + delta[w][s]) over neighbours w whose level-d+1
//     word has bit s (Neumaier-compensated FP32, terms split in two floats).
//   SSSP: Bellman-Ford by pull sweeps; changed[v] holds one bit per source whose
//     distance at v dropped in the last sweep; lane s relaxes dist[v][s] from the
//     neighbours whose changed bit s is set.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>
#include "engine.hpp"
#include "hook_calls.hpp"
#include "mg_device.hpp"

using namespace mg;

namespace {

constexpr int kBB = 256;


int grid_for(long long items, int per_block, int cap) {
    long long g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

__global__ void k_fill_u64(uint64_t *p, long long n, uint64_t v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32b(int *p, long long n, int v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64b(double *p, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32b(float *p, long long n, float v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

// seed: source i at vertex src[i] in slot (bit) slot[i], or bit i without a slot list.
// Several sources may share a vertex or a slot (a slot with several sources runs from
// their set, which for sources in disjoint components is one run per component).
__global__ void k_ms_seed(int n, const int *__restrict__ src, const int *__restrict__ slot, uint64_t *__restrict__ vis,
                          uint64_t *__restrict__ F, int *__restrict__ lev, double *__restrict__ sigma, int *__restrict__ dist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = slot ? slot[i] : i;
    const int v = src[i];
    atomicOr((unsigned long long *)&vis[v], 1ull << s);
    if (F) atomicOr((unsigned long long *)&F[v], 1ull << s);
    if (lev) lev[(long long)v * 64 + s] = 0;
    if (sigma) sigma[(long long)v * 64 + s] = 1.0;
    if (dist) dist[(long long)v * 64 + s] = 0;
}

// ---- compaction of nonzero words into (vertex, word) entries -----------------------
__global__ __launch_bounds__(kBB) void k_nz_count(int nv, const uint64_t *__restrict__ W, int *__restrict__ part) {
    const int per = (nv + gridDim.x - 1) / gridDim.x;
    const int v0 = blockIdx.x * per, v1 = min(nv, v0 + per);
    long long c = 0;
    for (int v = v0 + threadIdx.x; v < v1; v += kBB) c += W[v] != 0;
    __shared__ long long s_tmp[kBB / MG_WAVE];
    c = block_sum<kBB / MG_WAVE>(c, s_tmp);
    if (threadIdx.x == 0) part[blockIdx.x] = (int)c;
}

__global__ __launch_bounds__(1024) void k_nz_scan(int *__restrict__ part, int nparts, long long *__restrict__ total) {
    __shared__ int s_tmp[1024 / MG_WAVE + 1];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int b = 0; b < nparts; b += 1024) {
        const int i = b + threadIdx.x;
        const int v = i < nparts ? part[i] : 0;
        int t;
        const int pre = block_excl_scan<1024 / MG_WAVE>(v, s_tmp, &t);
        const int carry = s_carry;
        if (i < nparts) part[i] = carry + pre;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + t;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s_carry;
}

// (entries past `cap` are dropped: the host sized the log from the level's count)
__global__ __launch_bounds__(kBB) void k_nz_write(int nv, const uint64_t *__restrict__ W, const int *__restrict__ part,
                                                  long long base, long long cap, int *__restrict__ outv, uint64_t *__restrict__ outw) {
    const int per = (nv + gridDim.x - 1) / gridDim.x;
    const int v0 = blockIdx.x * per, v1 = min(nv, v0 + per);
    long long off = base + part[blockIdx.x];
    __shared__ int s_tmp[kBB / MG_WAVE + 1];
    for (int vb = v0; vb < v1; vb += kBB) {
        const int v = vb + threadIdx.x;
        const uint64_t w = v < v1 ? W[v] : 0ull;
        int t;
        const int pre = block_excl_scan<kBB / MG_WAVE>(w != 0 ? 1 : 0, s_tmp, &t);
        if (w && off + pre < cap) { outv[off + pre] = v; if (outw) outw[off + pre] = w; }
        off += t;
        __syncthreads();
    }
}

// ---- chunked work ------------------------------------------------------------------
// Every kernel below walks a static list of edge chunks: a vertex's edges are cut
// into chunks of at most kCH, so a hub is served by many wavefronts.  A vertex with
// one chunk writes its results directly; the chunks of a split vertex merge theirs
// with atomics (or, and, add, min) into state that starts the pass at its identity.
struct Chunk { int v, p0, p1, multi; };
constexpr int kCH = 1024;  // default edges per chunk (mg_options.batch_chunk)
#ifndef MG_HIT_STEP
#define MG_HIT_STEP 8
#endif
constexpr int kHitStep = MG_HIT_STEP;  // matched edges whose row gathers a wave issues together (BC)
#ifndef MG_BACK_HIT_STEP
#define MG_BACK_HIT_STEP 16
#endif
constexpr int kBackHitStep = MG_BACK_HIT_STEP;  // the same for the batched BC backward (4 / 8 / 16 / 32: 19.8 / 19.3 / 19.1 / 19.4 ms)
#ifndef MG_SSSP_HIT_STEP
#define MG_SSSP_HIT_STEP 4
#endif
// the same for the batched SSSP sweeps, whose 256-byte distance rows congest the memory
// system at 8 gathers in flight per wave (RMAT-22, 64 sources: 32.0 ms at 8, 27.4 at 4)
constexpr int kSsspHitStep = MG_SSSP_HIT_STEP;

// Wave w walks chunks w, w + nw, w + 2 nw, ... (nw waves in the grid) 64 at a time: each
// lane loads one chunk and its word (word_of(chunk); 0 = nothing to do), and the wave
// then runs body(chunk, word) for the nonzero ones in turn, from lane broadcasts.  A
// level whose frontier touches few chunks costs one load round per 64 chunks instead
// of one per chunk; the stride keeps a hub's consecutive chunks on different waves.
template <class WordOf, class Body>
__device__ __forceinline__ void for_active_chunks(long long nch, const Chunk *__restrict__ ch, WordOf word_of, Body body) {
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long base = wave; base < nch; base += nw * MG_WAVE) {
        const long long c = base + (long long)lane() * nw;
        Chunk k{0, 0, 0, 0};
        uint64_t word = 0ull;
        if (c < nch) {
            k = ch[c];
            word = word_of(k);
        }
        uint64_t act = ballot(word != 0ull);
        while (act) {
            const int l = __builtin_ctzll(act);
            act &= act - 1;
            const Chunk kl{readlane(k.v, l), readlane(k.p0, l), readlane(k.p1, l), readlane(k.multi, l)};
            body(kl, bcast64(word, l));
        }
    }
}

// Device-side level control (BFS and SSSP runs): the level kernels read the current
// level from here, so K levels can be captured in one hipGraph and replayed past the
// end of the run (every kernel returns at once when `done` is set).
struct BCtl {
    int it;       // level (BFS) or sweep (SSSP) being processed
    int done;
    int push;     // dense levels: push over all chunks (1) or pull (0)
    int sparse;   // this level walks the frontier's queue instead of every chunk
    int compact;  // build the current frontier's queue first (dense -> sparse)
    int fill;     // zero the whole next frontier first (it is not known clean)
    int prev_sparse;  // the previous level was sparse (its frontier buffer was cleared)
    int need_full;    // the sparse-only graph met a level it cannot run: replay the full graph
    int qn[2];    // queue lengths of F[0], F[1]
    int qvalid[2];    // the queue of F[b] lists its nonzero words
    int ndense, nsparse, npersist;  // levels run: dense, sparse, sparse by persistent launches
    int seq;      // host mirror only: the host check whose block it holds
    int thr;      // SSSP: sparse sweeps relax the (vertex, source) pairs below this distance
    long long n, mf;  // size and degree sum of the last produced frontier
    unsigned long long all;  // the slots in use
    int err;      // a persistent launch gave up a bounded wait: its results are not valid
    int big;      // the last persistent launch ran no level: the frontier is too large for one XCD
};

// A level kernel gets two frontier buffers (A, B).  Without a control block A is the
// current frontier and B the next one; with one they are F[0], F[1] and the level's
// parity picks: current = F[(it - 1) & 1], next = F[it & 1].
__device__ __forceinline__ void cur_next(const BCtl *ctl, uint64_t *A, uint64_t *B, uint64_t **cur, uint64_t **next) {
    if (!ctl) { *cur = A; *next = B; return; }
    const int it = ctl->it;
    *cur = ((it - 1) & 1) ? B : A;
    *next = (it & 1) ? B : A;
}
// dense-level kernels; want: -1 any dense level, 1 push levels only, 0 pull levels only
__device__ __forceinline__ bool level_runs(const BCtl *ctl, int want) {
    if (!ctl) return true;
    if (ctl->done || ctl->sparse || ctl->need_full) return false;
    return want < 0 || ctl->push == want;
}

// the frontier a workgroup produced, as partials: its size, its degree sum and its
// largest degree (a level may walk a queue only when no frontier vertex is split
// into several chunks, so no wave walks more than one chunk's edges)
__device__ __forceinline__ void put_parts(int cnt, long long mf, int mx, int *__restrict__ part_n,
                                          long long *__restrict__ part_e, int *__restrict__ part_x) {
    __shared__ long long s_tmp[kBB / MG_WAVE];
    const long long tn = block_sum<kBB / MG_WAVE>((long long)cnt, s_tmp);
    const long long te = block_sum<kBB / MG_WAVE>(mf, s_tmp);
    const long long tx = block_max<kBB / MG_WAVE>((long long)mx, s_tmp);
    if (threadIdx.x == 0) { part_n[blockIdx.x] = (int)tn; part_e[blockIdx.x] = te; part_x[blockIdx.x] = (int)tx; }
}

// push: a wave per chunk whose vertex u is in the frontier.  BFS: lanes over edges,
// new bits or-ed into Fn.  BC: lanes = sources, edges in turn, sigma[u][s] added
// into sigma[w][s] for every new bit s.
template <bool kBC>
__global__ __launch_bounds__(kBB) void k_mb_push(const BCtl *ctl, long long nch, const Chunk *__restrict__ ch, const int *__restrict__ col,
                                                 uint64_t *FA, const uint64_t *__restrict__ vis,
                                                 uint64_t *FB, double *__restrict__ sigma) {
    if (!level_runs(ctl, 1)) return;
    uint64_t *F, *Fn;
    cur_next(ctl, FA, FB, &F, &Fn);
    __shared__ int s_w[kBB];
    __shared__ uint64_t s_m[kBB];
    for_active_chunks(nch, ch, [&](const Chunk &k) { return F[k.v]; }, [&](const Chunk &k, const uint64_t fu) {
        if (!kBC) {
            for (int p = k.p0 + lane(); p < k.p1; p += MG_WAVE) {
                const int w = col[p];
                const uint64_t m = fu & ~vis[w];
                if (m) atomicOr((unsigned long long *)&Fn[w], (unsigned long long)m);
            }
        } else {
            // lanes load 64 edges and their visited words at once and mark the next
            // frontier; then the wave walks the edges with new bits, four at a time,
            // lanes = sources adding sigma[u][s] into sigma[w][s]
            const int wl = threadIdx.x & ~(MG_WAVE - 1);
            // the lane's message (BCHooks::get_partial_result: its path count)
            typename MG_BC_HOOKS::V1T lm;
            double su = 0.0;
            if ((fu >> lane()) & 1ull) MG_BC_HOOKS::get_partial_result(&lm, 0, &su, sigma[(long long)k.v * 64 + lane()]);
            for (int p0 = k.p0; p0 < k.p1; p0 += MG_WAVE) {
                const int p = p0 + lane();
                int w = 0;
                uint64_t m = 0;
                if (p < k.p1) { w = col[p]; m = fu & ~vis[w]; }
                if (m) atomicOr((unsigned long long *)&Fn[w], (unsigned long long)m);
                uint64_t hits = ballot(m != 0ull);
                if (!hits) continue;
                wave_sync();
                s_w[threadIdx.x] = w;
                s_m[threadIdx.x] = m;
                wave_sync();
                while (hits) {
                    int j[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) { j[q] = hits ? __builtin_ctzll(hits) : -1; hits &= hits - 1; }
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (j[q] >= 0 && ((s_m[wl + j[q]] >> lane()) & 1ull)) atomicAdd(&sigma[(long long)s_w[wl + j[q]] * 64 + lane()], su);
                }
            }
        }
    });
}

// pull, BFS: a wave per chunk of a vertex still missing bits; lanes or the frontier
// words of 64 neighbours at a time and stop once the missing bits are covered
#ifndef MG_HEAVY_STEP
#define MG_HEAVY_STEP 1
#endif
constexpr int kHeavyStep = MG_HEAVY_STEP;  // rounds of 64 in-edges the heavy BFS pull loads together (1 / 2 / 4: 1.90 / 1.97 / 2.05 ms per 64-source run: the early exit wastes the extra rounds)

__global__ __launch_bounds__(kBB) void k_mb_pull_bfs(const BCtl *ctl, long long nch, const Chunk *__restrict__ ch, uint64_t all,
                                                     const int *__restrict__ col, uint64_t *FA,
                                                     const uint64_t *__restrict__ vis, uint64_t *FB) {
    if (!level_runs(ctl, 0)) return;
    uint64_t *F, *Fn;
    cur_next(ctl, FA, FB, &F, &Fn);
    if (ctl) all = ctl->all;
    for_active_chunks(nch, ch, [&](const Chunk &k) { return all & ~vis[k.v]; }, [&](const Chunk &k, const uint64_t need) {
        uint64_t acc = 0;
        for (int p0 = k.p0; p0 < k.p1; p0 += kHeavyStep * MG_WAVE) {  // kHeavyStep rounds of loads in flight
            int u[kHeavyStep];
#pragma unroll
            for (int j = 0; j < kHeavyStep; ++j) {
                const int p = p0 + j * MG_WAVE + lane();
                u[j] = p < k.p1 ? ld_stream(&col[p]) : -1;
            }
            uint64_t f = 0ull;
#pragma unroll
            for (int j = 0; j < kHeavyStep; ++j) f |= u[j] >= 0 ? F[u[j]] & need : 0ull;
            acc |= wave_or(f);
            if (acc == need) break;
        }
        if (acc && lane() == 0) {
            if (k.multi) atomicOr((unsigned long long *)&Fn[k.v], (unsigned long long)acc);
            else Fn[k.v] = acc;
        }
    });
}

#ifndef MG_LIGHT_STEP
#define MG_LIGHT_STEP 2
#endif
// in-edges a lane of the light BFS pull loads together (RMAT-22, 64 sources, 1 / 2 / 4
// / 8: 1.93 / 1.89 / 1.90 / 1.93 ms per run)
constexpr int kLightStep = MG_LIGHT_STEP;

// pull, BFS, vertices of at most `light` in-edges (mg_options.batch_light, 32 by
// default): a lane per vertex walks its row
// alone (64 consecutive ids per wave, so the row loads of a step are close together)
// and stops once its missing bits are covered; k_mb_pull_bfs takes the heavier ones'
// chunks.  A wave per light vertex would leave most lanes idle and pay a cross-lane
// reduction per 64 edges for a handful of edges.
__global__ __launch_bounds__(kBB) void k_mb_pull_bfs_light(const BCtl *ctl, int nv, int light, const int *__restrict__ row,
                                                           const int *__restrict__ col, uint64_t all, uint64_t *FA,
                                                           const uint64_t *__restrict__ vis, uint64_t *FB) {
    if (!level_runs(ctl, 0)) return;
    uint64_t *F, *Fn;
    cur_next(ctl, FA, FB, &F, &Fn);
    if (ctl) all = ctl->all;
    for (int v = blockIdx.x * kBB + threadIdx.x; v < nv; v += gridDim.x * kBB) {
        const int p0 = row[v], p1 = row[v + 1];
        if (p1 == p0 || p1 - p0 > light) continue;
        const uint64_t need = all & ~vis[v];
        if (!need) continue;
        uint64_t acc = 0;
        // kLightStep edges per step: their column and frontier-word loads in flight together
        for (int p = p0; p < p1; p += kLightStep) {
            int u[kLightStep];
#pragma unroll
            for (int j = 0; j < kLightStep; ++j) u[j] = p + j < p1 ? ld_stream(&col[p + j]) : -1;
#pragma unroll
            for (int j = 0; j < kLightStep; ++j) acc |= u[j] >= 0 ? F[u[j]] & need : 0ull;
            if (acc == need) break;
        }
        if (acc) Fn[v] = acc;
    }
}

// pull, BC: lanes = sources; neighbours staged 64 at a time in LDS, then lane s sums
// sigma[u][s] over neighbours u whose frontier word has bit s
// Path counts per lane (lane = source) through BCHooks: every parent in a source's
// frontier sends get_partial_result(level it - 1, its count), combined with
// accumulate_nonAtomics; the vertex's new bits are the lanes that received any.
template <class H>
__global__ __launch_bounds__(kBB) void k_mb_pull_bc(int it, long long nch, const Chunk *__restrict__ ch, uint64_t all,
                                                    const int *__restrict__ col, const uint64_t *__restrict__ F,
                                                    const uint64_t *__restrict__ vis, uint64_t *__restrict__ Fn,
                                                    double *__restrict__ sigma) {
    __shared__ int s_u[kBB];
    __shared__ uint64_t s_f[kBB];
    const int wl = threadIdx.x & ~(MG_WAVE - 1);
    for_active_chunks(nch, ch, [&](const Chunk &k) { return all & ~vis[k.v]; }, [&](const Chunk &k, const uint64_t need) {
        uint64_t acc = 0;
        Acc<H> a;
        acc_init(a);
        for (int p0 = k.p0; p0 < k.p1; p0 += MG_WAVE) {
            const int p = p0 + lane();
            int u = 0;
            uint64_t f = 0;
            if (p < k.p1) { u = ld_stream(&col[p]); f = F[u] & need; }
            const uint64_t any = wave_or(f);
            if (!any) continue;
            acc |= any;
            wave_sync();
            s_u[threadIdx.x] = u;
            s_f[threadIdx.x] = f;
            wave_sync();
            uint64_t hits = ballot(f != 0ull);  // only the parents in some source's frontier
            while (hits) {  // kHitStep parents per step: their row gathers overlap
                int j[kHitStep];
#pragma unroll
                for (int q = 0; q < kHitStep; ++q) { j[q] = hits ? __builtin_ctzll(hits) : -1; hits &= hits - 1; }
                double t[kHitStep];
                bool on[kHitStep];
#pragma unroll
                for (int q = 0; q < kHitStep; ++q) {
                    on[q] = j[q] >= 0 && ((s_f[wl + j[q]] >> lane()) & 1ull);
                    t[q] = on[q] ? sigma[(long long)s_u[wl + j[q]] * 64 + lane()] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < kHitStep; ++q)
                    if (on[q]) acc_edge<H>(a, it - 1, t[q], 0, it);
            }
        }
        if (!acc) return;
        if ((acc >> lane()) & 1ull) {
            const long long kv = (long long)k.v * 64 + lane();
            if (k.multi) atomicAdd(&sigma[kv], a.s);
            else sigma[kv] = a.s;
        }
        if (lane() == 0) {
            if (k.multi) atomicOr((unsigned long long *)&Fn[k.v], (unsigned long long)acc);
            else Fn[k.v] = acc;
        }
    });
}

// after a push or pull level: keep only new bits, mark them visited, record levels,
// and leave the new frontier's size and degree sum as per-workgroup partials
// With a dense-level log (dlog: dcap words per vertex, one per dense level; BFS runs
// under the control block), a dense level records each vertex's new bits as one word
// in its slot, and k_mb_expand writes every level row once, whole, after the run:
// written level by level, a row takes several partial stores, each a read-modify-write
// of its lines in memory.  Without a log, or past its dcap slots, the rows are written
// here.
__global__ __launch_bounds__(kBB) void k_mb_apply(const BCtl *ctl, int nv, const int *__restrict__ row, uint64_t *__restrict__ vis,
                                                  uint64_t *FA, uint64_t *FB, int *__restrict__ lev, int it,
                                                  int *__restrict__ part_n, long long *__restrict__ part_e, int *__restrict__ part_x,
                                                  uint64_t *__restrict__ dlog, int *__restrict__ dlev, int dcap) {
    if (!level_runs(ctl, -1)) return;
    uint64_t *F, *Fn;
    cur_next(ctl, FA, FB, &F, &Fn);
    if (ctl) it = ctl->it;
    const int slot = ctl && dlog ? ctl->ndense - 1 : dcap;  // k_mb_ctl counted this level
    uint64_t *lw = slot >= 0 && slot < dcap ? dlog + (long long)slot * (nv + 1) : nullptr;
    if (lw && blockIdx.x == 0 && threadIdx.x == 0) dlev[slot] = it;
    int cnt = 0, mx = 0;
    long long mf = 0;
    // a lane per vertex for the words; the level rows are then written by the whole
    // wave, one vertex at a time (lanes = sources: one coalesced 256-byte row store),
    // not by each lane walking its own row (64 rows, i.e. 64 cache lines, per store)
    for (int v0 = blockIdx.x * kBB + (threadIdx.x & ~(MG_WAVE - 1)); v0 < nv; v0 += gridDim.x * kBB) {
        const int v = v0 + lane();
        uint64_t nw = 0;
        if (v < nv) {
            const uint64_t f = Fn[v];
            if (f) {
                const uint64_t vv = vis[v];
                nw = f & ~vv;
                if (nw != f) Fn[v] = nw;
                if (nw) {
                    vis[v] = vv | nw;
                    ++cnt;
                    const int dg = row[v + 1] - row[v];
                    mf += dg;
                    mx = max(mx, dg);
                }
            }
        }
        if (lw) {
            if (v < nv) lw[v] = nw;
            continue;
        }
        uint64_t todo = ballot(nw != 0ull);
        while (todo) {
            const int l = __builtin_ctzll(todo);
            todo &= todo - 1;
            const uint64_t m = (uint64_t)__shfl((long long)nw, l);
            if ((m >> lane()) & 1ull) lev[(long long)(v0 + l) * 64 + lane()] = it;
        }
    }
    put_parts(cnt, mf, mx, part_n, part_e, part_x);
}

// After a BFS run with a dense-level log: every level row written once (lanes =
// sources): a bit set at a logged dense level takes that level; a bit never visited
// is MG_INF; another visited bit keeps the value its sparse level, seed or unlogged
// dense level stored (its lane does not store).  A lane per vertex loads the words
// (coalesced); the wave then writes the 64 rows one by one from lane-broadcast copies.
constexpr int kDenseLog = 16;  // logged dense levels per BFS run
__global__ __launch_bounds__(kBB) void k_mb_expand(int nv, int nd, const uint64_t *__restrict__ vis,
                                                   const uint64_t *__restrict__ dlog, const int *__restrict__ dlev,
                                                   int *__restrict__ lev) {
    int dl[kDenseLog];
#pragma unroll
    for (int d = 0; d < kDenseLog; ++d) dl[d] = d < nd ? dlev[d] : 0;
    for (int v0 = blockIdx.x * kBB + (threadIdx.x & ~(MG_WAVE - 1)); v0 < nv; v0 += gridDim.x * kBB) {
        const int v = v0 + lane();
        const uint64_t vw = v < nv ? vis[v] : 0ull;
        uint64_t dw[kDenseLog];
#pragma unroll
        for (int d = 0; d < kDenseLog; ++d) dw[d] = (d < nd && v < nv) ? dlog[(long long)d * (nv + 1) + v] : 0ull;
        const int nl = min(MG_WAVE, nv - v0);
        for (int l = 0; l < nl; ++l) {  // l is wave-uniform: broadcasts by readlane
            const uint64_t bv = bcast64(vw, l);
            int lv = -1;
#pragma unroll
            for (int d = 0; d < kDenseLog; ++d) {
                if (d < nd) {  // uniform
                    const uint64_t w = bcast64(dw[d], l);
                    if (lv < 0 && ((w >> lane()) & 1ull)) lv = dl[d];
                }
            }
            // a visited bit of no logged level already holds its level: that lane
            // stores nothing (no read of the row)
            const bool visited = (bv >> lane()) & 1ull;
            if (!visited || lv >= 0) lev[(long long)(v0 + l) * 64 + lane()] = visited ? lv : (int)MG_INF;
        }
    }
}

// size and degree sum of the nonzero words of the next frontier (without a control
// block: of WB)
__global__ __launch_bounds__(kBB) void k_mb_count(const BCtl *ctl, int nv, const int *__restrict__ row, uint64_t *WA, uint64_t *WB,
                                                  int *__restrict__ part_n, long long *__restrict__ part_e, int *__restrict__ part_x) {
    if (!level_runs(ctl, -1)) return;
    uint64_t *cur, *W;
    cur_next(ctl, WA, WB, &cur, &W);
    int cnt = 0, mx = 0;
    long long mf = 0;
    for (int v = blockIdx.x * kBB + threadIdx.x; v < nv; v += gridDim.x * kBB)
        if (W[v]) { const int dg = row[v + 1] - row[v]; ++cnt; mf += dg; mx = max(mx, dg); }
    put_parts(cnt, mf, mx, part_n, part_e, part_x);
}

// zero the next frontier (a fill kernel: graphs record no memset nodes)
__global__ void k_mb_fill_next(const BCtl *ctl, uint64_t *FA, uint64_t *FB, long long n) {
    if (ctl->done || !ctl->fill || ctl->need_full) return;
    uint64_t *cur, *next;
    cur_next(ctl, FA, FB, &cur, &next);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) next[i] = 0ull;
}

__global__ void k_mb_ctl_init(BCtl *ctl, unsigned long long all) {
    BCtl c{};
    c.all = all;
    *ctl = c;
}

// between levels: total the partials of the last produced frontier; stop when it is
// empty, else open the next level and choose its direction (push while the frontier's
// degree sum is below E / alpha) and whether it walks a queue (sparse: push levels
// whose vertices plus edges are below (nv + E) / beta and whose vertices all fit one
// chunk, so per-level work follows the frontier, not the graph; high-diameter members
// of an uneven batch run thousands of such levels)
struct LevelRule { long long ne; int nv; float alpha, beta; int forced, light, delta; };

__device__ __forceinline__ void decide(BCtl &c, long long a, long long b, long long x, const LevelRule &r) {
    c.n = a;
    c.mf = b;
    c.need_full = 0;
    if (a == 0) { c.done = 1; return; }
    c.it += 1;
    c.thr = r.delta <= 0 ? 0x7FFFFFFF : (c.thr > 0x7FFFFFFF - r.delta ? 0x7FFFFFFF : c.thr + r.delta);
    int push = (double)b < (double)r.ne / r.alpha;
    if (r.forced == MG_MODE_SPARSE) push = 1;
    if (r.forced == MG_MODE_DENSE) push = 0;
    c.push = push;
    const int cur = (c.it - 1) & 1, next = c.it & 1;
    const int sparse = r.forced != MG_MODE_DENSE && push && (double)(a + b) * r.beta < (double)r.nv + (double)r.ne && x <= r.light;
    c.compact = sparse && !c.qvalid[cur];
    c.fill = !sparse || !c.prev_sparse;
    c.sparse = sparse;
    c.qn[next] = 0;
    c.qvalid[next] = sparse;
    c.prev_sparse = sparse;
}

__device__ __forceinline__ void sum_partials(const int *pn, const long long *pe, const int *px, int nparts, long long *a,
                                             long long *b, long long *c) {
    __shared__ long long s_tmp[256 / MG_WAVE];
    long long x = 0, y = 0, z = 0;
    for (int i = threadIdx.x; i < nparts; i += 256) { x += pn[i]; y += pe[i]; z = max(z, (long long)px[i]); }
    *a = block_sum<256 / MG_WAVE>(x, s_tmp);
    *b = block_sum<256 / MG_WAVE>(y, s_tmp);
    *c = block_max<256 / MG_WAVE>(z, s_tmp);
}

__global__ void k_mb_ctl(BCtl *ctl, const int *__restrict__ pn, const long long *__restrict__ pe, const int *__restrict__ px,
                         int nparts, LevelRule r) {
    if (ctl->done) return;
    long long a, b, x;
    sum_partials(pn, pe, px, nparts, &a, &b, &x);
    if (threadIdx.x != 0) return;
    BCtl c = *ctl;
    decide(c, a, b, x, r);
    if (!c.done) { if (c.sparse) ++c.nsparse; else ++c.ndense; }
    *ctl = c;
}

// the host's pinned copy of the control block after a replay: the block, then (after a
// system-scope fence) the check's sequence number, on which the host spins
__global__ void k_mb_mirror(const BCtl *ctl, BCtl *mirror, int seq) {
    BCtl c = *ctl;
    c.seq = seq - 1;
    *mirror = c;
    __threadfence_system();
    __hip_atomic_store(&mirror->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- sparse levels (queue of the frontier's vertices) -------------------------------
// dense -> sparse: the current frontier's queue by count / scan / ordered write over
// its nonzero words (the k_nz_* pattern), gated on ctl->compact
__global__ __launch_bounds__(kBB) void k_mbs_qcount(const BCtl *ctl, int nv, uint64_t *FA, uint64_t *FB, int *__restrict__ part) {
    if (ctl->done || !ctl->compact) return;
    const uint64_t *W = ((ctl->it - 1) & 1) ? FB : FA;
    const int per = (nv + gridDim.x - 1) / gridDim.x;
    const int v0 = blockIdx.x * per, v1 = min(nv, v0 + per);
    long long c = 0;
    for (int v = v0 + threadIdx.x; v < v1; v += kBB) c += W[v] != 0;
    __shared__ long long s_tmp[kBB / MG_WAVE];
    c = block_sum<kBB / MG_WAVE>(c, s_tmp);
    if (threadIdx.x == 0) part[blockIdx.x] = (int)c;
}
__global__ __launch_bounds__(1024) void k_mbs_qscan(BCtl *ctl, int *__restrict__ part, int nparts) {
    if (ctl->done || !ctl->compact) return;
    __shared__ int s_tmp[1024 / MG_WAVE + 1];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int b = 0; b < nparts; b += 1024) {
        const int i = b + threadIdx.x;
        const int v = i < nparts ? part[i] : 0;
        int t;
        const int pre = block_excl_scan<1024 / MG_WAVE>(v, s_tmp, &t);
        const int carry = s_carry;
        if (i < nparts) part[i] = carry + pre;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + t;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int cur = (ctl->it - 1) & 1;
        ctl->qn[cur] = s_carry;
        ctl->qvalid[cur] = 1;
    }
}
__global__ __launch_bounds__(kBB) void k_mbs_qwrite(const BCtl *ctl, int nv, uint64_t *FA, uint64_t *FB,
                                                    const int *__restrict__ part, int *QA, int *QB) {
    if (ctl->done || !ctl->compact) return;
    const int cur = (ctl->it - 1) & 1;
    const uint64_t *W = cur ? FB : FA;
    int *Q = cur ? QB : QA;
    const int per = (nv + gridDim.x - 1) / gridDim.x;
    const int v0 = blockIdx.x * per, v1 = min(nv, v0 + per);
    long long off = part[blockIdx.x];
    __shared__ int s_tmp[kBB / MG_WAVE + 1];
    for (int vb = v0; vb < v1; vb += kBB) {
        const int v = vb + threadIdx.x;
        const uint64_t w = v < v1 ? W[v] : 0ull;
        int t;
        const int pre = block_excl_scan<kBB / MG_WAVE>(w != 0 ? 1 : 0, s_tmp, &t);
        if (w) Q[off + pre] = v;
        off += t;
        __syncthreads();
    }
}

// Sparse levels take the queue 64 entries per wave, one per lane.  A lane whose vertex
// has little work (kLaneWork edges, or bit-edge pairs for SSSP) walks it alone, the
// wave stepping all such lanes together for the largest of their trip counts so that
// queue appends stay wave-aggregated; the heavier vertices of the 64 are then walked
// one at a time by the whole wave.  Small tails (grids, RMAT fringes: degree a few)
// thus keep 64 independent load chains in flight per wave instead of one.
constexpr int kLaneWork = 16;

// BFS: new bits are claimed on the visited word itself (atomicOr returns the bits
// that were already set), so each (vertex, source) pair gets its level exactly once,
// and the claimed bits go into the next frontier word; a vertex whose word was empty
// joins the next queue.  Every queued vertex is queued once and has at most one chunk,
// so its consumed word is cleared by the lane that owns it.  The thread counts the new
// frontier's vertices, degree sum and largest degree.
// kDup (the sparse-only graph's levels): every discoverer that claims new bits queues
// the vertex, without waiting for the next word's old value (one dependent device
// atomic less per level).  A vertex found by several discoverers is then queued
// several times: each copy walks the whole word (or finds it cleared by another), and
// the claims stay exclusive, so results are unchanged and copies do not multiply.
// Appends past the queue's capacity are dropped (the words still hold every bit) and
// the next level rebuilds the queue from the words instead.
template <bool kDup = false>
__device__ __forceinline__ bool bfs_visit(int it, int w, uint64_t fu, uint64_t *Fn, uint64_t *vis, int *__restrict__ lev) {
    const uint64_t m = fu & ~vis[w];
    if (!m) return false;
    const uint64_t nb = m & ~(uint64_t)atomicOr((unsigned long long *)&vis[w], (unsigned long long)m);
    if (!nb) return false;
    bool add = true;
    if (kDup) atomicOr((unsigned long long *)&Fn[w], (unsigned long long)nb);
    else add = atomicOr((unsigned long long *)&Fn[w], (unsigned long long)nb) == 0ull;
    uint64_t bits = nb;
    while (bits) { const int sl = __builtin_ctzll(bits); lev[(long long)w * 64 + sl] = it; bits &= bits - 1; }
    return add;
}

// Next-queue appends of a sparse level.  With a stage (the workgroup's list in LDS), a
// wave's appends take their slots there by one LDS atomic, and the workgroup moves the
// list to the queue at the end of the level with ONE atomic on the queue counter:
// atomics per wave on that one counter queue up behind each other.  Entries past the
// stage's capacity, and every append without a stage, go to the queue at once.  SSSP
// levels stage (a wave appends once per bit-edge step: uneven batch 194 -> 189 ms);
// BFS levels do not (one append per wave and step; staged, 55.7 -> 58.5 ms).
constexpr int kQStage = 4096;
struct QStage {
    int n;
    int v[kQStage];
};
struct QOut {
    int *Qn;     // the next queue
    int *qn;     // its length counter
    QStage *st;  // the workgroup's stage, or null
    int cap = 0x7FFFFFFF;  // entries the queue holds (appends past it are dropped: see sparse_bfs)
};

// atomicAdd on an int known to be in LDS (the cast keeps it a ds_add, where a pointer
// that may also be global would make it a flat atomic)
typedef __attribute__((address_space(3))) int mg_lds_int;
__device__ __forceinline__ int lds_atomic_add(int *p, int v) {
    return __hip_atomic_fetch_add((mg_lds_int *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void qstage_begin(QStage &st) {
    if (threadIdx.x == 0) st.n = 0;
    __syncthreads();
}
// every thread of the workgroup calls this once, after its last append of the level
__device__ __forceinline__ void qstage_flush(QStage &st, const QOut &o) {
    __shared__ int s_base;
    __syncthreads();
    const int n = min(st.n, kQStage);
    if (threadIdx.x == 0) s_base = n ? atomicAdd(o.qn, n) : 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n && s_base + i < o.cap; i += blockDim.x) o.Qn[s_base + i] = st.v[i];
}

// lane's entry w (when add) at position k of the wave's reserved range
__device__ __forceinline__ void qout_put(const QOut &o, bool add, int k, int w) {
    if (!o.st) {
        if (add && k < o.cap) o.Qn[k] = w;
        return;
    }
    const bool spill = add && k >= kQStage;
    if (add && !spill) o.st->v[k] = w;
    const int g = wave_append(o.qn, spill);
    if (spill && g < o.cap) o.Qn[g] = w;
}

__device__ __forceinline__ void queue_add(bool add, int w, const int *__restrict__ row, const QOut &o, int &cnt,
                                          long long &mf, int &mx) {
    const uint64_t m = ballot(add);
    if (!m) return;
    int base = 0;
    if (lane() == 0) base = o.st ? lds_atomic_add(&o.st->n, (int)__builtin_popcountll(m)) : atomicAdd(o.qn, (int)__builtin_popcountll(m));
    base = readlane(base, 0);
    qout_put(o, add, base + mbcnt(m), w);
    if (add) {
        const int dg = row[w + 1] - row[w];
        ++cnt;
        mf += dg;
        mx = max(mx, dg);
    }
}

// U appends per lane at once (lane l's j-th entry when add[j], of degree dg[j], loaded
// by the caller alongside its atomics): one counter atomic for the wave
template <int U>
__device__ __forceinline__ void queue_add_n(const bool *add, const int *w, const int *dgs, const QOut &o, int &cnt,
                                            long long &mf, int &mx) {
    uint64_t bm[U];
    int tot = 0;
#pragma unroll
    for (int j = 0; j < U; ++j) { bm[j] = ballot(add[j]); tot += __builtin_popcountll(bm[j]); }
    if (tot == 0) return;
    int base = 0;
    if (lane() == 0) base = o.st ? lds_atomic_add(&o.st->n, tot) : atomicAdd(o.qn, tot);
    base = readlane(base, 0);
#pragma unroll
    for (int j = 0; j < U; ++j) {
        qout_put(o, add[j], base + mbcnt(bm[j]), w[j]);
        if (add[j]) {
            const int dg = dgs[j];
            ++cnt;
            mf += dg;
            mx = max(mx, dg);
        }
        base += __builtin_popcountll(bm[j]);
    }
}

constexpr int kSparseLevels = 32;  // levels per replay of the sparse-only graph (at least)
#ifndef MG_VISIT
#define MG_VISIT 4
#endif
constexpr int kVisit = MG_VISIT;  // edges a lane of a sparse level visits together (uneven batch, 2 / 4 / 8: BFS 53.7 / 51.9 / 51.6 ms, SSSP at band 14 133.8 / 129.5 / 135.2 ms)

template <bool kDup = false>
__device__ __forceinline__ void sparse_bfs(int it, int n, const int *__restrict__ row, const int *__restrict__ col, uint64_t *F,
                                           uint64_t *Fn, uint64_t *vis, const int *Q, const QOut &o,
                                           int *__restrict__ lev, int &cnt, long long &mf, int &mx) {
    const int wave = (blockIdx.x * kBB + threadIdx.x) >> 6;
    const int nw = (gridDim.x * kBB) >> 6;
    for (int base = wave * MG_WAVE; base < n; base += nw * MG_WAVE) {
        const int i = base + lane();
        int u = -1, p0 = 0, p1 = 0;
        uint64_t fu = 0;
        if (i < n) { u = Q[i]; fu = F[u]; p0 = row[u]; p1 = row[u + 1]; }
        const bool alone = u >= 0 && p1 - p0 <= kLaneWork;
        int steps = alone ? p1 - p0 : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) steps = max(steps, __shfl_xor(steps, o));
        // kVisit edges per step, each stage's loads / atomics issued for all of them
        // before any result is used (one dependent chain per step, not per edge): the
        // claim on the visited word goes out with the neighbours' degree loads (no
        // load of the word first: an atomicOr of bits already set claims nothing)
        for (int k0 = 0; k0 < steps; k0 += kVisit) {
            int w[kVisit], dg[kVisit];
            uint64_t m[kVisit];
            bool add[kVisit];
#pragma unroll
            for (int j = 0; j < kVisit; ++j) w[j] = alone && p0 + k0 + j < p1 ? col[p0 + k0 + j] : -1;
#pragma unroll
            for (int j = 0; j < kVisit; ++j) {
                m[j] = w[j] >= 0 ? fu & ~(uint64_t)atomicOr((unsigned long long *)&vis[w[j]], (unsigned long long)fu) : 0ull;
                dg[j] = w[j] >= 0 ? row[w[j] + 1] - row[w[j]] : 0;
            }
#pragma unroll
            for (int j = 0; j < kVisit; ++j) {
                if (kDup) {
                    if (m[j]) atomicOr((unsigned long long *)&Fn[w[j]], (unsigned long long)m[j]);
                    add[j] = m[j] != 0ull;
                } else {
                    add[j] = m[j] && atomicOr((unsigned long long *)&Fn[w[j]], (unsigned long long)m[j]) == 0ull;
                }
            }
#pragma unroll
            for (int j = 0; j < kVisit; ++j)
                for (uint64_t bits = m[j]; bits; bits &= bits - 1) lev[(long long)w[j] * 64 + __builtin_ctzll(bits)] = it;
            queue_add_n<kVisit>(add, w, dg, o, cnt, mf, mx);
        }
        uint64_t heavy = ballot(u >= 0 && !alone);
        while (heavy) {  // the whole wave over one heavier vertex's edges
            const int l = __builtin_ctzll(heavy);
            heavy &= heavy - 1;
            const int q0 = __shfl(p0, l), q1 = __shfl(p1, l);
            const uint64_t fv = (uint64_t)__shfl((long long)fu, l);
            for (int pb = q0; pb < q1; pb += MG_WAVE) {
                const int pp = pb + lane();
                const int w = pp < q1 ? col[pp] : 0;
                const bool add = pp < q1 && bfs_visit<kDup>(it, w, fv, Fn, vis, lev);
                queue_add(add, w, row, o, cnt, mf, mx);
            }
        }
        if (u >= 0) F[u] = 0ull;
    }
}

// SSSP: a queued vertex u has changed bits (the sources whose distance at u dropped);
// along each edge (u, w) the distance of such a source at w is lowered to
// dist[u][s] + weight (read first, atomicMin only when it improves) and the source's
// bit is or-ed into w's next changed word; w joins the next queue when that word was
// empty.  A lane walks its vertex's (bit, edge) pairs alone when they are few; a
// heavier vertex is walked by the wave with lanes = sources, edges in turn.
// Pairs whose distance is not below `thr` are not relaxed in this sweep: their bits go
// back into the vertex's next changed word (near-far order: the threshold rises by a
// fixed band per level, so distances are settled roughly in increasing order instead of
// being lowered again and again as in plain Bellman-Ford; the result is the same).
// (Unlike BFS levels, SSSP levels queue a vertex only when its changed word was empty:
// an improvement is tested against a plain load of the distance, so copies of a vertex
// racing on one neighbour would both improve it and queue it twice, and the copies
// multiply along the wave: the uneven batch at band 14 took 445 ms instead of 132.)
template <class H>
__device__ __forceinline__ void sparse_sssp(int it, int n, int thr, const int *__restrict__ row, const int *__restrict__ col,
                                            const int *__restrict__ wgt, uint64_t *C, uint64_t *Cn, const int *Q, const QOut &o,
                                            int *__restrict__ dist, int &cnt, long long &mf, int &mx) {
    const int wave = (blockIdx.x * kBB + threadIdx.x) >> 6;
    const int nw = (gridDim.x * kBB) >> 6;
    for (int base = wave * MG_WAVE; base < n; base += nw * MG_WAVE) {
        const int i = base + lane();
        int u = -1, p0 = 0, p1 = 0;
        uint64_t cu = 0;
        if (i < n) { u = Q[i]; cu = C[u]; p0 = row[u]; p1 = row[u + 1]; }
        const bool alone = u >= 0 && __builtin_popcountll(cu) * (p1 - p0) <= kLaneWork;
        // bits still to walk, the lowest being the current source; a vertex without
        // out-edges (a sink of a directed graph, an isolated id) has nothing to relax,
        // and walking it would step into the next vertex's row
        uint64_t rem = alone && p1 > p0 ? cu : 0ull;
        uint64_t far = 0ull;               // bits left for a later sweep
        if (thr != 0x7FFFFFFF) {
            for (uint64_t t = rem; t; t &= t - 1) {
                const int sl = __builtin_ctzll(t);
                if (dist[(long long)u * 64 + sl] >= thr) far |= 1ull << sl;
            }
            rem &= ~far;
        }
        // one source (bit) at a time, its edges kVisit at a time: the column, weight and
        // distance loads, the relaxations and the changed-word atomics of a step are
        // issued together (the neighbours' degrees load with the distances)
        const int nchunk = (p1 - p0 + kVisit - 1) / kVisit;
        int steps = __builtin_popcountll(rem) * nchunk;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) steps = max(steps, __shfl_xor(steps, o));
        int p = p0, du = 0;
        if (rem) du = dist[(long long)u * 64 + __builtin_ctzll(rem)];
        for (int k = 0; k < steps; ++k) {
            const int sl = rem ? __builtin_ctzll(rem) : 0;
            int w[kVisit], wt[kVisit], cur[kVisit], nd[kVisit], dg[kVisit];
            bool imp[kVisit], add[kVisit];
#pragma unroll
            for (int j = 0; j < kVisit; ++j) {
                const bool ok = rem && p + j < p1;
                w[j] = ok ? col[p + j] : -1;
                wt[j] = ok ? wgt[p + j] : 0;
            }
#pragma unroll
            for (int j = 0; j < kVisit; ++j) {
                cur[j] = w[j] >= 0 ? dist[(long long)w[j] * 64 + sl] : 0;
                dg[j] = w[j] >= 0 ? row[w[j] + 1] - row[w[j]] : 0;
            }
#pragma unroll
            for (int j = 0; j < kVisit; ++j) {
                H::get_partial_result(&nd[j], du, wt[j]);
                imp[j] = w[j] >= 0 && nd[j] < cur[j];
                if (imp[j]) H::accumulate(&dist[(long long)w[j] * 64 + sl], nd[j], it);
            }
#pragma unroll
            for (int j = 0; j < kVisit; ++j)
                add[j] = imp[j] && atomicOr((unsigned long long *)&Cn[w[j]], 1ull << sl) == 0ull;
            queue_add_n<kVisit>(add, w, dg, o, cnt, mf, mx);
            if (rem) {
                p += kVisit;
                if (p >= p1) {
                    p = p0;
                    rem &= rem - 1;
                    if (rem) du = dist[(long long)u * 64 + __builtin_ctzll(rem)];
                }
            }
        }
        {  // far pairs wait in the next changed word
            const bool add = far && atomicOr((unsigned long long *)&Cn[u], (unsigned long long)far) == 0ull;
            queue_add(add, u, row, o, cnt, mf, mx);
        }
        uint64_t heavy = ballot(u >= 0 && !alone);
        while (heavy) {
            const int l = __builtin_ctzll(heavy);
            heavy &= heavy - 1;
            const int v = __shfl(u, l), q0 = __shfl(p0, l), q1 = __shfl(p1, l);
            const uint64_t cv = (uint64_t)__shfl((long long)cu, l);
            const bool bit = (cv >> lane()) & 1ull;
            const int dv0 = bit ? dist[(long long)v * 64 + lane()] : (int)MG_INF;
            const bool mine = bit && dv0 < thr;
            const int dv = mine ? dv0 : (int)MG_INF;
            {
                const uint64_t fv = ballot(bit && !mine);
                const bool add = lane() == 0 && fv && atomicOr((unsigned long long *)&Cn[v], (unsigned long long)fv) == 0ull;
                queue_add(add, v, row, o, cnt, mf, mx);
            }
            for (int pp = q0; pp < q1; ++pp) {
                const int w = col[pp];
                int nd;
                H::get_partial_result(&nd, dv, wgt[pp]);
                bool imp = false;
                if (mine) {
                    int *dw = &dist[(long long)w * 64 + lane()];
                    if (nd < *dw) { H::accumulate(dw, nd, it); imp = true; }
                }
                const uint64_t m = ballot(imp);
                bool add = false;
                if (lane() == 0 && m) add = atomicOr((unsigned long long *)&Cn[w], (unsigned long long)m) == 0ull;
                queue_add(add, w, row, o, cnt, mf, mx);
            }
        }
        if (u >= 0) C[u] = 0ull;
    }
}

// partials of workgroups [gridDim.x, nparts) are zero: a sparse level runs on a smaller
// grid than the dense kernels whose partial count the control kernel sums
__device__ __forceinline__ void zero_rest(int nparts, int *__restrict__ part_n, long long *__restrict__ part_e,
                                          int *__restrict__ part_x) {
    for (int i = gridDim.x + blockIdx.x * kBB + threadIdx.x; i < nparts; i += gridDim.x * kBB) {
        part_n[i] = 0;
        part_e[i] = 0;
        part_x[i] = 0;
    }
}

// one sparse level from a captured graph (gated on the control block), on a grid of
// about one workgroup per CU (queues are short: a 2048-workgroup launch would mostly
// dispatch idle workgroups)
__global__ __launch_bounds__(kBB) void k_mbs_level_bfs(BCtl *ctl, const int *__restrict__ row, const int *__restrict__ col,
                                                       uint64_t *FA, uint64_t *FB, uint64_t *vis, int *QA, int *QB,
                                                       int *__restrict__ lev, int *__restrict__ part_n,
                                                       long long *__restrict__ part_e, int *__restrict__ part_x, int nparts) {
    if (ctl->done || !ctl->sparse || ctl->need_full) return;
    const int it = ctl->it;
    const int cur = (it - 1) & 1;
    int cnt = 0, mx = 0;
    long long mf = 0;
    const QOut o{cur ? QA : QB, &ctl->qn[cur ^ 1], nullptr};
    sparse_bfs(it, ctl->qn[cur], row, col, cur ? FB : FA, cur ? FA : FB, vis, cur ? QB : QA, o, lev, cnt, mf, mx);
    put_parts(cnt, mf, mx, part_n, part_e, part_x);
    zero_rest(nparts, part_n, part_e, part_x);
}

__global__ __launch_bounds__(kBB) void k_mbs_level_sssp(BCtl *ctl, const int *__restrict__ row, const int *__restrict__ col,
                                                        const int *__restrict__ wgt, uint64_t *FA, uint64_t *FB, int *QA, int *QB,
                                                        int *__restrict__ dist, int *__restrict__ part_n,
                                                        long long *__restrict__ part_e, int *__restrict__ part_x, int nparts) {
    if (ctl->done || !ctl->sparse || ctl->need_full) return;
    const int cur = (ctl->it - 1) & 1;
    int cnt = 0, mx = 0;
    long long mf = 0;
    __shared__ QStage st;
    qstage_begin(st);
    const QOut o{cur ? QA : QB, &ctl->qn[cur ^ 1], &st};
    sparse_sssp<MG_SSSP_HOOKS>(ctl->it, ctl->qn[cur], ctl->thr, row, col, wgt, cur ? FB : FA, cur ? FA : FB, cur ? QB : QA, o, dist,
                               cnt, mf, mx);
    qstage_flush(st, o);
    put_parts(cnt, mf, mx, part_n, part_e, part_x);
    zero_rest(nparts, part_n, part_e, part_x);
}

// Sparse levels one graph node each (the sparse-only graph).  A level kernel takes the
// totals of the frontier the previous level produced (count = its queue length, degree
// sum, largest degree) from a triple-buffered FStat, which the previous kernel's
// workgroups filled with atomics, and applies the level rule itself, every workgroup
// alike; the control block is double-buffered (node k reads copy k & 1, workgroup 0
// writes the decided block to copy (k + 1) & 1 and clears the stats slot node k + 1
// will fill).  So no control kernel runs between levels: the kernel boundary alone
// orders one level's atomics before the next level's reads.  A level that cannot run
// sparse (or an empty frontier) passes block and stats on unchanged; the exit kernel
// then hands both back to the full graph, whose control kernel decides again.
struct FStat { int cnt; int mx; long long mf; };

__global__ void k_mbs_fused_enter(const BCtl *ctl, BCtl *c2, FStat *st, const int *__restrict__ pn,
                                  const long long *__restrict__ pe, const int *__restrict__ px, int nparts) {
    long long a, b, x;
    sum_partials(pn, pe, px, nparts, &a, &b, &x);
    if (threadIdx.x != 0) return;
    c2[0] = *ctl;
    st[0] = FStat{(int)a, (int)x, b};
    st[1] = FStat{0, 0, 0};
    st[2] = FStat{0, 0, 0};
}

template <bool kSSSP>
__global__ __launch_bounds__(kBB) void k_mbs_fused(int k, BCtl *c2, FStat *st, LevelRule r, const int *__restrict__ row,
                                                   const int *__restrict__ col, const int *__restrict__ wgt, uint64_t *FA,
                                                   uint64_t *FB, uint64_t *vis, int *QA, int *QB, int *__restrict__ out) {
    const BCtl c0 = c2[k & 1];
    const FStat s0 = st[k % 3];
    BCtl c = c0;
    bool run = !c.done && !c.need_full;
    if (run) {
        decide(c, s0.cnt, s0.mf, s0.mx, r);
        if (c.done) run = false;
        else if (!c.sparse || c.compact || c.fill) { c = c0; c.need_full = 1; run = false; }
        else if (s0.cnt > r.nv + 1) {  // appends were dropped: the full graph rebuilds the queue from the words
            c = c0;
            c.qvalid[c0.it & 1] = 0;
            c.need_full = 1;
            run = false;
        } else ++c.nsparse;
    }
    if (!run) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { c2[(k + 1) & 1] = c; st[(k + 1) % 3] = s0; }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) { c2[(k + 1) & 1] = c; st[(k + 2) % 3] = FStat{0, 0, 0}; }
    FStat *nx = &st[(k + 1) % 3];
    const int it = c.it, cur = (it - 1) & 1;
    int cnt = 0, mx = 0;
    long long mf = 0;
    if (kSSSP) {  // staged appends (BFS levels append once per wave and step: staging them measured slower)
        __shared__ QStage qs;
        qstage_begin(qs);
        const QOut o{cur ? QA : QB, &nx->cnt, &qs, r.nv + 1};
        sparse_sssp<MG_SSSP_HOOKS>(it, s0.cnt, c.thr, row, col, wgt, cur ? FB : FA, cur ? FA : FB, cur ? QB : QA, o, out, cnt, mf, mx);
        qstage_flush(qs, o);
    } else {
        const QOut o{cur ? QA : QB, &nx->cnt, nullptr, r.nv + 1};
        sparse_bfs<true>(it, s0.cnt, row, col, cur ? FB : FA, cur ? FA : FB, vis, cur ? QB : QA, o, out, cnt, mf, mx);
    }
    __shared__ long long s_tmp[kBB / MG_WAVE];
    const long long tm = block_sum<kBB / MG_WAVE>(mf, s_tmp);
    const long long tx = block_max<kBB / MG_WAVE>((long long)mx, s_tmp);
    if (threadIdx.x == 0) {
        if (tm) atomicAdd((unsigned long long *)&nx->mf, (unsigned long long)tm);
        if (tx) atomicMax(&nx->mx, (int)tx);
    }
}

// after K fused nodes: the control block (with the produced frontier's queue length)
// and the frontier's totals as partials, as a sparse level of the full graph leaves them
__global__ void k_mbs_fused_exit(int K, BCtl *ctl, const BCtl *c2, const FStat *st, int *__restrict__ pn,
                                 long long *__restrict__ pe, int *__restrict__ px, int nparts, int qcap) {
    FStat t = st[K % 3];
    const bool dropped = t.cnt > qcap;  // the last level's appends outgrew the queue
    if (dropped) t.cnt = qcap;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        pn[i] = i ? 0 : t.cnt;
        pe[i] = i ? 0 : t.mf;
        px[i] = i ? 0 : t.mx;
    }
    if (threadIdx.x == 0) {
        BCtl c = c2[K & 1];
        c.qn[c.it & 1] = t.cnt;
        if (dropped) c.qvalid[c.it & 1] = 0;  // the next sparse level rebuilds the queue from the words
        *ctl = c;
    }
}

// Persistent sparse SSSP levels (k_mbs_persist; BFS has its own, k_mbp_bfs, below):
// one cooperative launch (every workgroup resident) runs consecutive sparse levels
// with a grid barrier between them, so a long tail of
// small frontiers (the thousands of levels of a large grid in an uneven batch) costs
// a barrier per level instead of a control kernel, a level kernel and a graph node
// each.  It starts after a sparse level of the captured graph, decides every level
// with the control block's rule, and stops at the first level that is not sparse
// (need_full: the captured graph's control kernel re-decides from the partials it
// leaves), when the frontier is empty, or after max_levels levels.  Level L appends
// to the counters of slot L % 3; slot (L + 1) % 3 is cleared during level L, which no
// workgroup reads after the barrier that opened level L.
struct PCtl {
    int qn[3]; int mx[3]; long long mf[3];
    unsigned top;            // barrier: groups of workgroups arrived (zeroed before the launch)
    unsigned pad[15];
    unsigned grp[8 * 16];    // barrier: arrivals per group, one 64 B line each
    // entry-queue levels (k_mbp_bfs): an entry queue overflowed; the handback queue's
    // length, degree sum and largest degree
    int over, fq, fmx, err;
    long long fmf;
    unsigned xcc0, reg, xbar, pad3;  // its XCD (XCC id + 1), workgroups counted in, barrier
    int npart, pad4[3];
};

#ifndef MG_SPARSE_PER_CU
#define MG_SPARSE_PER_CU 1  // workgroups per CU of a sparse level
#endif
#ifndef MG_PERSIST_PER_CU
#define MG_PERSIST_PER_CU 1  // resident workgroups per CU of the persistent launch
#endif

// Grid barrier of a cooperative launch (every workgroup resident), in two levels so
// that no single address takes every arrival: workgroups count in on the counter of
// their group (blockIdx mod 8, the XCD a round-robin dispatch puts them on); the last
// of a group counts the group in on `top`, which every group leader polls.  Counters
// only grow, so barrier g waits for g arrivals per group member and g * groups on
// `top`.  Each thread makes its writes visible at device scope first, and the leader's
// fence after the wait drops stale cache lines before the workgroup goes on.
__device__ __forceinline__ void grid_barrier(PCtl *pc, unsigned &gen) {
    ++gen;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned groups = min(8u, gridDim.x);
        const unsigned g = blockIdx.x % groups;
        const unsigned members = (gridDim.x - g + groups - 1) / groups;
        const unsigned old = __hip_atomic_fetch_add(&pc->grp[g * 16], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == gen * members) __hip_atomic_fetch_add(&pc->top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (__hip_atomic_load(&pc->top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen * groups) __builtin_amdgcn_s_sleep(2);
        __threadfence();
    }
    __syncthreads();
}

struct PersistArgs {
    BCtl *ctl;
    PCtl *pc;
    LevelRule rule;
    const int *row, *col, *wgt;
    uint64_t *FA, *FB, *vis;
    int *QA, *QB;
    int *out;  // distances
    int *part_n;
    long long *part_e;
    int *part_x;
    int nparts, max_levels;
};

__global__ __launch_bounds__(kBB) void k_mbs_persist(PersistArgs a) {
    __shared__ long long s_tmp[kBB / MG_WAVE];
    unsigned gen = 0;
    long long n, mf, mx;
    {  // the frontier of the level before: per-workgroup partials of the captured graph
        long long x = 0, y = 0, z = 0;
        for (int i = threadIdx.x; i < a.nparts; i += kBB) { x += a.part_n[i]; y += a.part_e[i]; z = max(z, (long long)a.part_x[i]); }
        n = block_sum<kBB / MG_WAVE>(x, s_tmp);
        mf = block_sum<kBB / MG_WAVE>(y, s_tmp);
        mx = block_max<kBB / MG_WAVE>(z, s_tmp);
    }
    const BCtl c0 = *a.ctl;
    if (blockIdx.x == 0 && threadIdx.x < 3) { a.pc->qn[threadIdx.x] = 0; a.pc->mx[threadIdx.x] = 0; a.pc->mf[threadIdx.x] = 0; }
    grid_barrier(a.pc, gen);
    int L = c0.it;   // the last level run
    int ran = 0;     // levels run by this launch
    BCtl c = c0;
    bool handback = false;
    int thr = c0.thr;  // the threshold of the last level run
    for (;;) {
        c = c0;
        c.it = L;
        c.thr = thr;
        c.qvalid[L & 1] = 1;
        c.prev_sparse = 1;
        decide(c, n, mf, mx, a.rule);
        if (c.done) break;
        if (!c.sparse || c.compact || c.fill) { handback = true; break; }
        if (ran >= a.max_levels) break;
        const int it = L + 1, cur = L & 1, slot = it % 3;
        int cnt = 0, xm = 0;
        long long sm = 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {  // the counters of the next level
            const int ns = (it + 1) % 3;
            a.pc->qn[ns] = 0;
            a.pc->mx[ns] = 0;
            a.pc->mf[ns] = 0;
        }
        const QOut o{cur ? a.QA : a.QB, &a.pc->qn[slot], nullptr};
        sparse_sssp<MG_SSSP_HOOKS>(it, (int)n, c.thr, a.row, a.col, a.wgt, cur ? a.FB : a.FA, cur ? a.FA : a.FB, cur ? a.QB : a.QA,
                                   o, a.out, cnt, sm, xm);
        (void)cnt;
        const long long bs = block_sum<kBB / MG_WAVE>(sm, s_tmp);
        const long long bx = block_max<kBB / MG_WAVE>((long long)xm, s_tmp);
        if (threadIdx.x == 0) {
            if (bs) atomicAdd((unsigned long long *)&a.pc->mf[slot], (unsigned long long)bs);
            if (bx) atomicMax(&a.pc->mx[slot], (int)bx);
        }
        grid_barrier(a.pc, gen);
        n = __hip_atomic_load(&a.pc->qn[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mf = (long long)__hip_atomic_load((unsigned long long *)&a.pc->mf[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mx = __hip_atomic_load(&a.pc->mx[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        L = it;
        thr = c.thr;
        ++ran;
    }
    if (ran == 0) {  // nothing run: the state is the captured graph's
        if (handback && blockIdx.x == 0 && threadIdx.x == 0) a.ctl->need_full = 1;
        if (c.done && blockIdx.x == 0 && threadIdx.x == 0) { a.ctl->done = 1; a.ctl->n = 0; }
        return;
    }
    // leave the state a sparse level of the captured graph leaves: the last level's
    // frontier as partials (all in part 0) and its queue length in the control block
    for (int i = blockIdx.x * kBB + threadIdx.x; i < a.nparts; i += gridDim.x * kBB) {
        a.part_n[i] = i == 0 ? (int)n : 0;
        a.part_e[i] = i == 0 ? mf : 0;
        a.part_x[i] = i == 0 ? (int)mx : 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BCtl o = c0;
        o.it = L;
        o.n = n;
        o.mf = mf;
        o.push = 1;
        o.sparse = 1;
        o.compact = 0;
        o.fill = 0;
        o.prev_sparse = 1;
        o.qn[L & 1] = (int)n;
        o.qvalid[L & 1] = 1;
        o.done = c.done;
        o.need_full = handback ? 1 : 0;
        o.nsparse = c0.nsparse + ran;
        o.npersist = c0.npersist + ran;
        o.thr = thr;
        *a.ctl = o;
    }
}

// ---- persistent sparse BFS levels on one XCD, over entry queues ---------------------
// The long tails of batched BFS (thousands of levels of a few thousand vertices: the
// large grids of an uneven batch) are bound by the latency of a level, not by its
// bytes.  Here one cooperative launch runs consecutive sparse levels on the workgroups
// of ONE XCD (those whose XCC id is block 0's): they share that XCD's L2, so a level
// hands its frontier to the next through plain stores (written through each CU's L1
// into the L2) and L1-bypassing loads (sc0) that read the L2, and the barrier between
// levels is an L2 atomic counter, with no cache write-back or invalidation.  The other
// workgroups of the launch leave at once.
//
// A level's frontier is a list of entries (vertex, first edge, one past its last edge,
// the sources it carries), so a level loads no frontier word or row before it walks
// edges: a claim on the visited word (atomicOr) returns the new bits, and the
// discoverer, which loads the new vertex's row bounds alongside the claim, appends the
// entry.  Two discoverers of one vertex claim disjoint bits and append two entries.
// Levels run here while the frontier is small (vertices + edges at most kXcdWork per
// workgroup of the XCD); a larger or non-sparse level, an overflowing entry queue or
// max_levels hand the frontier back to the captured graphs as words, queue and
// partials.
typedef unsigned int mg_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t l2_rsrc(const void *p, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)min(bytes, 0x7FFFFFF0ll), 0x00020000);
}
// 64-bit word i of a buffer, read from the L2 (sc0: not from the CU's L1)
__device__ __forceinline__ uint64_t ld_l2(__amdgpu_buffer_rsrc_t r, int i) {
    const mg_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, 0, 1);
    return (uint64_t)v.x | ((uint64_t)v.y << 32);
}
// the current value of a counter that is only changed by atomics: read by an atomic
__device__ __forceinline__ unsigned long long ld_atomic(unsigned long long *p) {
    return __hip_atomic_fetch_add(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_atomic(int *p) { return __hip_atomic_fetch_add(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ unsigned ld_atomic(unsigned *p) { return __hip_atomic_fetch_add(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

constexpr int kXcdWork = 16 * kBB;            // frontier vertices + edges per workgroup of a persistent level
constexpr long long kSpinTicks = 100000000;   // bound of one barrier wait: 1 s of the 100 MHz wall clock

// Barrier of the XCD's P workgroups.  Each thread waits for its own memory operations
// to complete (stores are then in the L2), the workgroup counts in on an L2 counter and
// polls it by atomics.  The wait is bounded: past kSpinTicks the launch is marked
// failed (err) and every later wait returns at once, so a barrier cannot hang.
__device__ __forceinline__ bool xcd_sync(PCtl *pc, unsigned &gen, int P) {
    ++gen;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    __shared__ int s_err;
    if (threadIdx.x == 0) {
        atomicAdd(&pc->xbar, 1u);
        const long long t0 = wall_clock64();
        int err = ld_atomic(&pc->err);
        while (!err && ld_atomic(&pc->xbar) < gen * (unsigned)P) {
            __builtin_amdgcn_s_sleep(1);
            if (wall_clock64() - t0 > kSpinTicks) { atomicExch(&pc->err, 1); err = 1; }
            else err = ld_atomic(&pc->err);
        }
        s_err = err;
    }
    __syncthreads();
    return s_err == 0;
}

// entry queue: ea = vertex | first edge << 32, eb = one past its last edge, ec = its bits
struct EQ { uint64_t *ea, *eb, *ec; };

// handback of a vertex with bits: or-ed into its frontier word F, queued in Q when the
// word was empty (the form a sparse level of the captured graphs leaves)
__device__ __forceinline__ void eq_handback(bool on, int w, int deg, uint64_t bits, uint64_t *F, int *Q, PCtl *pc) {
    const bool add = on && atomicOr((unsigned long long *)&F[w], (unsigned long long)bits) == 0ull;
    const int k = wave_append(&pc->fq, add);
    if (add) {
        Q[k] = w;
        atomicAdd((unsigned long long *)&pc->fmf, (unsigned long long)deg);
        atomicMax(&pc->fmx, deg);
    }
}

// U candidate entries per lane (on[j]: vertex w[j] with edges [q0, q1) and bits m[j]):
// one counter atomic per wave; entries past the capacity are handed back at once (and
// the launch ends after this level)
template <int U>
__device__ __forceinline__ void eq_add_n(const bool *on, const int *w, const int *q0, const int *q1, const uint64_t *m, EQ out,
                                         int *en, int ecap, uint64_t *F, int *Q, PCtl *pc, long long &mf, int &mx) {
    uint64_t bm[U];
    int tot = 0;
#pragma unroll
    for (int j = 0; j < U; ++j) { bm[j] = ballot(on[j]); tot += __builtin_popcountll(bm[j]); }
    if (tot == 0) return;
    int base = 0;
    if (lane() == 0) base = atomicAdd(en, tot);
    base = readlane(base, 0);
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const int k = base + mbcnt(bm[j]);
        const bool fits = on[j] && k < ecap;
        if (fits) {
            out.ea[k] = (uint64_t)(uint32_t)w[j] | ((uint64_t)(uint32_t)q0[j] << 32);
            out.eb[k] = (uint64_t)(uint32_t)q1[j];
            out.ec[k] = m[j];
        }
        if (on[j]) {
            mf += q1[j] - q0[j];
            mx = max(mx, q1[j] - q0[j]);
        }
        const bool spill = on[j] && !fits;
        if (ballot(spill)) {
            if (spill) atomicExch(&pc->over, 1);
            eq_handback(spill, w[j], q1[j] - q0[j], m[j], F, Q, pc);
        }
        base += __builtin_popcountll(bm[j]);
    }
}

// one level: every entry's edges, claims on the visited words, level rows, next entries
// (wave wv of nw walks entries wv * 64 ..)
__device__ __forceinline__ void eq_bfs_level(int it, int n, int wv, int nw, const int *__restrict__ row, const int *__restrict__ col,
                                             uint64_t *vis, __amdgpu_buffer_rsrc_t ra, __amdgpu_buffer_rsrc_t rb,
                                             __amdgpu_buffer_rsrc_t rc, EQ out, int *en, int ecap, uint64_t *F, int *Q, PCtl *pc,
                                             int *__restrict__ lev, long long &mf, int &mx) {
    for (int base = wv * MG_WAVE; base < n; base += nw * MG_WAVE) {
        const int i = base + lane();
        int p0 = 0, p1 = 0;
        uint64_t fu = 0;
        if (i < n) {
            p0 = (int)(ld_l2(ra, i) >> 32);
            p1 = (int)ld_l2(rb, i);
            fu = ld_l2(rc, i);
        }
        const bool alone = fu && p1 - p0 <= kLaneWork;
        int steps = alone ? p1 - p0 : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) steps = max(steps, __shfl_xor(steps, o));
        // kVisit edges per step: the claims go out with the new vertices' row loads
        for (int k0 = 0; k0 < steps; k0 += kVisit) {
            int w[kVisit], q0[kVisit], q1[kVisit];
            uint64_t m[kVisit];
            bool on[kVisit];
#pragma unroll
            for (int j = 0; j < kVisit; ++j) w[j] = alone && p0 + k0 + j < p1 ? col[p0 + k0 + j] : -1;
#pragma unroll
            for (int j = 0; j < kVisit; ++j) {
                m[j] = w[j] >= 0 ? fu & ~(uint64_t)atomicOr((unsigned long long *)&vis[w[j]], (unsigned long long)fu) : 0ull;
                q0[j] = w[j] >= 0 ? row[w[j]] : 0;
                q1[j] = w[j] >= 0 ? row[w[j] + 1] : 0;
            }
#pragma unroll
            for (int j = 0; j < kVisit; ++j) {
                on[j] = m[j] != 0ull;
                for (uint64_t bits = m[j]; bits; bits &= bits - 1) lev[(long long)w[j] * 64 + __builtin_ctzll(bits)] = it;
            }
            eq_add_n<kVisit>(on, w, q0, q1, m, out, en, ecap, F, Q, pc, mf, mx);
        }
        uint64_t heavy = ballot(fu && !alone);
        while (heavy) {  // the whole wave over one heavier vertex's edges
            const int l = __builtin_ctzll(heavy);
            heavy &= heavy - 1;
            const int r0 = __shfl(p0, l), r1 = __shfl(p1, l);
            const uint64_t fv = (uint64_t)__shfl((long long)fu, l);
            for (int pb = r0; pb < r1; pb += MG_WAVE) {
                const int pp = pb + lane();
                int w = -1, q0 = 0, q1 = 0;
                uint64_t m = 0ull;
                if (pp < r1) {
                    w = col[pp];
                    m = fv & ~(uint64_t)atomicOr((unsigned long long *)&vis[w], (unsigned long long)fv);
                    q0 = row[w];
                    q1 = row[w + 1];
                }
                bool on = m != 0ull;
                for (uint64_t bits = m; bits; bits &= bits - 1) lev[(long long)w * 64 + __builtin_ctzll(bits)] = it;
                eq_add_n<1>(&on, &w, &q0, &q1, &m, out, en, ecap, F, Q, pc, mf, mx);
            }
        }
    }
}

struct EqArgs {
    BCtl *ctl;
    PCtl *pc;
    LevelRule rule;
    const int *row, *col;
    uint64_t *FA, *FB, *vis;
    int *QA, *QB;
    int *lev;
    int *part_n;
    long long *part_e;
    int *part_x;
    EQ q[2];
    int ecap, nparts, max_levels;
};

// Persistent sparse BFS levels (the contract of k_mbs_persist): entered after a sparse
// level of the captured graphs (frontier words F[L & 1] and queue Q[L & 1]), left with
// the last level's frontier in the same form and the control block updated.
__global__ __launch_bounds__(kBB) void k_mbp_bfs(EqArgs a) {
    __shared__ long long s_tmp[kBB / MG_WAVE];
    __shared__ long long s_v[4];
    long long n, mf, mx;
    {
        long long x = 0, y = 0, z = 0;
        for (int i = threadIdx.x; i < a.nparts; i += kBB) { x += a.part_n[i]; y += a.part_e[i]; z = max(z, (long long)a.part_x[i]); }
        n = block_sum<kBB / MG_WAVE>(x, s_tmp);
        mf = block_sum<kBB / MG_WAVE>(y, s_tmp);
        mx = block_max<kBB / MG_WAVE>(z, s_tmp);
    }
    const BCtl c0 = *a.ctl;
    int L = c0.it;
    // the XCD's workgroups: block 0 publishes its XCC id, the workgroups on that XCC take
    // ranks, and every workgroup counts in on `reg` (all bounded waits, across XCDs)
    if (threadIdx.x == 0) {
        const unsigned x = (__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((16 - 1) << 11)) & 0xFu) + 1u;
        if (blockIdx.x == 0) atomicExch(&a.pc->xcc0, x);
        const long long t0 = wall_clock64();
        unsigned x0;
        while ((x0 = ld_atomic(&a.pc->xcc0)) == 0u && wall_clock64() - t0 < kSpinTicks) __builtin_amdgcn_s_sleep(1);
        const int rank = x0 == x ? atomicAdd(&a.pc->npart, 1) : -1;
        __builtin_amdgcn_s_waitcnt(0);
        atomicAdd(&a.pc->reg, 1u);
        int P = 0;
        if (rank >= 0) {
            while (ld_atomic(&a.pc->reg) < gridDim.x && wall_clock64() - t0 < kSpinTicks) __builtin_amdgcn_s_sleep(1);
            if (ld_atomic(&a.pc->reg) < gridDim.x) atomicExch(&a.pc->err, 1);
            P = ld_atomic(&a.pc->npart);
        }
        s_v[0] = rank;
        s_v[1] = P;
    }
    __syncthreads();
    const int rank = (int)s_v[0], P = (int)s_v[1];
    if (rank < 0) return;
    auto next_block = [&](int Lv) {
        BCtl c = c0;
        c.it = Lv;
        c.qvalid[Lv & 1] = 1;
        c.prev_sparse = 1;
        decide(c, n, mf, mx, a.rule);
        return c;
    };
    const long long small = (long long)kXcdWork * P;
    BCtl c = next_block(L);
    if (c.done || !c.sparse || c.compact || c.fill || n + mf > small || a.max_levels <= 0) {
        if (rank == 0 && threadIdx.x == 0) {  // nothing run: the state is the captured graph's
            if (c.done) { a.ctl->done = 1; a.ctl->n = 0; }
            else if (c.sparse && !c.compact && !c.fill) a.ctl->big = 1;  // a sparse level, too large for one XCD
            else a.ctl->need_full = 1;
        }
        return;
    }
    unsigned gen = 0;
    const int gt = rank * kBB + threadIdx.x, gn = P * kBB;
    const int wv = gt >> 6, nw = gn >> 6;
    {  // the queue of level L as entries (its words cleared: both word buffers end clean)
        uint64_t *F = (L & 1) ? a.FB : a.FA;
        const int *Q = (L & 1) ? a.QB : a.QA;
        const EQ e = a.q[L & 1];
        for (int i = gt; i < (int)n; i += gn) {
            const int w = Q[i];
            const uint64_t bits = atomicExch((unsigned long long *)&F[w], 0ull);
            e.ea[i] = (uint64_t)(uint32_t)w | ((uint64_t)(uint32_t)a.row[w] << 32);
            e.eb[i] = (uint64_t)(uint32_t)a.row[w + 1];
            e.ec[i] = bits;
        }
    }
    bool ok = xcd_sync(a.pc, gen, P);
    const long long qbytes = (long long)a.ecap * 8;
    int ran = 0;
    bool handback = false, over = false;
    while (ok) {
        const int it = L + 1, slot = it % 3;
        if (rank == 0 && threadIdx.x == 0) {  // the counters of the level after
            const int ns = (it + 1) % 3;
            atomicExch(&a.pc->qn[ns], 0);
            atomicExch(&a.pc->mx[ns], 0);
            atomicExch((unsigned long long *)&a.pc->mf[ns], 0ull);
        }
        long long sm = 0;
        int xm = 0;
        uint64_t *F = (it & 1) ? a.FB : a.FA;  // spills: the level's frontier words and queue
        int *Q = (it & 1) ? a.QB : a.QA;
        const EQ in = a.q[L & 1];
        eq_bfs_level(it, (int)min(n, (long long)a.ecap), wv, nw, a.row, a.col, a.vis, l2_rsrc(in.ea, qbytes), l2_rsrc(in.eb, qbytes),
                     l2_rsrc(in.ec, qbytes), a.q[it & 1], &a.pc->qn[slot], a.ecap, F, Q, a.pc, a.lev, sm, xm);
        const long long bs = block_sum<kBB / MG_WAVE>(sm, s_tmp);
        const long long bx = block_max<kBB / MG_WAVE>((long long)xm, s_tmp);
        if (threadIdx.x == 0) {
            if (bs) atomicAdd((unsigned long long *)&a.pc->mf[slot], (unsigned long long)bs);
            if (bx) atomicMax(&a.pc->mx[slot], (int)bx);
        }
        ok = xcd_sync(a.pc, gen, P);
        if (threadIdx.x == 0) {
            s_v[0] = ld_atomic(&a.pc->qn[slot]);
            s_v[1] = (long long)ld_atomic((unsigned long long *)&a.pc->mf[slot]);
            s_v[2] = ld_atomic(&a.pc->mx[slot]);
            s_v[3] = ld_atomic(&a.pc->over);
        }
        __syncthreads();
        n = s_v[0];
        mf = s_v[1];
        mx = s_v[2];
        over = s_v[3] != 0;
        __syncthreads();
        L = it;
        ++ran;
        if (over || !ok) break;
        c = next_block(L);
        if (c.done) break;
        if (!c.sparse || c.compact || c.fill || n + mf > small) { handback = true; break; }
        if (ran >= a.max_levels) break;
    }
    // hand the last level's entries back as frontier words and queue (spilled entries
    // are there already)
    {
        uint64_t *F = (L & 1) ? a.FB : a.FA;
        int *Q = (L & 1) ? a.QB : a.QA;
        const EQ e = a.q[L & 1];
        const __amdgpu_buffer_rsrc_t ra = l2_rsrc(e.ea, qbytes), rb = l2_rsrc(e.eb, qbytes), rc = l2_rsrc(e.ec, qbytes);
        const int ne = (int)min(n, (long long)a.ecap);
        for (int b0 = wv * MG_WAVE; b0 < ne; b0 += nw * MG_WAVE) {
            const int i = b0 + lane();
            int w = 0, deg = 0;
            uint64_t bits = 0;
            if (i < ne) {
                const uint64_t x = ld_l2(ra, i);
                w = (int)(uint32_t)x;
                deg = (int)ld_l2(rb, i) - (int)(x >> 32);
                bits = ld_l2(rc, i);
            }
            eq_handback(i < ne, w, deg, bits, F, Q, a.pc);
        }
    }
    xcd_sync(a.pc, gen, P);
    if (threadIdx.x == 0) {
        s_v[0] = ld_atomic(&a.pc->fq);
        s_v[1] = (long long)ld_atomic((unsigned long long *)&a.pc->fmf);
        s_v[2] = ld_atomic(&a.pc->fmx);
        s_v[3] = ld_atomic(&a.pc->err);
    }
    __syncthreads();
    const int qn = (int)s_v[0];
    const long long qf = s_v[1];
    const int qx = (int)s_v[2];
    for (int i = gt; i < a.nparts; i += gn) {
        a.part_n[i] = i == 0 ? qn : 0;
        a.part_e[i] = i == 0 ? qf : 0;
        a.part_x[i] = i == 0 ? qx : 0;
    }
    if (rank == 0 && threadIdx.x == 0) {
        BCtl o = c0;
        o.it = L;
        o.n = qn;
        o.mf = qf;
        o.push = 1;
        o.sparse = 1;
        o.compact = 0;
        o.fill = 0;
        o.prev_sparse = 1;
        o.qn[L & 1] = qn;
        o.qvalid[L & 1] = 1;
        o.done = qn == 0;
        o.need_full = (handback || over) && qn > 0 ? 1 : 0;
        o.nsparse = c0.nsparse + ran;
        o.npersist = c0.npersist + ran;
        o.err = s_v[3] != 0;
        o.big = 0;
        *a.ctl = o;
    }
}

__global__ void k_sum_parts(const int *__restrict__ pn, const long long *__restrict__ pe, int n, long long *__restrict__ out) {
    __shared__ long long s_tmp[256 / MG_WAVE];
    long long a = 0, b = 0;
    for (int i = threadIdx.x; i < n; i += 256) { a += pn[i]; if (pe) b += pe[i]; }
    a = block_sum<256 / MG_WAVE>(a, s_tmp);
    b = block_sum<256 / MG_WAVE>(b, s_tmp);
    if (threadIdx.x == 0) { out[0] = a; out[1] = b; }
}

// ---- BC backward -------------------------------------------------------------------
__global__ void k_ms_scatter_words(long long n, const int *__restrict__ lv, const uint64_t *__restrict__ lw,
                                   uint64_t *__restrict__ D, int clear) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        D[lv[i]] = clear ? 0ull : lw[i];
}

// term sigma[v][s] * c[w][s] with c = (1 + delta) / sigma of the successor, computed in
// FP64 and stored in FP32 (like delta) once per (vertex, source) when its level is
// finished (k_ms_cw): a matched successor costs one 256-byte row, not two rows and a
// division per source
template <class H>
__device__ __forceinline__ void ms_term(KSum &acc, double sv, float cw) {
    double t;
    H::back_term(&t, sv, (double)cw);
    H::back_accumulate(&acc, t);
}

// c[v][s] = (1 + delta[v][s]) / sigma[v][s] for the vertices split into several chunks
// (the list mv, fixed at build) with bit s in the finished level's words E: k_mb_back
// writes the c of every other vertex itself
template <class H>
__global__ void k_ms_cw(int n, const int *__restrict__ mv, const uint64_t *__restrict__ E, const double *__restrict__ sigma,
                        const float *__restrict__ delta, float *__restrict__ cw) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int i = wave; i < n; i += nw) {
        const int v = mv[i];
        if (!((E[v] >> lane()) & 1ull)) continue;
        const long long kv = (long long)v * 64 + lane();
        H::back_message(&cw[kv], delta[kv], sigma[kv]);
    }
}

// E = words of level d, D = words of level d+1; a wave per chunk of a vertex in E.
// Lanes load 64 edges' successor words at a time; then only the edges with a common
// bit are walked (lanes = sources), each one gather of the successor's c row.
template <class H>
__global__ __launch_bounds__(kBB) void k_mb_back(long long nch, const Chunk *__restrict__ ch, const int *__restrict__ col,
                                                 const uint64_t *__restrict__ E, const uint64_t *__restrict__ D,
                                                 const double *__restrict__ sigma, float *__restrict__ cw,
                                                 float *__restrict__ delta) {
    __shared__ int s_u[kBB];
    __shared__ uint64_t s_f[kBB];
    const int wl = threadIdx.x & ~(MG_WAVE - 1);
    for_active_chunks(nch, ch, [&](const Chunk &k) { return E[k.v]; }, [&](const Chunk &k, const uint64_t word) {
        const bool mine = (word >> lane()) & 1ull;
        const double sv = mine ? sigma[(long long)k.v * 64 + lane()] : 0.0;
        KSum acc;
        for (int p0 = k.p0; p0 < k.p1; p0 += MG_WAVE) {
            const int p = p0 + lane();
            int w = 0;
            uint64_t f = 0;
            if (p < k.p1) { w = ld_stream(&col[p]); f = D[w] & word; }
            uint64_t hits = ballot(f != 0ull);
            if (!hits) continue;
            wave_sync();
            s_u[threadIdx.x] = w;
            s_f[threadIdx.x] = f;
            wave_sync();
            while (hits) {  // kBackHitStep matched edges per step: their row gathers overlap
                int j[kBackHitStep];
#pragma unroll
                for (int q = 0; q < kBackHitStep; ++q) { j[q] = hits ? __builtin_ctzll(hits) : -1; hits &= hits - 1; }
                float c[kBackHitStep];
#pragma unroll
                for (int q = 0; q < kBackHitStep; ++q)
                    c[q] = (j[q] >= 0 && ((s_f[wl + j[q]] >> lane()) & 1ull)) ? cw[(long long)s_u[wl + j[q]] * 64 + lane()] : 0.f;
#pragma unroll
                for (int q = 0; q < kBackHitStep; ++q)
                    if (c[q] != 0.f) ms_term<H>(acc, sv, c[q]);
            }
        }
        if (mine) {
            const long long kv = (long long)k.v * 64 + lane();
            if (k.multi) {
                atomicAdd(&delta[kv], acc.value());
            } else {  // final: also its c for the level above
                H::back_update(&delta[kv], &cw[kv], acc.value(), sv);
            }
        }
    });
}

// bc[v] += sum over slots s of delta[v][s]; the slots v is a source of (level 0) hold
// zero, as the backward pass stops at level 1
__global__ __launch_bounds__(256) void k_ms_bc_sum(int nv, int nslots, const float *__restrict__ delta, float *__restrict__ bc) {
    // a wave takes 64 vertices: their rows pass through an LDS tile (one coalesced row
    // load each), then lane l sums row l alone, with no cross-lane reduction per vertex
    // (one vertex a wave with a wave-wide compensated reduction: 0.39 ms at RMAT-22)
    __shared__ float s_t[256 / MG_WAVE][MG_WAVE * (MG_WAVE + 1)];
    float *t = s_t[threadIdx.x / MG_WAVE];
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long v0 = wave * MG_WAVE; v0 < nv; v0 += nw * MG_WAVE) {
        const int nl = (int)min((long long)MG_WAVE, nv - v0);
        wave_sync();  // the previous tile is consumed
        for (int l = 0; l < nl; ++l) t[l * (MG_WAVE + 1) + lane()] = delta[(v0 + l) * 64 + lane()];
        wave_sync();
        if (lane() < nl) {
            KSum k;
            for (int q = 0; q < nslots; ++q) k.add(t[lane() * (MG_WAVE + 1) + q]);
            bc[v0 + lane()] += k.value();
        }
    }
}

// ---- SSSP --------------------------------------------------------------------------
// a wave per chunk, lanes = sources: relax dist[v][s] from neighbours whose changed
// bit s is set; the vertex's new changed word is the ballot of improvements
// Hooks per lane: candidates from get_partial_result, combined in a register with
// accumulate_nonAtomics (starting at initialize's identity), applied with
// update_condition, or with accumulate (atomicMin) on a vertex split over chunks.
template <class H>
__global__ __launch_bounds__(kBB) void k_mb_sssp(const BCtl *ctl, long long nch, const Chunk *__restrict__ ch, const int *__restrict__ col,
                                                 const int *__restrict__ wgt, uint64_t *CA,
                                                 uint64_t *CB, int *__restrict__ dist) {
    if (!level_runs(ctl, 0)) return;
    const int it = ctl ? ctl->it : 0;
    uint64_t *C, *Cn;
    cur_next(ctl, CA, CB, &C, &Cn);
    __shared__ int s_u[kBB], s_w[kBB];
    __shared__ uint64_t s_f[kBB];
    const int wl = threadIdx.x & ~(MG_WAVE - 1);
    const long long wave = ((long long)blockIdx.x * kBB + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * kBB) >> 6;
    for (long long c = wave; c < nch; c += nw) {
        const Chunk k = ch[c];
        const long long kv = (long long)k.v * 64 + lane();
        const int cur = dist[kv];
        int best;
        H::initialize(&best);
        for (int p0 = k.p0; p0 < k.p1; p0 += MG_WAVE) {
            const int p = p0 + lane();
            int u = 0, wu = 0;
            uint64_t f = 0;
            if (p < k.p1) { u = ld_stream(&col[p]); f = C[u]; }
            if (!wave_or(f)) continue;
            if (p < k.p1) wu = ld_stream(&wgt[p]);
            wave_sync();
            s_u[threadIdx.x] = u;
            s_w[threadIdx.x] = wu;
            s_f[threadIdx.x] = f;
            wave_sync();
            uint64_t hits = ballot(f != 0ull);  // only neighbours that changed for some source
            while (hits) {  // kSsspHitStep neighbours per step: their row gathers overlap
                int j[kSsspHitStep];
#pragma unroll
                for (int q = 0; q < kSsspHitStep; ++q) { j[q] = hits ? __builtin_ctzll(hits) : -1; hits &= hits - 1; }
                int du[kSsspHitStep];
                bool on[kSsspHitStep];
#pragma unroll
                for (int q = 0; q < kSsspHitStep; ++q) {
                    on[q] = j[q] >= 0 && ((s_f[wl + j[q]] >> lane()) & 1ull);
                    du[q] = on[q] ? dist[(long long)s_u[wl + j[q]] * 64 + lane()] : 0;
                }
#pragma unroll
                for (int q = 0; q < kSsspHitStep; ++q)
                    if (on[q]) {
                        int cand;
                        H::get_partial_result(&cand, du[q], s_w[wl + j[q]]);
                        H::accumulate_nonAtomics(&best, cand, it);
                    }
            }
        }
        bool imp;
        if (k.multi) {  // other chunks of the vertex write too
            imp = best < cur;
            if (imp) H::accumulate(&dist[kv], best, it);
        } else {
            int old = cur, stamp;
            imp = H::update_condition(&old, &dist[kv], best, it, &stamp);
        }
        const uint64_t m = ballot(imp);
        if (lane() == 0 && m) {
            if (k.multi) atomicOr((unsigned long long *)&Cn[k.v], (unsigned long long)m);
            else Cn[k.v] = m;
        }
    }
}

// push sweep (few changed vertices): a wave per chunk of a vertex u with changed bits,
// lanes = sources; along each edge (u, w) lane s lowers dist[w][s] to
// dist[u][s] + weight (read first, atomicMin only when it improves) and the
// improvements are or-ed into w's next changed word
template <class H>
__global__ __launch_bounds__(kBB) void k_mb_sssp_push(const BCtl *ctl, long long nch, const Chunk *__restrict__ ch,
                                                      const int *__restrict__ col, const int *__restrict__ wgt, uint64_t *CA,
                                                      uint64_t *CB, int *__restrict__ dist) {
    if (!level_runs(ctl, 1)) return;
    const int it = ctl ? ctl->it : 0;
    uint64_t *C, *Cn;
    cur_next(ctl, CA, CB, &C, &Cn);
    for_active_chunks(nch, ch, [&](const Chunk &k) { return C[k.v]; }, [&](const Chunk &k, const uint64_t cu) {
        const bool mine = (cu >> lane()) & 1ull;
        const int du = mine ? dist[(long long)k.v * 64 + lane()] : (int)MG_INF;
        // edges kHitStep at a time: their distance-row reads are in flight together
        for (int p0 = k.p0; p0 < k.p1; p0 += kHitStep) {
            int w[kHitStep], nd[kHitStep], dw[kHitStep];
#pragma unroll
            for (int q = 0; q < kHitStep; ++q) {
                const int p = min(p0 + q, k.p1 - 1);
                w[q] = col[p];
                if (p0 + q < k.p1) H::get_partial_result(&nd[q], du, wgt[p]);
                else H::initialize(&nd[q]);
            }
#pragma unroll
            for (int q = 0; q < kHitStep; ++q) dw[q] = mine ? dist[(long long)w[q] * 64 + lane()] : (int)MG_INF;
#pragma unroll
            for (int q = 0; q < kHitStep; ++q) {
                bool imp = false;
                if (mine && nd[q] < dw[q]) { H::accumulate(&dist[(long long)w[q] * 64 + lane()], nd[q], it); imp = true; }
                const uint64_t m = ballot(imp);
                if (lane() == 0 && m) atomicOr((unsigned long long *)&Cn[w[q]], (unsigned long long)m);
            }
        }
    });
}

// vertex-major [v][64] -> source-major [s][v]
// (row s of the output is lane perm[s]; no perm: lane s)
template <class T>
__global__ void k_ms_transpose(int nv, int nsrc, const int *__restrict__ perm, const T *__restrict__ in, T *__restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)nv * nsrc;
         i += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(i / nv), v = (int)(i % nv);
        out[i] = in[(long long)v * 64 + (perm ? perm[s] : s)];
    }
}

}  // namespace

struct mg_batch {
    int kind = 0;
    int nv = 0;
    long long ne = 0;
    int max_sources = 64;
    mg_options opt{};
    hipStream_t stream = nullptr;
    int *row = nullptr, *col = nullptr, *w = nullptr;
    uint64_t *vis = nullptr, *F[2] = {nullptr, nullptr}, *D = nullptr, *E = nullptr;
    int *Q[2] = {nullptr, nullptr};   // queues of F[0], F[1] on sparse levels
    int *lev = nullptr, *dist = nullptr, *src = nullptr, *slot = nullptr;
    int seed_cap = 0, nseeds = 0;  // device capacity of src/slot, sources of the last run
    bool slotted = false;          // the seeds carry a slot list (the caller's, or the lane order below)
    // Lane order of a run without a slot list: sources sorted by their 2-hop reach
    // (sum of their out-neighbours' out-degrees, `reach`), so sources whose levels run
    // alike sit in neighbouring lanes.  A matched edge of a BC pull or backward level
    // then needs source bits from fewer 16-lane groups, i.e. fewer 128-byte lines of the
    // gathered sigma / c row (RMAT-20, 64 sources: 23 % fewer sigma lines, 12 % fewer c
    // lines).  perm[i] = lane of output row i (mg_batch_get undoes the order).
    std::vector<long long> reach;
    int *perm = nullptr;
    bool permuted = false;
    uint64_t all = 0;              // the slots in use
    double *sigma = nullptr;      // BC path counts
    float *delta = nullptr, *bc = nullptr;
    float *cw = nullptr;          // BC: (1 + delta) / sigma of the finished levels
    int *mv = nullptr;            // BC: the vertices whose edges span several chunks
    int nmv = 0;
    Chunk *chunks = nullptr;      // static edge chunks
    long long nch = 0;
    // a directed graph's in-edges for pull levels (transposed CSR, its own chunks);
    // null on a symmetric graph, whose rows serve both directions
    Chunk *tchunks = nullptr;
    long long ntch = 0;
    int *trow = nullptr, *tcol = nullptr, *tw = nullptr;
    const Chunk *in_chunks() const { return tchunks ? tchunks : chunks; }
    long long in_nch() const { return tchunks ? ntch : nch; }
    const int *in_row() const { return tchunks ? trow : row; }
    const int *in_col() const { return tchunks ? tcol : col; }
    const int *in_w() const { return tchunks ? tw : w; }
    // BFS pulls: in-edge chunks of the vertices above `light` in-edges only (the
    // lighter ones are walked a lane per vertex)
    Chunk *hchunks = nullptr;
    long long nhch = 0;
    int light = 32;               // BFS pulls: the lane / wave split (mg_options.batch_light)
    // BFS: new-bit words of the first kDenseLog dense levels ([level][nv + 1]) and their
    // level numbers, expanded into the level rows after the run (k_mb_expand)
    uint64_t *dlog = nullptr;
    int *dlev = nullptr;
    int *logv = nullptr;          // (vertex, word) entries of every level (BC)
    uint64_t *logw = nullptr;
    long long log_cap = 0;
    int *part_n = nullptr, *part_x = nullptr;  // per-workgroup frontier size, largest degree
    long long *part_e = nullptr, *tot = nullptr, *h_tot = nullptr;
    int nsrc = 0, levels = 0;     // slots of the last run, levels it swept
    BCtl *ctl = nullptr, *h_ctl = nullptr;  // device level control (BFS, SSSP)
    BCtl *ctl2 = nullptr;                   // the sparse-only graph's two control copies
    FStat *fst = nullptr;                   // and its three frontier-total slots
    hipGraphExec_t gexec = nullptr;         // K captured levels (all level kinds)
    hipGraphExec_t gsparse = nullptr;       // K captured sparse levels (one node each)
    hipGraphExec_t gsparse_long = nullptr;  // the same with kSparseLevels levels, for long tails
    int gexec_levels = 0;
    PCtl *pc = nullptr;           // counters of the persistent sparse levels
    int persist_grid = 0;         // workgroups of its cooperative launch (0: not available)
    uint64_t *eq = nullptr;       // BFS: the entry queues of the persistent levels (2 x 3 x ecap words)
    int ecap = 0;
    long long stats[5] = {0, 0, 0, 0, 0};  // of the last BFS / SSSP run (mg_batch_stats)
    int auto_delta = 0;           // SSSP near-far band by default: a quarter of the mean weight
    int mirror_seq = 0;           // host checks of the control block so far (k_mb_mirror)
    int cus = 256;                // compute units of the device
};

static void batch_free(mg_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->gsparse) (void)hipGraphExecDestroy(b->gsparse);
    if (b->gsparse_long) (void)hipGraphExecDestroy(b->gsparse_long);
    void *ptrs[] = {b->Q[0], b->Q[1], b->row, b->col, b->w, b->vis, b->F[0], b->F[1], b->D, b->E, b->lev, b->dist, b->src, b->slot, b->sigma,
                    b->delta, b->bc, b->chunks, b->logv, b->logw, b->cw, b->part_n, b->part_x, b->part_e, b->tot, b->ctl, b->pc,
                    b->tchunks, b->trow, b->tcol, b->tw, b->hchunks, b->dlog, b->dlev, b->perm, b->mv, b->ctl2, b->fst, b->eq};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    if (b->h_tot) (void)hipHostFree(b->h_tot);
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int mg_batch_build(const mg_graph *g, int kind, int max_sources, const mg_options *opt, mg_batch **out) {
    if (!g || !out) return mg_fail(MG_ERR_ARG, "null argument");
    *out = nullptr;
    if (kind < MG_BATCH_BFS || kind > MG_BATCH_SSSP) return mg_fail(MG_ERR_ARG, "unknown batch kind %d", kind);
    if (max_sources < 1 || max_sources > 64) return mg_fail(MG_ERR_LIMIT, "a batch holds 1..64 sources (asked %d)", max_sources);
    if (g->wide()) return mg_fail(MG_ERR_LIMIT, "batched runs need int32 row offsets (%lld edges)", (long long)g->ne);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return mg_fail(MG_ERR_DEVICE, "no HIP device available");
    mg_batch *b = new mg_batch();
    b->kind = kind;
    b->nv = g->nv;
    b->ne = g->ne;
    b->max_sources = max_sources;
    if (opt) b->opt = *opt; else mg_default_options(&b->opt);
    if (b->opt.batch_light > 0) b->light = std::min(b->opt.batch_light, 1024);
    // Defaults: a grid of 4096 workgroups (the most the per-workgroup partial arrays
    // hold); chunks of 1024 edges for BFS, of 128 for BC and SSSP, whose dense levels
    // walk every matched edge's row per chunk, so a wave on a long chunk of a hub held
    // up the end of each sweep (RMAT-22, 64 sources: SSSP 54.5 -> 31.8 ms, BC 22.6 ->
    // 19.4 ms against 1024-edge chunks on 2048 workgroups; BFS, which stops at the
    // first parent, 1.94 ms with 1024 and 2.24 with 128)
    if (b->opt.grid_blocks <= 0 || b->opt.grid_blocks > 4096) b->opt.grid_blocks = 4096;
    if (b->opt.batch_chunk <= 0) b->opt.batch_chunk = kind == MG_BATCH_BFS ? kCH : kind == MG_BATCH_SSSP ? 64 : 128;
    auto fail = [&](int rc) { batch_free(b); return rc; };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(mg_fail(MG_ERR_DEVICE, "stream"));
    // static edge chunks (a hub's edges are cut into pieces of at most `chunk`)
    const int chunk = b->opt.batch_chunk > 0 ? std::max(64, b->opt.batch_chunk) : kCH;
    auto cut = [&](const int32_t *row, std::vector<Chunk> &ch) {
        ch.reserve((size_t)g->nv + (size_t)(g->ne / chunk) + 1);
        for (int v = 0; v < g->nv; ++v) {
            const int p0 = row[v], p1 = row[v + 1];
            if (p1 == p0) continue;
            const int multi = (p1 - p0) > chunk ? 1 : 0;
            for (int p = p0; p < p1; p += chunk) ch.push_back(Chunk{v, p, std::min(p1, p + chunk), multi});
        }
    };
    std::vector<Chunk> ch, tch;
    cut(g->row.data(), ch);
    b->nch = (long long)ch.size();
    // directed: pull levels read in-edges from the transposed CSR (counting sort by
    // destination; scanning sources in order leaves every row sorted by source)
    std::vector<int32_t> trow, tcol, tw;
    if (!g->symmetric && g->nv > 0) {
        trow.assign((size_t)g->nv + 1, 0);
        tcol.resize((size_t)g->ne);
        if (kind == MG_BATCH_SSSP) tw.resize((size_t)g->ne);
        for (long long p = 0; p < g->ne; ++p) ++trow[(size_t)g->col[p] + 1];
        for (int v = 0; v < g->nv; ++v) trow[v + 1] += trow[v];
        std::vector<int32_t> cur(trow.begin(), trow.end() - 1);
        for (int u = 0; u < g->nv; ++u)
            for (int p = g->row[u]; p < g->row[u + 1]; ++p) {
                const int32_t q = cur[g->col[p]]++;
                tcol[q] = u;
                if (!tw.empty()) tw[q] = g->w[p];
            }
        cut(trow.data(), tch);
        b->ntch = (long long)tch.size();
    }
    std::vector<Chunk> hch;
    if (kind == MG_BATCH_BFS)
        for (const Chunk &c : (trow.empty() ? ch : tch)) {
            const int32_t *r = trow.empty() ? g->row.data() : trow.data();
            if (r[c.v + 1] - r[c.v] > b->light) hch.push_back(c);
        }
    b->nhch = (long long)hch.size();
    if (b->opt.batch_lanes == 0) {  // 2-hop reach of every vertex, for the lane order
        b->reach.assign((size_t)g->nv, 0);
        for (int v = 0; v < g->nv; ++v) {
            long long r = 0;
            for (int p = g->row[v]; p < g->row[v + 1]; ++p) r += g->row[g->col[p] + 1] - g->row[g->col[p]];
            b->reach[v] = r;
        }
    }
    const long long nv = b->nv, ne = std::max<long long>(1, b->ne), slots = nv * 64 + 64;
    bool ok = hipMalloc(&b->row, sizeof(int) * (nv + 1)) == hipSuccess && hipMalloc(&b->col, sizeof(int) * ne) == hipSuccess &&
              hipMalloc(&b->chunks, sizeof(Chunk) * std::max<long long>(1, b->nch)) == hipSuccess &&
              hipMalloc(&b->vis, sizeof(uint64_t) * (nv + 1)) == hipSuccess &&
              hipMalloc(&b->F[0], sizeof(uint64_t) * (nv + 1)) == hipSuccess &&
              hipMalloc(&b->F[1], sizeof(uint64_t) * (nv + 1)) == hipSuccess &&
              hipMalloc(&b->part_n, sizeof(int) * 4096) == hipSuccess &&
              hipMalloc(&b->part_x, sizeof(int) * 4096) == hipSuccess &&
              hipMalloc(&b->part_e, sizeof(long long) * 4096) == hipSuccess &&
              hipMalloc(&b->tot, sizeof(long long) * 4) == hipSuccess && hipMalloc(&b->perm, sizeof(int) * 64) == hipSuccess &&
              hipHostMalloc(&b->h_tot, sizeof(long long) * 4, hipHostMallocDefault) == hipSuccess &&
              hipMalloc(&b->ctl, sizeof(BCtl)) == hipSuccess && hipMalloc(&b->pc, sizeof(PCtl)) == hipSuccess &&
              hipMalloc(&b->ctl2, 2 * sizeof(BCtl)) == hipSuccess && hipMalloc(&b->fst, 3 * sizeof(FStat)) == hipSuccess &&
              hipMalloc(&b->Q[0], sizeof(int) * (nv + 1)) == hipSuccess && hipMalloc(&b->Q[1], sizeof(int) * (nv + 1)) == hipSuccess &&
              hipHostMalloc(&b->h_ctl, sizeof(BCtl), hipHostMallocCoherent) == hipSuccess;  // written by k_mb_mirror
    if (ok && kind != MG_BATCH_SSSP) ok = hipMalloc(&b->lev, sizeof(int) * slots) == hipSuccess;
    if (ok && kind == MG_BATCH_SSSP) {
        ok = hipMalloc(&b->w, sizeof(int) * ne) == hipSuccess && hipMalloc(&b->dist, sizeof(int) * slots) == hipSuccess;
        double sw = 0.0;
        for (long long p = 0; p < g->ne; ++p) sw += (double)g->w[p];
        b->auto_delta = g->ne > 0 ? std::max(1, (int)(sw / (double)g->ne / 4.0)) : 1;
    }
    if (ok && kind == MG_BATCH_BFS) {
        ok = hipMalloc(&b->hchunks, sizeof(Chunk) * std::max<long long>(1, b->nhch)) == hipSuccess &&
             hipMalloc(&b->dlog, sizeof(uint64_t) * kDenseLog * (nv + 1)) == hipSuccess &&
             hipMalloc(&b->dlev, sizeof(int) * kDenseLog) == hipSuccess;
        if (ok && b->nhch > 0 && hipMemcpy(b->hchunks, hch.data(), sizeof(Chunk) * b->nhch, hipMemcpyHostToDevice) != hipSuccess)
            return fail(mg_fail(MG_ERR_DEVICE, "batch: chunk upload failed"));
    }
    if (ok && !trow.empty()) {
        ok = hipMalloc(&b->tchunks, sizeof(Chunk) * std::max<long long>(1, b->ntch)) == hipSuccess &&
             hipMalloc(&b->trow, sizeof(int) * (nv + 1)) == hipSuccess &&
             hipMalloc(&b->tcol, sizeof(int) * ne) == hipSuccess &&
             (tw.empty() || hipMalloc(&b->tw, sizeof(int) * ne) == hipSuccess);
        if (ok && ((b->ntch > 0 && hipMemcpy(b->tchunks, tch.data(), sizeof(Chunk) * b->ntch, hipMemcpyHostToDevice) != hipSuccess) ||
                   hipMemcpy(b->trow, trow.data(), sizeof(int) * (nv + 1), hipMemcpyHostToDevice) != hipSuccess ||
                   (g->ne > 0 && hipMemcpy(b->tcol, tcol.data(), sizeof(int) * g->ne, hipMemcpyHostToDevice) != hipSuccess) ||
                   (b->tw && g->ne > 0 && hipMemcpy(b->tw, tw.data(), sizeof(int) * g->ne, hipMemcpyHostToDevice) != hipSuccess)))
            return fail(mg_fail(MG_ERR_DEVICE, "batch: upload of the transposed CSR failed"));
    }
    if (ok && kind == MG_BATCH_BC) {
        b->log_cap = nv * 8 + 1024;  // grown on demand
        ok = hipMalloc(&b->sigma, sizeof(double) * slots) == hipSuccess && hipMalloc(&b->delta, sizeof(float) * slots) == hipSuccess &&
             hipMalloc(&b->cw, sizeof(float) * slots) == hipSuccess &&
             hipMalloc(&b->bc, sizeof(float) * (nv + 1)) == hipSuccess && hipMalloc(&b->D, sizeof(uint64_t) * (nv + 1)) == hipSuccess &&
             hipMalloc(&b->E, sizeof(uint64_t) * (nv + 1)) == hipSuccess &&
             hipMalloc(&b->logv, sizeof(int) * b->log_cap) == hipSuccess &&
             hipMalloc(&b->logw, sizeof(uint64_t) * b->log_cap) == hipSuccess;
        std::vector<int> mv;
        for (const Chunk &c : ch)
            if (c.multi && c.p0 == g->row[c.v]) mv.push_back(c.v);
        b->nmv = (int)mv.size();
        if (ok) ok = hipMalloc(&b->mv, sizeof(int) * std::max<size_t>(1, mv.size())) == hipSuccess;
        if (ok && !mv.empty() && hipMemcpy(b->mv, mv.data(), sizeof(int) * mv.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(mg_fail(MG_ERR_DEVICE, "batch: upload of the split-vertex list failed"));
    }
    if (!ok) return fail(mg_fail(MG_ERR_NOMEM, "batch: device allocation failed"));
    {
        int dev = 0, cu = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            cu > 0)
            b->cus = cu;
    }
    if (kind != MG_BATCH_BC && b->opt.persist_levels != 0) {  // one resident workgroup per CU
        int dev = 0, cus = 0, per = 0;
        const void *fn = kind == MG_BATCH_SSSP ? (const void *)k_mbs_persist : (const void *)k_mbp_bfs;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, kBB, 0) == hipSuccess && per > 0 && cus > 0)
            b->persist_grid = cus * std::min(per, MG_PERSIST_PER_CU);
        if (b->persist_grid > 0 && kind == MG_BATCH_BFS) {
            b->ecap = std::min<int>(nv + 1024, 1 << 27);  // entry words are addressed with 31-bit byte offsets
            if (hipMalloc(&b->eq, sizeof(uint64_t) * 6 * (size_t)b->ecap) != hipSuccess)
                return fail(mg_fail(MG_ERR_NOMEM, "batch: entry queues of the persistent levels"));
        }
    }
    if (hipMemcpy(b->row, g->row.data(), sizeof(int) * (nv + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (g->ne > 0 && hipMemcpy(b->col, g->col.data(), sizeof(int) * g->ne, hipMemcpyHostToDevice) != hipSuccess) ||
        (b->w && g->ne > 0 && hipMemcpy(b->w, g->w.data(), sizeof(int) * g->ne, hipMemcpyHostToDevice) != hipSuccess) ||
        (b->nch > 0 && hipMemcpy(b->chunks, ch.data(), sizeof(Chunk) * b->nch, hipMemcpyHostToDevice) != hipSuccess))
        return fail(mg_fail(MG_ERR_DEVICE, "batch: upload failed"));
    *out = b;
    return MG_OK;
}

extern "C" int mg_batch_release(mg_batch *b) {
    if (!b) return mg_fail(MG_ERR_ARG, "null batch");
    batch_free(b);
    return MG_OK;
}

// compact the n nonzero words of W (n: the frontier size the level's counts gave)
// into the log at `base`
static int compact(mg_batch *b, const uint64_t *W, long long base, long long n) {
    hipStream_t s = b->stream;
    const int gq = grid_for(b->nv, 2048, 2048);
    if (base + n > b->log_cap) {
        long long cap = std::max(b->log_cap * 2, base + n + 1024);
        int *nv2 = nullptr;
        uint64_t *nw2 = nullptr;
        MG_HIP(hipMalloc(&nv2, sizeof(int) * cap));
        MG_HIP(hipMalloc(&nw2, sizeof(uint64_t) * cap));
        if (base > 0) {
            MG_HIP(hipMemcpyAsync(nv2, b->logv, sizeof(int) * base, hipMemcpyDeviceToDevice, s));
            MG_HIP(hipMemcpyAsync(nw2, b->logw, sizeof(uint64_t) * base, hipMemcpyDeviceToDevice, s));
        }
        MG_HIP(hipStreamSynchronize(s));
        hipFree(b->logv);
        hipFree(b->logw);
        b->logv = nv2;
        b->logw = nw2;
        b->log_cap = cap;
    }
    hipLaunchKernelGGL(k_nz_count, dim3(gq), dim3(kBB), 0, s, b->nv, W, b->part_n);
    hipLaunchKernelGGL(k_nz_scan, dim3(1), dim3(1024), 0, s, b->part_n, gq, b->tot + 2);
    hipLaunchKernelGGL(k_nz_write, dim3(gq), dim3(kBB), 0, s, b->nv, W, b->part_n, base, b->log_cap, b->logv, b->logw);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// frontier size and degree sum from the per-workgroup partials
static int read_counts(mg_batch *b, int nparts, long long *n, long long *mf) {
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, b->stream, b->part_n, b->part_e, nparts, b->tot);
    MG_HIP(hipMemcpyAsync(b->h_tot, b->tot, sizeof(long long) * 2, hipMemcpyDeviceToHost, b->stream));
    MG_HIP(hipStreamSynchronize(b->stream));
    *n = b->h_tot[0];
    if (mf) *mf = b->h_tot[1];
    return MG_OK;
}

static int run_bfs_bc(mg_batch *b) {
    hipStream_t s = b->stream;
    const int nv = b->nv;
    const bool bc = b->kind == MG_BATCH_BC;
    const uint64_t all = b->all;
    const int G = b->opt.grid_blocks > 0 ? b->opt.grid_blocks : 2048;
    const float alpha = b->opt.alpha > 0 ? b->opt.alpha : 15.0f;
    const long long slots = (long long)nv * 64 + 64;
    const int gv = grid_for(nv, kBB, G);
    const int gc = grid_for(b->nch, kBB / MG_WAVE, G);
    const int gci = grid_for(b->in_nch(), kBB / MG_WAVE, G);  // pulls walk in-edge chunks
    const int gch = grid_for(b->nhch, kBB / MG_WAVE, G);     // BFS pulls: heavy chunks only
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, b->vis, (long long)nv + 1, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, b->F[0], (long long)nv + 1, 0ull);
    hipLaunchKernelGGL(k_fill_i32b, dim3(grid_for(slots, 256, 8192)), dim3(256), 0, s, b->lev, slots, (int)MG_INF);
    if (bc) {
        hipLaunchKernelGGL(k_fill_f64b, dim3(grid_for(slots, 256, 8192)), dim3(256), 0, s, b->sigma, slots, 0.0);
        hipLaunchKernelGGL(k_fill_f32b, dim3(grid_for(slots, 256, 8192)), dim3(256), 0, s, b->delta, slots, 0.0f);
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, b->D, (long long)nv + 1, 0ull);
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, b->E, (long long)nv + 1, 0ull);
    }
    hipLaunchKernelGGL(k_ms_seed, dim3(grid_for(b->nseeds, 256, 4096)), dim3(256), 0, s, b->nseeds, b->src,
                       b->slotted ? b->slot : nullptr, b->vis, b->F[0], b->lev, bc ? b->sigma : nullptr, (int *)nullptr);
    std::vector<long long> off{0};  // BC level log offsets
    long long n = 0, mf = 0;
    hipLaunchKernelGGL(k_mb_count, dim3(gv), dim3(kBB), 0, s, (const BCtl *)nullptr, nv, b->row, b->F[1], b->F[0], b->part_n, b->part_e, b->part_x);
    MG_TRY(read_counts(b, gv, &n, &mf));
    if (bc) { MG_TRY(compact(b, b->F[0], 0, n)); off.push_back(n); }
    int cur = 0, it = 1;
    while (n > 0) {
        uint64_t *F = b->F[cur], *Fn = b->F[cur ^ 1];
        bool push = (double)mf < (double)b->ne / alpha;
        if (b->opt.mode == MG_MODE_SPARSE) push = true;
        if (b->opt.mode == MG_MODE_DENSE) push = false;
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, Fn, (long long)nv + 1, 0ull);
        if (push) {
            if (bc) hipLaunchKernelGGL((k_mb_push<true>), dim3(gc), dim3(kBB), 0, s, (const BCtl *)nullptr, b->nch, b->chunks, b->col, F, b->vis, Fn, b->sigma);
            else hipLaunchKernelGGL((k_mb_push<false>), dim3(gc), dim3(kBB), 0, s, (const BCtl *)nullptr, b->nch, b->chunks, b->col, F, b->vis, Fn, b->sigma);
        } else {
            if (bc) hipLaunchKernelGGL(k_mb_pull_bc<MG_BC_HOOKS>, dim3(gci), dim3(kBB), 0, s, it, b->in_nch(), b->in_chunks(), all, b->in_col(), F, b->vis, Fn, b->sigma);
            else {
                hipLaunchKernelGGL(k_mb_pull_bfs_light, dim3(gv), dim3(kBB), 0, s, (const BCtl *)nullptr, nv, b->light, b->in_row(), b->in_col(),
                                   all, F, b->vis, Fn);
                hipLaunchKernelGGL(k_mb_pull_bfs, dim3(gch), dim3(kBB), 0, s, (const BCtl *)nullptr, b->nhch, b->hchunks, all,
                                   b->in_col(), F, b->vis, Fn);
            }
        }
        hipLaunchKernelGGL(k_mb_apply, dim3(gv), dim3(kBB), 0, s, (const BCtl *)nullptr, nv, b->row, b->vis, F, Fn, b->lev, it, b->part_n, b->part_e, b->part_x,
                           (uint64_t *)nullptr, (int *)nullptr, 0);
        MG_TRY(read_counts(b, gv, &n, &mf));
        if (bc && n > 0) { MG_TRY(compact(b, Fn, off.back(), n)); off.push_back(off.back() + n); }
        cur ^= 1;
        ++it;
    }
    b->levels = it - 1;
    if (!bc) return MG_OK;
    // backward: levels L-1 .. 1, entries off[l] .. off[l+1] (level 0 holds the sources,
    // whose own dependency BC leaves out).  E holds level l's words, D level l + 1's;
    // after a level D is cleared and the buffers swap roles.
    const int L = (int)off.size() - 1;
    uint64_t *E = b->E, *D = b->D;
    for (int l = L - 1; l >= 1; --l) {
        const long long n0 = off[l + 1] - off[l];
        const long long n1 = (l + 1 < L) ? off[l + 2] - off[l + 1] : 0;
        if (n0 > 0) {
            hipLaunchKernelGGL(k_ms_scatter_words, dim3(grid_for(n0, 256, 4096)), dim3(256), 0, s, n0, b->logv + off[l], b->logw + off[l], E, 0);
            hipLaunchKernelGGL(k_mb_back<MG_BC_HOOKS>, dim3(gc), dim3(kBB), 0, s, b->nch, b->chunks, b->col, E, D, b->sigma, b->cw, b->delta);
            if (b->nmv > 0)
                hipLaunchKernelGGL(k_ms_cw<MG_BC_HOOKS>, dim3(grid_for(b->nmv, 4, 1024)), dim3(256), 0, s, b->nmv, b->mv, E, b->sigma, b->delta, b->cw);
        }
        if (n1 > 0) hipLaunchKernelGGL(k_ms_scatter_words, dim3(grid_for(n1, 256, 4096)), dim3(256), 0, s, n1, b->logv + off[l + 1], b->logw + off[l + 1], D, 1);
        std::swap(E, D);
    }
    hipLaunchKernelGGL(k_ms_bc_sum, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, nv, b->nsrc, b->delta, b->bc);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// the level rule of BFS / SSSP runs (the control kernels' and the persistent kernel's)
static LevelRule rule_of(const mg_batch *b) {
    const float alpha = b->opt.alpha > 0 ? b->opt.alpha : 15.0f;
    const float beta = b->opt.beta > 0 ? b->opt.beta : 18.0f;
    const int chunk = b->opt.batch_chunk > 0 ? std::max(64, b->opt.batch_chunk) : kCH;
    const int delta = b->kind != MG_BATCH_SSSP ? 0 : b->opt.sssp_delta > 0 ? b->opt.sssp_delta : b->opt.sssp_delta < 0 ? 0 : b->auto_delta;
    return LevelRule{b->ne, b->nv, alpha, beta, b->opt.mode, chunk, delta};
}

// BFS (without BC) and SSSP: levels run under the device control block, K at a time
// from a captured hipGraph; the host reads the control block once per K levels
static int capture_levels(mg_batch *b, int K, bool sparse_only, hipGraphExec_t *out) {
    hipStream_t s = b->stream;
    const int nv = b->nv;
    const int G = b->opt.grid_blocks > 0 ? b->opt.grid_blocks : 2048;
    const int gv = grid_for(nv, kBB, G);
    const int gc = grid_for(b->nch, kBB / MG_WAVE, G);
    const int gci = grid_for(b->in_nch(), kBB / MG_WAVE, G);  // pulls walk in-edge chunks
    const int gch = grid_for(b->nhch, kBB / MG_WAVE, G);     // BFS pulls: heavy chunks only
    const bool sssp = b->kind == MG_BATCH_SSSP;
    hipGraph_t g = nullptr;
    MG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const LevelRule rule = rule_of(b);
    const int gq = grid_for(nv, 2048, 2048);
    // sparse levels: mg_options.sparse_grid workgroups, else MG_SPARSE_PER_CU per CU
    const int gsp = std::min(gv, std::max(1, b->opt.sparse_grid > 0 ? b->opt.sparse_grid : b->cus * MG_SPARSE_PER_CU));
    auto sparse_level = [&]() {
        if (sssp)
            hipLaunchKernelGGL(k_mbs_level_sssp, dim3(gsp), dim3(kBB), 0, s, b->ctl, b->row, b->col, b->w, b->F[0], b->F[1], b->Q[0],
                               b->Q[1], b->dist, b->part_n, b->part_e, b->part_x, gv);
        else
            hipLaunchKernelGGL(k_mbs_level_bfs, dim3(gsp), dim3(kBB), 0, s, b->ctl, b->row, b->col, b->F[0], b->F[1], b->vis, b->Q[0],
                               b->Q[1], b->lev, b->part_n, b->part_e, b->part_x, gv);
    };
    if (sparse_only) {  // one node per level, between an entry and an exit node
        hipLaunchKernelGGL(k_mbs_fused_enter, dim3(1), dim3(256), 0, s, (const BCtl *)b->ctl, b->ctl2, b->fst, b->part_n, b->part_e,
                           b->part_x, gv);
        for (int k = 0; k < K; ++k) {
            if (sssp)
                hipLaunchKernelGGL(k_mbs_fused<true>, dim3(gsp), dim3(kBB), 0, s, k, b->ctl2, b->fst, rule, b->row, b->col, b->w,
                                   b->F[0], b->F[1], b->vis, b->Q[0], b->Q[1], b->dist);
            else
                hipLaunchKernelGGL(k_mbs_fused<false>, dim3(gsp), dim3(kBB), 0, s, k, b->ctl2, b->fst, rule, b->row, b->col,
                                   (const int *)nullptr, b->F[0], b->F[1], b->vis, b->Q[0], b->Q[1], b->lev);
        }
        hipLaunchKernelGGL(k_mbs_fused_exit, dim3(1), dim3(256), 0, s, K, b->ctl, (const BCtl *)b->ctl2, (const FStat *)b->fst,
                           b->part_n, b->part_e, b->part_x, gv, (int)b->nv + 1);
    }
    for (int k = 0; k < K && !sparse_only; ++k) {
        hipLaunchKernelGGL(k_mb_ctl, dim3(1), dim3(256), 0, s, b->ctl, b->part_n, b->part_e, b->part_x, gv, rule);
        hipLaunchKernelGGL(k_mbs_qcount, dim3(gq), dim3(kBB), 0, s, (const BCtl *)b->ctl, nv, b->F[0], b->F[1], b->part_n);
        hipLaunchKernelGGL(k_mbs_qscan, dim3(1), dim3(1024), 0, s, b->ctl, b->part_n, gq);
        hipLaunchKernelGGL(k_mbs_qwrite, dim3(gq), dim3(kBB), 0, s, (const BCtl *)b->ctl, nv, b->F[0], b->F[1], b->part_n, b->Q[0],
                           b->Q[1]);
        hipLaunchKernelGGL(k_mb_fill_next, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, (const BCtl *)b->ctl, b->F[0],
                           b->F[1], (long long)nv + 1);
        if (sssp) {
            hipLaunchKernelGGL(k_mb_sssp_push<MG_SSSP_HOOKS>, dim3(gc), dim3(kBB), 0, s, (const BCtl *)b->ctl, b->nch, b->chunks, b->col, b->w,
                               b->F[0], b->F[1], b->dist);
            hipLaunchKernelGGL(k_mb_sssp<MG_SSSP_HOOKS>, dim3(gci), dim3(kBB), 0, s, (const BCtl *)b->ctl, b->in_nch(), b->in_chunks(), b->in_col(),
                               b->in_w(), b->F[0], b->F[1], b->dist);
            hipLaunchKernelGGL(k_mb_count, dim3(gv), dim3(kBB), 0, s, (const BCtl *)b->ctl, nv, b->row, b->F[0], b->F[1], b->part_n,
                               b->part_e, b->part_x);
        } else {
            hipLaunchKernelGGL((k_mb_push<false>), dim3(gc), dim3(kBB), 0, s, (const BCtl *)b->ctl, b->nch, b->chunks, b->col,
                               b->F[0], b->vis, b->F[1], (double *)nullptr);
            hipLaunchKernelGGL(k_mb_pull_bfs_light, dim3(gv), dim3(kBB), 0, s, (const BCtl *)b->ctl, nv, b->light, b->in_row(), b->in_col(),
                               0ull, b->F[0], b->vis, b->F[1]);
            hipLaunchKernelGGL(k_mb_pull_bfs, dim3(gch), dim3(kBB), 0, s, (const BCtl *)b->ctl, b->nhch, b->hchunks, 0ull,
                               b->in_col(), b->F[0], b->vis, b->F[1]);
            hipLaunchKernelGGL(k_mb_apply, dim3(gv), dim3(kBB), 0, s, (const BCtl *)b->ctl, nv, b->row, b->vis, b->F[0], b->F[1],
                               b->lev, 0, b->part_n, b->part_e, b->part_x, b->dlog, b->dlev, b->dlog ? kDenseLog : 0);
        }
        sparse_level();
    }
    hipError_t ce = hipStreamEndCapture(s, &g);
    if (ce != hipSuccess) return mg_fail(MG_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) return mg_fail(MG_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    return MG_OK;
}

static int run_levels(mg_batch *b) {
    hipStream_t s = b->stream;
    const int nv = b->nv;
    const bool sssp = b->kind == MG_BATCH_SSSP;
    const int G = b->opt.grid_blocks > 0 ? b->opt.grid_blocks : 2048;
    const int gv = grid_for(nv, kBB, G);
    const long long slots = (long long)nv * 64 + 64;
    const int K = b->opt.graph_levels > 0 ? b->opt.graph_levels : 8;
    if (!b->gexec || b->gexec_levels != K) {  // captured once per batch, on an idle stream
        if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        if (b->gsparse) { hipGraphExecDestroy(b->gsparse); b->gsparse = nullptr; }
        if (b->gsparse_long) { hipGraphExecDestroy(b->gsparse_long); b->gsparse_long = nullptr; }
        MG_HIP(hipStreamSynchronize(s));
        MG_TRY(capture_levels(b, K, false, &b->gexec));
        MG_TRY(capture_levels(b, K, true, &b->gsparse));
        if (K < kSparseLevels) MG_TRY(capture_levels(b, kSparseLevels, true, &b->gsparse_long));
        b->gexec_levels = K;
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, b->vis, (long long)nv + 1, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv + 1, 256, 4096)), dim3(256), 0, s, b->F[0], (long long)nv + 1, 0ull);
    // BFS level rows need no clearing: k_mb_expand writes every one after the run
    if (sssp) hipLaunchKernelGGL(k_fill_i32b, dim3(grid_for(slots, 256, 8192)), dim3(256), 0, s, b->dist, slots, (int)MG_INF);
    else if (!b->dlog) hipLaunchKernelGGL(k_fill_i32b, dim3(grid_for(slots, 256, 8192)), dim3(256), 0, s, b->lev, slots, (int)MG_INF);
    hipLaunchKernelGGL(k_ms_seed, dim3(grid_for(b->nseeds, 256, 4096)), dim3(256), 0, s, b->nseeds, b->src,
                       b->slotted ? b->slot : nullptr, b->vis, b->F[0], sssp ? (int *)nullptr : b->lev, (double *)nullptr,
                       sssp ? b->dist : (int *)nullptr);
    hipLaunchKernelGGL(k_mb_ctl_init, dim3(1), dim3(1), 0, s, b->ctl, (unsigned long long)b->all);
    hipLaunchKernelGGL(k_mb_count, dim3(gv), dim3(kBB), 0, s, (const BCtl *)nullptr, nv, b->row, b->F[1], b->F[0], b->part_n,
                       b->part_e, b->part_x);
    MG_HIP(hipGetLastError());
    memset(b->h_ctl, 0, sizeof(BCtl));  // the first replay is always the full graph
    const int max_levels = b->opt.persist_levels > 0 ? b->opt.persist_levels : (1 << 30);  // -1: unbounded
    long long trips = 0;
    bool was_sparse = false;  // the last replay was the sparse-only graph and ran on
    bool skip_persist = false;  // the last persistent launch found the frontier too large
    for (;;) {
        // after a sparse level the persistent kernel (or the sparse-only graph) runs on
        // until a level needs more
        const bool sparse_next = b->h_ctl->it > 0 && b->h_ctl->sparse && !b->h_ctl->need_full;
        const bool persist = sparse_next && b->persist_grid > 0 && !skip_persist;
        if (persist) {
            hipLaunchKernelGGL(k_fill_i32b, dim3(1), dim3(256), 0, s, (int *)b->pc, (long long)(sizeof(PCtl) / sizeof(int)), 0);
            if (sssp) {
                PersistArgs pa{b->ctl, b->pc, rule_of(b), b->row, b->col, b->w, b->F[0], b->F[1], b->vis, b->Q[0], b->Q[1],
                               b->dist, b->part_n, b->part_e, b->part_x, gv, max_levels};
                void *args[] = {&pa};
                MG_HIP(hipLaunchCooperativeKernel((const void *)k_mbs_persist, dim3(b->persist_grid), dim3(kBB), args, 0, s));
            } else {
                const size_t c = (size_t)b->ecap;
                EqArgs ea{b->ctl, b->pc, rule_of(b), b->row, b->col, b->F[0], b->F[1], b->vis, b->Q[0], b->Q[1], b->lev,
                          b->part_n, b->part_e, b->part_x,
                          {EQ{b->eq, b->eq + c, b->eq + 2 * c}, EQ{b->eq + 3 * c, b->eq + 4 * c, b->eq + 5 * c}},
                          b->ecap, gv, max_levels};
                void *args[] = {&ea};
                MG_HIP(hipLaunchCooperativeKernel((const void *)k_mbp_bfs, dim3(b->persist_grid), dim3(kBB), args, 0, s));
            }
        } else {
            // a tail that outlasted one short sparse replay gets the long one: fewer
            // host checks; a short tail does not pay the long graph's idle nodes
            hipGraphExec_t ge = !sparse_next ? b->gexec : (was_sparse && b->gsparse_long) ? b->gsparse_long : b->gsparse;
            MG_HIP(hipGraphLaunch(ge, s));
        }
        was_sparse = sparse_next;
        // the control block through the pinned mirror: spin on its sequence number
        // (after ~2 ms, block on the stream, which also reports a failed kernel)
        const int seq = ++b->mirror_seq;
        hipLaunchKernelGGL(k_mb_mirror, dim3(1), dim3(1), 0, s, (const BCtl *)b->ctl, b->h_ctl, seq);
        MG_HIP(hipGetLastError());
        {
            const auto t0 = std::chrono::steady_clock::now();
            for (unsigned k = 0; __atomic_load_n(&b->h_ctl->seq, __ATOMIC_ACQUIRE) != seq; ++k)
                if ((k & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
                    MG_HIP(hipStreamSynchronize(s));
                    if (__atomic_load_n(&b->h_ctl->seq, __ATOMIC_ACQUIRE) != seq)
                        MG_HIP(hipMemcpy(b->h_ctl, b->ctl, sizeof(BCtl), hipMemcpyDeviceToHost));
                    break;
                }
        }
        ++trips;
        if (b->h_ctl->err) return mg_fail(MG_ERR_DEVICE, "persistent levels: a barrier wait passed its bound");
        // too large for the persistent launch: one replay of the sparse-only graph first
        skip_persist = persist && b->h_ctl->big;
        if (b->h_ctl->done) break;
    }
    if (!sssp && b->dlog)
        hipLaunchKernelGGL(k_mb_expand, dim3(gv), dim3(kBB), 0, s, nv, std::min(b->h_ctl->ndense, kDenseLog), b->vis, b->dlog,
                           b->dlev, b->lev);
    MG_HIP(hipGetLastError());
    b->levels = b->h_ctl->it;
    b->stats[0] = b->levels;
    b->stats[1] = b->h_ctl->ndense;
    b->stats[2] = b->h_ctl->nsparse;
    b->stats[3] = b->h_ctl->npersist;
    b->stats[4] = trips;
    return MG_OK;
}

static int batch_run(mg_batch *b, int n, const int32_t *sources, const int32_t *slots, float *ms) {
    if (!b || (n > 0 && !sources)) return mg_fail(MG_ERR_ARG, "null argument");
    if (n < 1) return mg_fail(MG_ERR_ARG, "a batch needs at least one source");
    if (!slots && n > b->max_sources) return mg_fail(MG_ERR_LIMIT, "batch of %d sources (1..%d)", n, b->max_sources);
    int nslots = slots ? 0 : n;
    uint64_t all = 0;
    for (int i = 0; i < n; ++i) {
        if (sources[i] < 0 || sources[i] >= b->nv) return mg_fail(MG_ERR_ARG, "source %d outside [0,%d)", sources[i], b->nv);
        const int k = slots ? slots[i] : i;
        if (k < 0 || k >= b->max_sources) return mg_fail(MG_ERR_LIMIT, "slot %d outside [0,%d)", k, b->max_sources);
        nslots = std::max(nslots, k + 1);
        all |= 1ull << k;
    }
    if (n > b->seed_cap) {
        if (b->src) hipFree(b->src);
        if (b->slot) hipFree(b->slot);
        b->src = b->slot = nullptr;
        b->seed_cap = 0;
        const int cap = std::max(64, n);
        MG_HIP(hipMalloc(&b->src, sizeof(int) * cap));
        MG_HIP(hipMalloc(&b->slot, sizeof(int) * cap));
        b->seed_cap = cap;
    }
    MG_HIP(hipMemcpy(b->src, sources, sizeof(int) * n, hipMemcpyHostToDevice));
    if (slots) MG_HIP(hipMemcpy(b->slot, slots, sizeof(int) * n, hipMemcpyHostToDevice));
    b->slotted = slots != nullptr;
    b->permuted = false;
    if (!slots && !b->reach.empty() && n > 1) {  // sources in order of 2-hop reach (stable: ties keep theirs)
        int ord[64], lane_of[64];
        for (int i = 0; i < n; ++i) ord[i] = i;
        std::stable_sort(ord, ord + n, [&](int a, int c) { return b->reach[sources[a]] > b->reach[sources[c]]; });
        bool ident = true;
        for (int k = 0; k < n; ++k) { lane_of[ord[k]] = k; ident = ident && ord[k] == k; }
        if (!ident) {
            MG_HIP(hipMemcpy(b->slot, lane_of, sizeof(int) * n, hipMemcpyHostToDevice));
            MG_HIP(hipMemcpy(b->perm, lane_of, sizeof(int) * n, hipMemcpyHostToDevice));
            b->slotted = b->permuted = true;
        }
    }
    b->nseeds = n;
    b->nsrc = nslots;
    b->all = all;
    if (b->kind == MG_BATCH_BC)
        hipLaunchKernelGGL(k_fill_f32b, dim3(grid_for(b->nv + 1, 256, 4096)), dim3(256), 0, b->stream, b->bc, (long long)b->nv + 1, 0.0f);
    MG_HIP(hipStreamSynchronize(b->stream));
    hipEvent_t e0, e1;
    MG_HIP(hipEventCreate(&e0));
    MG_HIP(hipEventCreate(&e1));
    MG_HIP(hipEventRecord(e0, b->stream));
    int rc = b->kind == MG_BATCH_BC ? run_bfs_bc(b) : run_levels(b);
    hipEventRecord(e1, b->stream);
    hipError_t se = hipEventSynchronize(e1);
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != MG_OK) return rc;
    if (se != hipSuccess) return mg_fail(MG_ERR_DEVICE, "batch run failed: %s", hipGetErrorString(se));
    if (ms) *ms = t;
    return MG_OK;
}

extern "C" int mg_batch_run(mg_batch *b, int nsrc, const int32_t *sources, float *ms) {
    return batch_run(b, nsrc, sources, nullptr, ms);
}

extern "C" int mg_batch_run_slots(mg_batch *b, int n, const int32_t *sources, const int32_t *slots, float *ms) {
    if (!slots) return mg_fail(MG_ERR_ARG, "null slot list");
    return batch_run(b, n, sources, slots, ms);
}

extern "C" int mg_batch_get(mg_batch *b, int which, void *dst, int64_t bytes) {
    if (!b || !dst) return mg_fail(MG_ERR_ARG, "null argument");
    const int nv = b->nv, ns = b->nsrc;
    hipStream_t s = b->stream;
    if (which == MG_RES_LEVEL || which == MG_RES_SIGMA) {
        const bool sig = which == MG_RES_SIGMA;
        if (sig && b->kind != MG_BATCH_BC) return mg_fail(MG_ERR_ARG, "sigma needs a BC batch");
        const size_t esz = sig ? sizeof(double) : sizeof(int);
        const int64_t need = (int64_t)esz * nv * ns;
        if (bytes < need) return mg_fail(MG_ERR_ARG, "buffer of %lld bytes, need %lld", (long long)bytes, (long long)need);
        void *tmp = nullptr;
        MG_HIP(hipMalloc(&tmp, std::max<int64_t>(need, 8)));
        const long long n = (long long)nv * ns;
        const int *perm = b->permuted ? b->perm : nullptr;
        if (sig) hipLaunchKernelGGL((k_ms_transpose<double>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, nv, ns, perm, b->sigma, (double *)tmp);
        else hipLaunchKernelGGL((k_ms_transpose<int>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, nv, ns, perm,
                                b->kind == MG_BATCH_SSSP ? b->dist : b->lev, (int *)tmp);
        hipError_t e1 = hipStreamSynchronize(s);
        hipError_t e2 = need > 0 ? hipMemcpy(dst, tmp, need, hipMemcpyDeviceToHost) : hipSuccess;
        hipFree(tmp);
        if (e1 != hipSuccess || e2 != hipSuccess) return mg_fail(MG_ERR_DEVICE, "batch result copy failed");
        return MG_OK;
    }
    if (which == MG_RES_DELTA) {
        if (b->kind != MG_BATCH_BC) return mg_fail(MG_ERR_ARG, "bc needs a BC batch");
        if (bytes < (int64_t)sizeof(float) * nv) return mg_fail(MG_ERR_ARG, "buffer too small");
        MG_HIP(hipStreamSynchronize(s));
        MG_HIP(hipMemcpy(dst, b->bc, sizeof(float) * nv, hipMemcpyDeviceToHost));
        return MG_OK;
    }
    return mg_fail(MG_ERR_ARG, "unknown batch result %d", which);
}

extern "C" int mg_batch_stats(const mg_batch *b, int64_t *out, int n) {
    if (!b || !out) return mg_fail(MG_ERR_ARG, "null argument");
    if (n < 0 || n > 5) return mg_fail(MG_ERR_ARG, "stats: 0..5 entries (asked %d)", n);
    for (int i = 0; i < n; ++i) out[i] = b->stats[i];
    return MG_OK;
}

extern "C" int mg_batch_levels(const mg_batch *b, int *levels) {
    if (!b || !levels) return mg_fail(MG_ERR_ARG, "null argument");
    *levels = b->levels;
    return MG_OK;
}
