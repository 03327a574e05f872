// Engine internals shared by the HIP translation units: the device control block,
// the engine handle (role of struct HYB + struct vector_data, code/common.h:129-199)
// and launch/profiling helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "mg_internal.hpp"

#define MG_HIP(x)                                                                                      \
    do {                                                                                               \
        hipError_t _e = (x);                                                                           \
        if (_e != hipSuccess)                                                                          \
            return mg_fail(MG_ERR_DEVICE, "%s: %s (%s:%d)", #x, hipGetErrorString(_e), __FILE__, __LINE__); \
    } while (0)

#define MG_TRY(x)                 \
    do {                          \
        int _r = (x);             \
        if (_r != MG_OK) return _r; \
    } while (0)

namespace mg {

enum { MODE_TD = 0, MODE_BU = 1, MODE_DONE = 2 };

constexpr int kBlock = 256;          // threads per workgroup of the frontier kernels
constexpr int kHugeChunk = 4096;     // edges per chunk of a split high-degree vertex
constexpr int kTile = 1024;          // frontier entries per degree-scan tile
constexpr int kEPB = 2048;           // edges per workgroup chunk of the edge-balanced push (at most)
constexpr int kEPBMin = 256;         // ... and at least
constexpr long long kSmallMF = 16384;  // frontiers with fewer edges use the single-pass push
// Edges per chunk of the edge-balanced push for a frontier of `total` edges: enough to
// make about 2048 chunks, rounded up to a multiple of 256, within [kEPBMin, kEPB].  A
// hub's 163 K edges then spread over 640 workgroups instead of 80.
__host__ __device__ constexpr int lb_epb(long long total) {
    long long e = (total + 2047) / 2048;
    e = (e + kEPBMin - 1) / kEPBMin * kEPBMin;
    return (int)(e < kEPBMin ? kEPBMin : e > kEPB ? kEPB : e);
}
// the most chunks a frontier of at most ne edges is cut into
__host__ __device__ constexpr long long lb_max_chunks(long long ne) {
    const long long c = (ne + kEPB - 1) / kEPB;
    return c > 2048 ? c : 2048;
}
// Direction rule of the single-source engines (mg_options alpha / beta left at 0): push
// -> pull when the frontier's edges exceed the unexplored edges / kAlpha, pull -> push
// when the frontier has fewer than nv / kBeta vertices.  RMAT-22 from 8 random sources
// against 15 / 18: BFS 0.485 -> 0.347 ms, SSSP 1.685 -> 1.536 ms, BC 1.875 -> 1.794 ms;
// from the hub all within 2 %.
constexpr float kAlpha = 30.0f;
constexpr float kBeta = 36.0f;

// Device-resident control block.  One small kernel advances it between levels, so
// a sequence of levels needs no host round trip and can be captured in a hipGraph;
// in eager mode the host reads it back once per level (the reference's D2H of its
// 4 queue counters, process.cu:1341).
struct Ctl {
    int iter;        // level / iteration being produced (the hooks' `attribute`)
    int mode;        // MODE_TD, MODE_BU or MODE_DONE
    int in_begin;    // current frontier = q[in_begin, in_begin + in_n)
    int in_n;
    int out_begin;   // next frontier appended at q[out_begin + ...]
    int out_n;       // append counter of the level in flight
    int nchunks;     // split high-degree vertices of the level in flight
    int convert;     // 1: build bitmaps from levels before a pull level
    int bm_cur;      // bitmap slot holding the current frontier
    int levels;      // levels produced so far (BC log length)
    int changed;     // dense-sweep "something changed" flag
    int lb;          // 1: the push level uses the edge-balanced kernels
    int lb_total;    // edges of the current frontier (edge-balanced path)
    int lb_chunks;   // edge chunks of the current frontier
    int need_queue;  // 1: build the new frontier's queue from its bitmap (after a pull level)
    int qsrc;        // bitmap slot holding the frontier produced by the last pull level
    int msg_iter;    // SSSP over the partitioned layout: the level whose input messages
                     // the last pull sweep wrote (else they are rebuilt from distances)
    long long out_mf;  // degree sum of the appended vertices
    long long mf;      // degree sum of the current frontier
    long long mu;      // degree sum of unexplored vertices
    double resid;      // PageRank FP64 L1 residual of the last iteration
    long long edges;   // edges traversed (statistics)
    int seq;           // host mirror only: the level whose control block it holds
    int sat;           // SSSP layout sweeps with 16-bit messages: a finite distance above
                       // 65534 could not be sent (the host widens the messages)
};

struct Params {
    int nv;
    int sssp;        // 1: SSSP, where the dense level is an atomic-free pull relaxation
    int need_log;    // 1: every level's vertices are kept in the queue log (BC)
    int log_queue;   // 1: frontiers appended to one log (BFS/BC); 0: ping-pong halves (SSSP)
    int forced;      // MG_MODE_*
    int huge;        // split threshold
    long long ne;
    float alpha, beta;
};

struct HugeChunk { int v, j; };
// pull chunk of a high-degree vertex: edges [p0, p1) of v, k = v's slot in the huge list
// (OT: the type of edge offsets, int, or long long for graphs past 2^31 - 1 edges)
template <class OT>
struct PullChunkT { int v; OT p0, p1; int k; };
using PullChunk = PullChunkT<int>;
constexpr int kPullHuge = 2048;   // pull levels split vertices above this degree (default of pull_split)
constexpr int kPullChunk = 2048;  // edges per pull chunk

struct KStat {
    std::string name;
    long long launches = 0;
    double ms = 0.0, bytes = 0.0;
};

// the control block of a run from `source` of out-degree deg_src: level 1 next, the
// queue holding the source alone (written by k_seed on the device and by the host for
// its own copy, so a reset needs no read back)
__host__ __device__ inline Ctl seed_ctl(int deg_src, long long ne, int out_begin) {
    Ctl c{};
    c.iter = 1; c.mode = MODE_TD; c.in_begin = 0; c.in_n = 1; c.out_begin = out_begin; c.out_n = 0;
    c.nchunks = 0; c.convert = 0; c.bm_cur = 0; c.levels = 1; c.changed = 0; c.msg_iter = -1;
    c.lb = deg_src > kSmallMF ? 1 : 0; c.lb_total = 0; c.lb_chunks = 0; c.need_queue = 0; c.qsrc = 0;
    c.out_mf = 0; c.mf = deg_src; c.mu = ne - deg_src; c.resid = 0.0; c.edges = 0; c.seq = 0; c.sat = 0;
    if (c.lb) {  // the source's edges, already scanned (k_seed writes the scan's tables)
        c.lb_total = deg_src;
        c.lb_chunks = (deg_src + lb_epb(deg_src) - 1) / lb_epb(deg_src);
    }
    return c;
}

__host__ __device__ inline void advance(Ctl &c, const Params &p) {
    // record the level just produced (BC level log)
    const int produced = c.out_n;
    c.mu -= c.out_mf;
    if (c.mu < 0) c.mu = 0;
    if (produced == 0) { c.mode = MODE_DONE; c.in_n = 0; return; }
    const int prev = c.mode;
    int next;
    if (p.forced == MG_MODE_SPARSE) next = MODE_TD;
    else if (p.forced == MG_MODE_DENSE) next = MODE_BU;
    else if (p.sssp) next = ((double)c.out_mf > (double)p.ne / (double)p.alpha) ? MODE_BU : MODE_TD;
    else if (prev == MODE_TD) next = ((double)c.out_mf > (double)c.mu / (double)p.alpha) ? MODE_BU : MODE_TD;
    else next = ((double)produced < (double)p.nv / (double)p.beta) ? MODE_TD : MODE_BU;
    c.convert = (next == MODE_BU && prev != MODE_BU) ? 1 : 0;
    // a pull level leaves its frontier as a bitmap (slot bm_cur ^ 1, or 0 for SSSP);
    // a queue is built from it when the next level pushes or when levels are logged
    c.qsrc = p.sssp ? 0 : (c.bm_cur ^ 1);
    c.need_queue = (prev == MODE_BU && (p.need_log || next == MODE_TD)) ? 1 : 0;
    if (prev == MODE_BU && next == MODE_BU) c.bm_cur ^= 1;
    c.in_begin = c.out_begin;
    c.in_n = produced;
    if (p.log_queue) c.out_begin = c.out_begin + produced;
    else c.out_begin = (c.out_begin == 0) ? p.nv : 0;
    c.mf = c.out_mf;
    c.out_n = 0;
    c.out_mf = 0;
    c.nchunks = 0;
    c.lb = c.mf > kSmallMF ? 1 : 0;
    c.iter += 1;
    c.levels = c.iter;
    c.mode = next;
}

}  // namespace mg

// PageRank partitioned layout (the gen_structure.cu role, built on the GPU).
struct mg_pb_layout {
    int plog = 14;            // log2 partition size when forced by the options
    int P = 0;                // ids per partition
    int np = 0;               // partitions
    int G = 16;               // entries per mapper group
    long long slots = 0;      // edge slots of the destination-major stream (tiles padded to G)
    long long pairs = 0;      // message slots: one per (source vertex, destination partition)
    uint16_t *a_src = nullptr;   // per message slot, source-major order: local source id (0xFFFF = padding)
    int *amap = nullptr;         // per group of G source-major message slots: their destination-major position
    float *msg = nullptr;        // per message slot, destination-major order
    uint16_t *b_dst = nullptr;   // per edge slot: local destination id | (starts a new message) << 15
    int *gbase = nullptr;        // per group of G edge slots (+1): message index of the group's first slot
    // work units (chunks of the streams), built on the host from the partition starts
    int nA = 0, nB = 0, nApply = 0;
    int4 *unitsA = nullptr;      // {partition, first message slot, end, 0}
    int4 *unitsB = nullptr;      // {partition, first edge slot, end, split}
    int *split_parts = nullptr;  // destination partitions processed by more than one unit
    int *apart = nullptr;        // np + 1: first source-major message slot of every partition
    // SSSP sweeps over the layout
    void *b_w = nullptr;         // per edge slot: its weight (destination-major, like b_dst), uint8 or uint16
    bool w8 = false;             // b_w holds bytes (every weight <= 255)
    int *inv = nullptr;          // internal id -> compacted id (engine ids)
    int4 *unitsS = nullptr;      // every source partition {p, a0, a1, 0}: messages rebuilt from distances
    int nS = 0;
};

struct mg_engine {
    int algo = 0;
    mg_options opt{};
    int nv = 0;
    long long ne = 0;
    int source = 0;
    hipStream_t stream = nullptr;
    // CSR on the device (the renumbered CSR of HYB: _ncsc_v/_ncsc_e/_ncsc_ev)
    int *row = nullptr, *col = nullptr, *w = nullptr;
    // 64-bit row offsets instead of row on a graph past 2^31 - 1 edges (or with
    // mg_options.wide_offsets): BFS and SSSP on symmetric graphs
    long long *row64 = nullptr;
    bool wide = false;
    // out-degrees on the host (frontier engines): a reset seeds its control block from
    // the source's degree without reading the device
    std::vector<int32_t> hdeg;
    // in-edges of a directed graph (the transposed CSR, rows sorted by source, weights
    // carried along) for pull levels; null on a symmetric graph, whose rows are both
    int *trow = nullptr, *tcol = nullptr, *tw = nullptr;
    const int *in_row() const { return trow ? trow : row; }
    const int *in_col() const { return trow ? tcol : col; }
    const int *in_w() const { return trow ? tw : w; }
    // run state (struct vector_data)
    int *level = nullptr;      // BFS/BC level, SSSP distance, CC label (vv)
    int *vfc = nullptr;        // SSSP enqueue stamp / PR active flag
    double *sigma = nullptr;   // BC path counts
    float *delta = nullptr;    // BC dependencies
    double *dacc = nullptr;    // BC split-vertex partial sums
    uint8_t *lev8 = nullptr;   // BC backward: levels as bytes
    float *bc_cw = nullptr;    // BC backward: (1 + delta[w]) / sigma[w] of finished vertices (FP32)
    int *q = nullptr;          // frontier queue / level log
    long long qcap = 0;
    uint64_t *bm[2] = {nullptr, nullptr};
    void *pull_ch = nullptr;   // pull chunks of high-degree vertices (BFS, BC): PullChunkT<int> or <long long>
    int *pull_hv = nullptr, *pull_found = nullptr;
    double *pull_sum = nullptr;
    int pull_nch = 0, pull_nh = 0;
    uint64_t *vis = nullptr;
    mg::HugeChunk *chunks = nullptr;
    long long chunk_cap = 0;
    int *eoff = nullptr;       // frontier degree scan (tile-local exclusive offsets)
    int *tsum = nullptr;       // per-tile degree sums, scanned in place
    int *cstart = nullptr;     // first frontier entry of every edge chunk
    int *pcstart = nullptr;    // first row of every chunk of the whole CSR (SSSP pull)
    int pnch = 0;
    int *part_n = nullptr;     // per-workgroup partial counts of a pull level
    long long *part_mf = nullptr;
    int *lvl_off = nullptr;    // BC level log offsets (levels + 2)
    long long *lvl_mf = nullptr;  // BC: degree sum of every level
    int *aux = nullptr;        // small device counters (BC split-vertex chunk count)
    mg::Ctl *ctl = nullptr;
    mg::Ctl *h_ctl = nullptr;  // pinned mirror
    bool mirror = true;        // k_ctl writes h_ctl itself (false: h_ctl is pageable, copied back)
    int ctl_seq = 0;           // levels whose control block k_ctl mirrored (eager steps)
    // SSSP layout sweeps: 16-bit distance messages while every finite distance sent
    // fits (pb16 for the current run, pb16_ok when the engine allows them), 32-bit
    // otherwise; pb_force makes the next sweep rebuild every message
    bool pb16 = false, pb16_ok = false, pb_force = false;
    mg::Params P{};
    // CC: upper-triangular CSR
    int *urow = nullptr, *ucol = nullptr;
    long long ne_u = 0;
    bool cc_afforest = false;
    // PageRank
    float *rank = nullptr, *cval = nullptr;
    int *deg = nullptr;
    double *psum = nullptr;
    mg_pb_layout pb;
    bool pr_pb = false;               // PageRank over the partitioned layout
    bool sssp_pb = false;             // SSSP pull sweeps over the partitioned layout
    long long pr_nvi = 0;             // internal (renumbered, padded) vertex count
    std::vector<int> pr_perm_h;       // compacted id -> internal id
    // host-side progress
    int last_mode = -1;        // mode of the level the last mg_step ran: 0 dense (pull), 1 sparse (push)
    int iter = 1;
    bool ready = false;
    bool done = false;
    int levels = 0;
    long long edges_traversed = 0;
    // hipGraph of `graph_levels` levels
    hipGraphExec_t gexec = nullptr;
    int gexec_levels = 0;
    // profiling
    bool prof = false;
    std::string prof_only;     // profiling: only kernels whose name starts with this (empty: all)
    std::vector<mg::KStat> stats;
    std::vector<hipEvent_t> ev_pool;
    struct Pending { int stat; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
};

// instrumentation helpers (engine.hip)
int mg_prof_begin(mg_engine *e, const char *name, double bytes, int *slot);
int mg_prof_end(mg_engine *e, int slot);
int mg_prof_flush(mg_engine *e);

// per-algorithm entry points
int mg_frontier_build(mg_engine *e, const mg_graph *g);
int mg_frontier_reset(mg_engine *e, int source);
int mg_frontier_level(mg_engine *e);        // one eager level; updates e->done
int mg_frontier_capture(mg_engine *e, int levels, hipGraphExec_t *out);
int mg_bc_backward_impl(mg_engine *e);
int mg_cc_build(mg_engine *e, const mg_graph *g);
int mg_cc_reset(mg_engine *e);
int mg_cc_step(mg_engine *e, int *more);
int mg_pr_build(mg_engine *e, const mg_graph *g);
int mg_pr_reset(mg_engine *e);
int mg_pr_step(mg_engine *e, int *more);
int mg_pr_capture(mg_engine *e, int iters, hipGraphExec_t *out);
int mg_pb_reset(mg_engine *e);
int mg_pb_build(mg_engine *e, const mg_graph *g);
int mg_ps_sweep(mg_engine *e);
int mg_pb_build_alloc(mg_engine *e, const mg_graph *g);
int mg_pb_iteration(mg_engine *e);
void mg_free_dev(void *p);
