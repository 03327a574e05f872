/*
 * multigraph_amd.h — C ABI of the MI355X (gfx950) frontier-processing engine.
 *
 * Every entry point returns an int status (MG_OK == 0, negative on error) and
 * leaves a message for mg_last_error().  No torch or HIP types cross this
 * boundary: graphs and engines are opaque handles, inputs and outputs are plain
 * host pointers and sizes.  Vertex ids on this boundary are the reader's
 * compacted ids (the i-th non-isolated vertex of the input, ascending original
 * id), which is the id space the reference validates in
 * (code/read_input.cu:115-139, code/validate.cu:310-316).
 *
 * Reference interface replaced by each group of entry points:
 *   graph input ............ generate_CSC            code/read_input.cu:11-254, code/common.h:215
 *   layout build ........... generate_MultiGraph(S)  code/gen_structure.cu:2230-2947, :99-139
 *   release ................ release_MultiGraph      code/release_structure.cu:4-19
 *   init / reset ........... initValue / resetValue  code/user_function.h:3-208
 *   one engine step ........ MultiGraph_V_E_V, _NOTIDEM, _E_E, _V_V  code/common.h:228-233
 *   BC backward ............ MultiGraph_V_E_V_pull   code/process2.cu:263-325
 *   results ................ struct OUTPUT r/r2/r3   code/common.h:205-213
 */
#ifndef MULTIGRAPH_AMD_H
#define MULTIGRAPH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define MG_OK            0
#define MG_ERR_ARG      -1   /* bad argument (null pointer, id out of range, ...) */
#define MG_ERR_IO       -2   /* file missing or malformed                         */
#define MG_ERR_RANGE    -3   /* vertex id in the input outside [1, nv]            */
#define MG_ERR_NOMEM    -4   /* host or device allocation failed                  */
#define MG_ERR_DEVICE   -5   /* HIP runtime error (no device, launch failure)     */
#define MG_ERR_LIMIT    -6   /* a size limit of this build was exceeded           */
#define MG_ERR_STATE    -7   /* call out of order (e.g. step before reset)        */

/* ---- algorithms (the reference's -D{BFS,SSSP,BC,CC,PR_T,PR_D}, attribute.h) */
#define MG_BFS   0
#define MG_SSSP  1
#define MG_BC    2
#define MG_CC    3
#define MG_PR_T  4   /* DATA_DRIVEN_PR     (install.sh:23) */
#define MG_PR_D  5   /* TOPOLOGY_DRIVEN_PR (install.sh:24) */

/* ---- reader flags (compile-time -DSYM / -DNSYM / -DCC of the reference) ---- */
#define MG_READ_SYM      1   /* add (v,u,w) for every (u,v,w): read_input.cu:88-95 (-DSYM) */
#define MG_READ_CC       2   /* keep only src < dst edges:     read_input.cu:111-113     */
#define MG_READ_HEADER   4   /* honour the MTX header's symmetric flag instead of forcing */

/* ---- engine modes (the reference's -DSPARSE_MODE / -DDENSE_MODE / hybrid) - */
#define MG_MODE_HYBRID   0
#define MG_MODE_SPARSE   1   /* queue push only                    */
#define MG_MODE_DENSE    2   /* bitmap pull / partitioned sweeps   */

/* ---- result selectors (struct OUTPUT r, r2, r3) ------------------------------ */
#define MG_RES_LEVEL     0   /* int32: BFS/BC level, SSSP distance, CC label (min id in component) */
#define MG_RES_SIGMA     1   /* double: BC shortest-path counts                                  */
#define MG_RES_DELTA     2   /* float : BC dependency                                            */
#define MG_RES_RANK      3   /* float : PageRank                                                 */

#define MG_INF 0x64646464    /* unreached level / distance (common.h:37-39,55-56) */

typedef struct mg_graph  mg_graph;    /* host CSR after reading   (struct csc_package) */
typedef struct mg_engine mg_engine;   /* device layout + run state (struct HYB + vector_data) */

/* Tunables of one engine.  Zero fields take the defaults; the autotuner fills them. */
typedef struct mg_options {
    int    mode;            /* MG_MODE_*                                                */
    float  alpha;           /* push->pull when frontier edges > unexplored edges / alpha
                               (0, the default: 30 for engines, 15 for batched runs)   */
    float  beta;            /* pull->push when frontier vertices < nv / beta; batched
                               BFS / SSSP push levels walk a queue of the frontier when
                               its vertices + edges < (nv + E) / beta (0, the default:
                               36 for engines, 18 for batched runs)                     */
    int    huge_degree;     /* vertices at or above this degree are split across blocks */
    int    grid_blocks;     /* workgroups of the grid-stride frontier kernels (0, the
                               default: 2048, BC and CC engines 4096; batched runs:
                               4096, also their maximum)                                */
    int    graph_levels;    /* levels / iterations captured per hipGraph launch (0: BFS,
                               SSSP, BC step eagerly; PageRank replays 8 per host check) */
    float  pr_damping;      /* PageRank damping (0.85, common.h:59)                     */
    float  pr_tolerance;    /* PageRank tolerance (0.01, common.h:60)                   */
    int    pr_max_iter;     /* PageRank iteration cap (0: unbounded, as the reference)  */
    int    part_log2;       /* log2 of the vertex-partition size of the tiled layout    */
    int    batch_chunk;     /* batched runs: edges per static chunk, at least 64 (0:
                               1024 for BFS, 128 for BC, 64 for SSSP)                   */
    int    persist_levels;  /* batched BFS / SSSP: run sparse levels in a persistent
                               cooperative launch, at most this many per launch (-1:
                               unbounded; 0, the default: every level from the captured
                               graphs, which measures faster on MI355X)                 */
    int    pull_layout;     /* SSSP dense sweeps: 0, the default, over the partitioned
                               layout (distance messages, 16-bit weights; symmetric graphs
                               with weights in [0, 65535]) with 16-bit messages until a
                               distance above 65534 must be sent, then 32-bit; 1 over the
                               CSR; 2 over the layout with 32-bit messages throughout   */
    int    sssp_delta;      /* batched SSSP: sparse sweeps relax only the (vertex, source)
                               pairs below a distance threshold that rises by this much
                               per level, the others wait (0: a quarter of the mean edge
                               weight; -1: every pair, plain Bellman-Ford)              */
    int    batch_lanes;     /* batched runs without a slot list: 0, the default, puts the
                               sources in lanes ordered by their 2-hop reach (sources whose
                               levels run alike share cache lines of the per-vertex rows;
                               results come back in the caller's order); 1 keeps source i
                               in lane i                                                */
    /* Bin edges and block shape of the frontier kernels (0: the default in brackets):   */
    int    lane_edges;      /* BC pull and backward: rows of at most this many edges are
                               walked by one lane, longer ones by a wave [32]            */
    int    pull_split;      /* BFS / BC pull: rows above this many in-edges are cut into
                               2048-edge chunks, each a wave's [2048]                    */
    int    back_split;      /* BC backward: rows of at least this many edges are cut into
                               back_chunk-edge chunks, each a workgroup's [2048]         */
    int    back_chunk;      /* BC backward: edges per chunk [1024]                       */
    int    block_threads;   /* workgroup size of the pull, edge-balanced push and BC
                               backward kernels: 256, 512 or 1024 [256]                 */
    int    batch_light;     /* batched BFS pull: vertices of at most this many in-edges
                               are walked by one lane, heavier ones by a wave [32]      */
    int    sparse_grid;     /* batched BFS / SSSP: workgroups of a sparse (queue) level
                               (0: one per CU)                                          */
    int    wide_offsets;    /* 1: the engine keeps 64-bit row offsets on the device even
                               when the graph's fit int32 (a wide graph always has them;
                               BFS and SSSP run on such engines, symmetric graphs)      */
} mg_options;

/* ---- errors ---------------------------------------------------------------- */
const char *mg_last_error(void);
const char *mg_version(void);

/* ---- graph input (host) ---------------------------------------------------- */
/* generate_CSC (read_input.cu:11-254): read a Matrix Market file, symmetrize
 * (flags & MG_READ_SYM), sort, drop duplicates, compact isolated vertices, build
 * CSR and remap `source` into the compacted id space.  Pattern files get weights
 * 1 + splitmix64(weight_seed, line) % 64 (the reference uses time-seeded rand()). */
int mg_graph_load_mtx(const char *path, int source, int flags, uint64_t weight_seed,
                      mg_graph **out);
/* Same semantics for an in-memory list of 0-based entries (one per MTX line). */
int mg_graph_from_coo(int nv, int64_t nnz, const int32_t *src, const int32_t *dst,
                      const int32_t *val, int source, int flags, mg_graph **out);
/* A CSR built by the caller (row[nv + 1] offsets, col, optional weights, 1 when NULL):
 * ids are kept as given (no compaction), every row is sorted by column with a repeated
 * column keeping its first entry, and the graph is symmetric when every stored edge
 * has its reverse with the same weight.  The reference has no such entry: its graphs come from
 * generate_CSC only (read_input.cu:11-254); this is the CSR it builds (:151-171). */
int mg_graph_from_csr(int nv, const int64_t *row, const int32_t *col, const int32_t *weight,
                      mg_graph **out);
/* RMAT (a,b,c, d = 1-a-b-c), 2^scale vertices, edge_factor * 2^scale draws, self-loops
 * dropped, weights 1 + hash(weight_seed, min(u,v), max(u,v)) % 64; then read as above. */
int mg_graph_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                      uint64_t seed, uint64_t weight_seed, int flags, mg_graph **out);
/* 2-D 4-neighbour grid, width x height, id = y*width + x (edges +x and +y, both ways
 * with MG_READ_SYM).  Generators honour MG_READ_SYM and MG_READ_CC as the reader does. */
int mg_graph_gen_grid(int width, int height, uint64_t weight_seed, int flags, mg_graph **out);
/* Disjoint union of graphs (vertex ids offset in order); used for uneven batches. */
int mg_graph_union(int count, const mg_graph *const *graphs, mg_graph **out);
int mg_graph_write_mtx(const mg_graph *g, const char *path);
int mg_graph_info(const mg_graph *g, int32_t *nv, int64_t *ne, int32_t *start_p);
/* 1 when every stored edge has its reverse with the same weight (the reader's -DSYM
 * graphs): rows double as in-edges; else the engines also keep the transposed CSR for
 * pull levels.  No reference counterpart (its -DNSYM path is unfinished,
 * read_input.cu:42-44). */
int mg_graph_symmetric(const mg_graph *g, int *symmetric);
/* The CSR's arrays (the reference's csc_v / csc_e / csc_ev).  Row offsets are int32
 * as csc_v is while the graph has at most 2^31 - 1 edges; past that (a "wide" graph,
 * scale 26-27 RMAT) the row request fails with MG_ERR_LIMIT and mg_graph_csr64 copies
 * the 64-bit offsets (nv + 1 of them) into `row`.  The reference has no such graphs
 * (int csc_v, read_input.cu:135-171). */
int mg_graph_csr(const mg_graph *g, const int32_t **row, const int32_t **col,
                 const int32_t **weight);
int mg_graph_csr64(const mg_graph *g, int64_t *row);
/* Original (pre-compaction) id of every compacted vertex, length nv. */
int mg_graph_orig_ids(const mg_graph *g, const int32_t **orig);
/* Map an original vertex id to the compacted id exactly as read_input.cu:130 does. */
int mg_graph_remap(const mg_graph *g, int orig_id, int32_t *compact_id);
int mg_graph_max_degree_vertex(const mg_graph *g, int32_t *v);
void mg_graph_free(mg_graph *g);

/* ---- device engine --------------------------------------------------------- */
void mg_default_options(mg_options *opt);
int mg_device_count(int *n);
/* generate_MultiGraph / generate_MultiGraphS: upload the CSR, build the
 * algorithm's device layout on the GPU, allocate run state (initValue). */
int mg_build(const mg_graph *g, int algo, const mg_options *opt, mg_engine **out);
int mg_release(mg_engine *e);
/* resetValue: seed the run from `source` (compacted id; ignored for CC and PR). */
int mg_reset(mg_engine *e, int source);
/* One engine call (MultiGraph_V_E_V / _NOTIDEM / _E_E / _V_V): returns 1 when it
 * must be called again, 0 when the run is finished, < 0 on error.  *iter is the
 * level / iteration number and is advanced by the callee, as in the reference. */
int mg_step(mg_engine *e, int *iter);
/* Mode of the level the last mg_step ran: 0 dense (pull / all-edge pass), 1 sparse
 * (push over the frontier queue), -1 none; the first field of the reference's TRACE
 * lines ("0 %d %f" / "1 %d %f", process.cu:1300-1305,1349-1353). */
int mg_step_mode(const mg_engine *e, int *mode);
/* BC dependency pass over the logged levels (MultiGraph_V_E_V_pull). */
int mg_bc_backward(mg_engine *e, int iter);
/* reset + all steps (+ backward for BC), timed with HIP events on the engine's
 * stream; uses captured hipGraphs when opt.graph_levels > 0 (always for PageRank on the
 * partitioned layout, unless profiling is on). */
int mg_run(mg_engine *e, int source, float *ms);
int mg_sync(mg_engine *e);
int mg_iterations(const mg_engine *e, int *iters);
/* Copy a result (MG_RES_*) into host memory, indexed by compacted id. */
int mg_get_result(mg_engine *e, int which, void *host_dst, int64_t bytes);

/* ---- batched multi-source runs (one bit per source packed per vertex) ------ */
#define MG_BATCH_BFS  0
#define MG_BATCH_BC   1
#define MG_BATCH_SSSP 2
typedef struct mg_batch mg_batch;
int mg_batch_build(const mg_graph *g, int kind, int max_sources, const mg_options *opt,
                   mg_batch **out);
int mg_batch_run(mg_batch *b, int nsrc, const int32_t *sources, float *ms);
/* Sources with explicit slots (bits): slot[i] in [0, max_sources).  Sources that share a
 * slot run from their set; on a disjoint union (mg_graph_union) sources of different
 * member graphs can share slots, so 32 graphs x 4 sources fill 4 slots of one run. */
int mg_batch_run_slots(mg_batch *b, int n, const int32_t *sources, const int32_t *slots, float *ms);
/* levels / distances: int32 [nslots][nv] (slot-major); sigma: double [nslots][nv];
 * bc (MG_RES_DELTA): float [nv], dependencies summed over slots, each vertex's own slots left out */
int mg_batch_get(mg_batch *b, int which, void *host_dst, int64_t bytes);
int mg_batch_release(mg_batch *b);
/* levels swept by the last mg_batch_run (BFS/BC: deepest level + 1; SSSP: sweeps) */
int mg_batch_levels(const mg_batch *b, int *levels);
/* How the last BFS / SSSP mg_batch_run spent its levels: out[0] levels, out[1] dense
 * levels (every chunk), out[2] sparse levels (a queue of the frontier), out[3] of
 * those run by persistent launches, out[4] host round trips (control-block reads).
 * n entries are written (at most 5).  No counterpart in the reference: its level
 * loop prints a per-iteration trace instead (code/process.cu:1300-1305). */
int mg_batch_stats(const mg_batch *b, int64_t *out, int n);

/* ---- autotuner ------------------------------------------------------------- */
/* Pick the tunables of mg_options (push/pull thresholds alpha and beta, grid shape,
 * levels per captured graph, PageRank partition size, batched chunk size) for one
 * graph and algorithm by timing candidates on the graph itself, one coordinate at a
 * time (each candidate: a build, a warm-up run, the median of `reps` timed runs).
 * Replaces the reference's per-run wall-clock mode sampling (process.cu:1161-1255)
 * with a one-time search.  *best starts from *best's own values (zero fields: the
 * defaults); ms_default / ms_tuned receive the timed cost before and after. */
int mg_autotune(const mg_graph *g, int algo, int nsrc, const int32_t *sources, int reps,
                mg_options *best, float *ms_default, float *ms_tuned);
/* Same for a batch kind over one run of `sources` (with `slots`, or NULL). */
int mg_batch_autotune(const mg_graph *g, int kind, int n, const int32_t *sources,
                      const int32_t *slots, int reps, mg_options *best, float *ms_default,
                      float *ms_tuned);

/* ---- allocation-free entry point ------------------------------------------- */
/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The BFS
 * of mg_build + mg_run (same kernels, same levels) over a CSR the caller already holds on
 * the current device: row[nv + 1] int32 offsets, col[ne] int32, symmetric unless mode is
 * MG_MODE_SPARSE (pull levels read rows as in-edges); levels land in the caller's device
 * array level[nv] (MG_INF = unreached).  The queue, bitmaps, degree-scan tables, partial
 * counters and control block live in one fixed device array of the library (about 40 MB,
 * held by every GPU the library runs on until the process ends), so the call allocates
 * nothing; nv above 4,194,304 or ne above 2^31 - 1 is refused with MG_ERR_LIMIT.  `stream`
 * is a hipStream_t (NULL: the null stream); the call returns when the levels are written.
 * Calls that overlap on one GPU share that array and race: serialise them. */
int mg_bfs_noalloc(int nv, int64_t ne, const int32_t *row, const int32_t *col, int source, int mode,
                   int32_t *level, void *stream);

/* ---- instrumentation -------------------------------------------------------- */
/* Per-kernel accounting of the last mg_run: launches, summed device time (ms,
 * from HIP events around every launch on the engine stream when profiling is on)
 * and algorithmic bytes.  name is the kernel's function name. */
int mg_set_profiling(mg_engine *e, int on);
/* Profile only the kernels whose name starts with prefix (NULL or "": all), so a timed
 * run pays events around those launches alone.  No reference counterpart (the
 * reference times whole iterations, process.cu:1300-1305). */
int mg_set_profile_filter(mg_engine *e, const char *prefix);
int mg_kernel_stats(mg_engine *e, int idx, char *name, int name_len, int64_t *launches,
                    double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* MULTIGRAPH_AMD_H */
