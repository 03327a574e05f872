// Host-side mirror of the reference's engine interface over the C ABI.
//
// The names, argument order and return conventions follow code/common.h:215-235 so a
// driver reads like code/main.cu: build once, then per repetition `resetValue` and
// `while (MultiGraph_V_E_V(...));`, then `verifyResults`.  The algorithm is chosen at
// compile time with the reference's macros (-DBFS, -DSSSP, -DBC, -DCC, -DPR_T, -DPR_D,
// attribute.h:1-57); everything device-side is behind include/multigraph_amd.h.
//
// Errors: the reference prints and exits on a bad vertex id (read_input.cu:73-76)
// and leaves other failures unchecked; here every failure prints the library's
// message to stderr and exits with status 1 (status 0 for the out-of-range id, as
// the reference does).
#pragma once
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "multigraph_amd.h"

#if defined(PR_T) || defined(PR_D)
typedef float V1T;
#else
typedef int V1T;
#endif
typedef double V2T;  // BC path counts (FP64; the reference's int overflows)
typedef float V3T;   // BC dependencies

inline void mg_check(int rc, const char *what) {
    if (rc < 0) {
        fprintf(stdout, "%s\n", mg_last_error());
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, mg_last_error());
        exit(rc == MG_ERR_RANGE ? 0 : 1);
    }
}

struct csc_package {          // code/common.h:170-182
    mg_graph *g = nullptr;
    int nv = 0, ne = 0, start_p = 0, edges = 0;
    const int *csc_v = nullptr, *csc_e = nullptr, *csc_ev = nullptr;
};

struct HYB {                  // code/common.h:129-168 (the device layout lives in the engine)
    mg_engine *e = nullptr;
    int nv = 0, ne = 0, start_point = 0;
};
typedef HYB MULTI_SPARSE;     // code/common.h:123-127

struct vector_data {          // code/common.h:184-199 (run state lives in the engine)
    int source = 0;
};

struct OUTPUT {               // code/common.h:205-213, here host copies
    std::vector<V1T> r;
    std::vector<V2T> r2;
    std::vector<V3T> r3;
};

inline int mg_compat_algo() {
#if defined(BFS)
    return MG_BFS;
#elif defined(SSSP)
    return MG_SSSP;
#elif defined(BC)
    return MG_BC;
#elif defined(CC)
    return MG_CC;
#elif defined(PR_T)
    return MG_PR_T;
#elif defined(PR_D)
    return MG_PR_D;
#else
#error "define one of BFS SSSP BC CC PR_T PR_D"
#endif
}

// read_input.cu:11-254; argv[1] = file, argv[2] = source (original id)
inline void generate_CSC(struct csc_package *inp, char **argv) {
    int flags = MG_READ_SYM;
#if defined(CC)
    flags |= MG_READ_CC;
#endif
    mg_check(mg_graph_load_mtx(argv[1], atoi(argv[2]), flags, 1, &inp->g), "generate_CSC");
    int32_t nv, sp;
    int64_t ne;
    mg_graph_info(inp->g, &nv, &ne, &sp);
    inp->nv = nv; inp->ne = (int)ne; inp->start_p = sp; inp->edges = (int)ne;
    const int32_t *row, *col, *w;
    mg_graph_csr(inp->g, &row, &col, &w);
    inp->csc_v = row; inp->csc_e = col; inp->csc_ev = w;
}

inline mg_options mg_compat_options() {
    mg_options o;
    mg_default_options(&o);
#if defined(SPARSE_MODE)
    o.mode = MG_MODE_SPARSE;
#elif defined(DENSE_MODE)
    o.mode = MG_MODE_DENSE;
#endif
    if (const char *s = getenv("MG_PR_TOLERANCE")) o.pr_tolerance = (float)atof(s);
    if (const char *s = getenv("MG_PR_MAX_ITER")) o.pr_max_iter = atoi(s);
    if (const char *s = getenv("MG_GRAPH_LEVELS")) o.graph_levels = atoi(s);
    return o;
}

// gen_structure.cu:2230-2947 / :99-139
inline void generate_MultiGraph(struct HYB *m, struct csc_package *inp) {
    mg_options o = mg_compat_options();
    mg_check(mg_build(inp->g, mg_compat_algo(), &o, &m->e), "generate_MultiGraph");
    m->nv = inp->nv; m->ne = inp->ne; m->start_point = inp->start_p;
}
inline void generate_MultiGraphS(MULTI_SPARSE *m, struct csc_package *inp) { generate_MultiGraph(m, inp); }

inline void release_MultiGraph(struct HYB *m) { mg_release(m->e); m->e = nullptr; }

// user_function.h:3-208: allocation happens in generate_MultiGraph; both seed the run
inline void initValue(int source_vertex, struct HYB *, struct vector_data *vd) { vd->source = source_vertex; }
inline void resetValue(int source_vertex, struct HYB *m, struct vector_data *vd) {
    vd->source = source_vertex;
    mg_check(mg_reset(m->e, source_vertex), "resetValue");
}

inline void mg_fill_output(struct HYB *m, struct OUTPUT *out) {
    out->r.resize(m->nv);
#if defined(PR_T) || defined(PR_D)
    mg_check(mg_get_result(m->e, MG_RES_RANK, out->r.data(), (int64_t)sizeof(V1T) * m->nv), "result");
#else
    mg_check(mg_get_result(m->e, MG_RES_LEVEL, out->r.data(), (int64_t)sizeof(V1T) * m->nv), "result");
#endif
#if defined(BC)
    out->r2.resize(m->nv);
    out->r3.resize(m->nv);
    mg_check(mg_get_result(m->e, MG_RES_SIGMA, out->r2.data(), (int64_t)sizeof(V2T) * m->nv), "result");
    mg_check(mg_get_result(m->e, MG_RES_DELTA, out->r3.data(), (int64_t)sizeof(V3T) * m->nv), "result");
#endif
}

// function.cu:4-12: wall clock in seconds
inline double rtclock() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// process.cu:1079 / processi.cu:682 / process.cu:1051,1012: nonzero = call again.
// With MG_TRACE set (the reference's -DTRACE), every level prints "<mode> <level> <ms>"
// as process.cu:1300-1305,1349-1353 do: mode 0 dense, 1 sparse, the level counted from
// 0, milliseconds between device synchronisations around it.
inline int mg_compat_step(struct HYB *m, int *iter) {
    static const bool trace = getenv("MG_TRACE") != nullptr;
    const int level = iter ? *iter : 1;
    double t0 = 0.0;
    if (trace) {
        mg_check(mg_sync(m->e), "sync");
        t0 = rtclock();
    }
    int rc = mg_step(m->e, iter);
    mg_check(rc, "MultiGraph step");
    if (trace) {
        mg_check(mg_sync(m->e), "sync");
        int mode = -1;
        mg_check(mg_step_mode(m->e, &mode), "step mode");
        if (mode >= 0) printf("%d %d %f\n", mode, level - 1, (rtclock() - t0) * 1000.0);
    }
    return rc;
}
inline int MultiGraph_V_E_V(struct HYB *m, struct csc_package *, struct vector_data *, struct OUTPUT *, int *iter) {
    return mg_compat_step(m, iter);
}
inline int MultiGraph_V_E_V_NOTIDEM(struct HYB *m, struct csc_package *, struct vector_data *, struct OUTPUT *, int *iter) {
    return mg_compat_step(m, iter);
}
inline int MultiGraph_E_E(MULTI_SPARSE *m, struct vector_data *, struct OUTPUT *, int *iter) { return mg_compat_step(m, iter); }
inline int MultiGraph_V_V(MULTI_SPARSE *, struct vector_data *, struct vector_data *, struct OUTPUT *, int *) {
    return 0;  // the union-find hook step already leaves every label at its root
}
// process2.cu:263-325: called once after the forward pass, returns 0
inline int MultiGraph_V_E_V_pull(struct HYB *m, struct csc_package *, struct vector_data *, struct OUTPUT *, int *iter) {
    mg_check(mg_bc_backward(m->e, iter ? *iter : 0), "MultiGraph_V_E_V_pull");
    return 0;
}
