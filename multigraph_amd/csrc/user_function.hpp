// Algorithm plug-in surface: the role of code/user_function.h and code/attribute.h.
//
// An algorithm is a struct carrying its type bindings (attribute.h: V1T, V2T, E1T)
// and the reference's hooks with the same names and meaning:
//   initialize              identity of the reduction            user_function.h:211-230
//   get_partial_result      message along one edge               user_function.h:232-262
//   accumulate              atomic-safe combine                   user_function.h:264-291
//   accumulate_nonAtomics   thread-private combine                user_function.h:294-330
//   update_condition        apply at a vertex, "did it change?"   user_function.h:332-384
//   fused_update_condition  push-side apply + enqueue predicate   user_function.h:386-447
// `attribute` is the iteration number and `frontier` the per-vertex stamp (vfc),
// exactly as in the reference.  The engine's HIP kernels are templates over these
// structs and call the hooks at the reference's points, so editing a hook changes
// every path that runs the algorithm:
//   push levels (k_push, k_push_lb)        get_partial_result + fused_update_condition
//   pull levels (k_pull + its chunks,      initialize (register / slot identity),
//     k_pull_sssp, k_ps_gather,            get_partial_result, accumulate_nonAtomics
//     k_pb_gather, k_pr_gather)            (register), accumulate (LDS / global slot),
//                                          update_condition (apply), fused_update_
//                                          condition where two workgroups share a vertex
//   batched SSSP / BC forward (k_mb_*)     the same hooks per lane (lane = source)
// Where several lanes' register accumulators are merged, the merge is
// accumulate_nonAtomics of one accumulator into another, so the hooks must form a
// commutative, associative combine (all of the reference's do).
//
// Which struct serves which algorithm is chosen below by MG_BFS_HOOKS, MG_SSSP_HOOKS,
// MG_BC_HOOKS, MG_PR_HOOKS and MG_CC_HOOKS; a maintainer's own hooks file, passed at build time as
// -DMG_USER_FUNCTION='"file.hpp"' (make USER_FUNCTION=file.hpp), is included after the
// defaults and may derive from them and re-point those names.
#pragma once
#include "mg_device.hpp"

namespace mg {

// ---------------------------------------------------------------- BFS (attribute.h:1-7)
struct BFSHooks {
    using V1T = int;
    static constexpr bool kV2 = false, kE1 = false;
    // every message of one level carries the same value, so the first one a pull level
    // finds for a vertex is final and the rest of its parents need not be scanned
    static constexpr bool kFirstWins = true;
    __device__ static void initialize(V1T *loc) { *loc = MG_INF; }
    __device__ static void get_partial_result(V1T *loc, V1T data) { *loc = data + 1; }
    __device__ static void accumulate(V1T *loc, V1T data, int attribute) {
        if (data <= attribute && data < *loc) atomicMin(loc, data);
    }
    __device__ static void accumulate_nonAtomics(V1T *loc, V1T data, int attribute) {
        if (data <= attribute && data < *loc) *loc = data;
    }
    __device__ static int update_condition(V1T *old_loc, V1T *new_loc, V1T data, int, int *) {
        if (data < *old_loc) { *new_loc = data; return 1; }
        return 0;
    }
    // The reference writes without an atomic and tolerates duplicate queue entries
    // (user_function.h:392-398); here the write is an atomicMin whose old value tells
    // exactly one edge that it claimed the vertex, so every frontier is duplicate-free.
    __device__ static int fused_update_condition(V1T *old_loc, V1T *new_loc, V1T data, int attribute, int *) {
        if (data <= attribute && data < *old_loc) return atomicMin(new_loc, data) > data;
        return 0;
    }
};

// --------------------------------------------------------------- SSSP (attribute.h:9-17)
struct SSSPHooks {
    using V1T = int;
    using E1T = int;
    static constexpr bool kV2 = false, kE1 = true;
    static constexpr bool kFirstWins = false;
    __device__ static void initialize(V1T *loc) { *loc = MG_INF; }
    __device__ static void get_partial_result(V1T *loc, V1T data, E1T weight) { *loc = data + weight; }
    __device__ static void accumulate(V1T *loc, V1T data, int) { atomicMin(loc, data); }
    __device__ static void accumulate_nonAtomics(V1T *loc, V1T data, int) { if (data < *loc) *loc = data; }
    __device__ static int update_condition(V1T *old_loc, V1T *new_loc, V1T data, int attribute, int *frontier) {
        if (data < *old_loc) { *new_loc = data; *frontier = attribute; return 1; }
        return 0;
    }
    // user_function.h:400-407: lower the distance, enqueue once per iteration via the stamp
    __device__ static int fused_update_condition(V1T *old_loc, V1T *new_loc, V1T data, int attribute, int *frontier) {
        if (data < *old_loc) {
            atomicMin(new_loc, data);
            if (atomicMax(frontier, attribute) < attribute) return 1;
        }
        return 0;
    }
};

// ------------------------------------------------------ BC forward (attribute.h:19-29)
// V2 = shortest-path count sigma.  The reference binds it to int (attribute.h:23-24)
// and overflows on large graphs; here it is FP64, exact for counts below 2^53, so
// the atomic sums are order-independent and bit-reproducible.
struct BCHooks {
    using V1T = int;
    using V2T = double;
    using V3T = float;
    static constexpr bool kV2 = true, kE1 = false;
    static constexpr bool kFirstWins = false;  // every parent's path count is summed
    __device__ static void initialize(V1T *loc, V2T *loc2) { *loc = MG_INF; *loc2 = 0.0; }
    __device__ static void get_partial_result(V1T *loc, V1T data, V2T *loc2, V2T data2) {
        *loc = data + 1; *loc2 = data2;
    }
    __device__ static void accumulate(V1T *loc, V1T data, int attribute, V2T *loc2, V2T data2) {
        if (data <= attribute && data < *loc) atomicMin(loc, data);
        if (data == attribute && (*loc == MG_INF || *loc == attribute)) atomicAdd(loc2, data2);
    }
    __device__ static void accumulate_nonAtomics(V1T *loc, V1T data, int attribute, V2T *loc2, V2T data2) {
        if (data <= attribute && data < *loc) *loc = data;
        if (data == attribute && (*loc == MG_INF || *loc == attribute)) *loc2 += data2;
    }
    // user_function.h:365-373 stores into old_loc with old_loc == new_loc at its call
    // site; storing into new_loc lets the engine pass the old value in a register
    __device__ static int update_condition(V1T *old_loc, V1T *new_loc, V1T data, int, int *, V2T *, V2T *new_loc2, V2T data2) {
        if (data < *old_loc) { *new_loc = data; *new_loc2 = data2; return 1; }
        return 0;
    }
    // user_function.h:409-417
    __device__ static int fused_update_condition(V1T *old_loc, V1T *new_loc, V1T, int attribute, int *,
                                                 V2T *, V2T *new_loc2, V2T data2) {
        int o = *old_loc;
        if (o == MG_INF || o == attribute) {
            atomicAdd(new_loc2, data2);
            if (atomicCAS(new_loc, (int)MG_INF, attribute) == (int)MG_INF) return 1;
        }
        return 0;
    }
    // Backward (dependency) pass.  The reference has no hooks here: process2.cu:135-259
    // and validate.cu:170-185 inline delta[v] = sum over successors w of
    // sigma[v] / sigma[w] * (1 + delta[w]).  The same rule as hooks: a finished vertex w
    // publishes c[w] = (1 + delta[w]) / sigma[w] (back_message), an edge v -> w adds
    // sigma[v] * c[w] (back_term) into a compensated FP32 sum (back_accumulate: the
    // term split into two floats, Neumaier's sum), and the vertex applies it
    // (back_update: delta, then its own c for the level above).
    // c is stored in FP32 (CT): one gather per successor edge of half the bytes; its
    // relative rounding (6e-8) bounds each term's, far inside BC's 1e-6 tolerance
    using CT = float;
    __device__ static void back_message(CT *c, V3T delta, V2T sigma) { *c = (CT)((1.0 + (double)delta) / sigma); }
    __device__ static void back_term(double *t, V2T sigma_v, double c_w) { *t = sigma_v * c_w; }
    __device__ static void back_accumulate(KSum *acc, double t) {
        const float hi = (float)t;
        acc->add(hi);
        acc->add((float)(t - (double)hi));
    }
    __device__ static void back_update(V3T *delta, CT *c, V3T dv, V2T sigma) {
        *delta = dv;
        back_message(c, dv, sigma);
    }
};

// ------------------------------------------------------------------ CC (attribute.h:31-39)
// The reference hooks along active edges by iteration parity and pointer-jumps
// (user_function.h:374-380, :418-436).  Here hooking is an asynchronous union of
// roots that always hangs the larger root under the smaller one, so parent[v] <= v
// holds throughout and the final root of every component is its minimum id.
struct CCHooks {
    using V1T = int;
    // root of v with path halving: every visited vertex is re-pointed at its
    // grandparent.  Concurrent writers only ever store an ancestor, and roots are
    // changed only by the hook's CAS, so the halving writes need no atomics.
    __device__ static int find(V1T *parent, int v) {
        int p = parent[v];
        while (p != v) {
            const int gp = parent[p];
            if (gp != p) parent[v] = gp;
            v = p;
            p = gp;
        }
        return v;
    }
    // "hook": returns 1 when this call merged two trees
    __device__ static int fused_update_condition(V1T *parent, int u, int v) {
        int ru = find(parent, u), rv = find(parent, v);
        for (;;) {
            if (ru == rv) return 0;
            if (ru < rv) { int t = ru; ru = rv; rv = t; }  // ru is the larger root
            int old = atomicCAS(&parent[ru], ru, rv);
            if (old == ru) return 1;
            ru = find(parent, old);
            rv = find(parent, rv);
        }
    }
    // root of v without writing: used by the final compression, where a halving
    // write from one thread could re-point another thread's vertex, which already
    // holds its root, back at an intermediate ancestor
    __device__ static int find_ro(const V1T *parent, int v) {
        int p = parent[v];
        while (p != v) { v = p; p = parent[v]; }
        return v;
    }
    // pointer jumping: parent[v] = root(v), written only by v's own thread;
    // returns 1 when it changed
    __device__ static int update_condition(V1T *parent, int v) {
        int p = parent[v];
        if (p == v) return 0;
        int r = find_ro(parent, p);
        if (r != p) { parent[v] = r; return 1; }
        return 0;
    }
};

// ---------------------------------------------------------- PageRank (attribute.h:41-57)
// V1T rank, V2T degree.  Messages are FP32 r/deg (the division the reference does in
// _phase1, processi.cu:410-461); sums are accumulated in FP64 and the rank is rounded
// to FP32 once per iteration.
struct PRHooks {
    using V1T = float;
    using V2T = int;
    __device__ static void initialize(double *loc) { *loc = 0.0; }
    __device__ static void get_partial_result(float *loc, float rank, int deg) { *loc = deg > 0 ? rank / (float)deg : 0.f; }
    __device__ static void accumulate(double *loc, double data) { atomicAdd(loc, data); }
    __device__ static void accumulate_nonAtomics(double *loc, double data) { *loc += data; }
    // PR_T / PR_D rules of processi.cu:636-654 (variant 0 / 1); returns 1 when the
    // vertex counts as changed.  f1 is computed from the FP64 sum.
    __device__ static int update_condition(float *rank, int *active, double sum, int variant, float damping, float tol,
                                           float *f1_out) {
        float f1 = (float)((1.0 - (double)damping) + (double)damping * sum);
        float f2 = *rank;
        *f1_out = f1;
        if (variant == 0) {
            if (fabsf(f1 - f2) > tol) { *rank = f1; return 1; }
            return 0;
        }
        if (*active == 1 && fabsf(f1 - f2) > tol * f2) { *rank = f1; return 1; }
        *rank = f1;
        *active = 0;
        return 0;
    }
};

}  // namespace mg

#ifdef MG_USER_FUNCTION
#include MG_USER_FUNCTION
#endif
#ifndef MG_BFS_HOOKS
#define MG_BFS_HOOKS mg::BFSHooks
#endif
#ifndef MG_SSSP_HOOKS
#define MG_SSSP_HOOKS mg::SSSPHooks
#endif
#ifndef MG_BC_HOOKS
#define MG_BC_HOOKS mg::BCHooks
#endif
#ifndef MG_PR_HOOKS
#define MG_PR_HOOKS mg::PRHooks
#endif
#ifndef MG_CC_HOOKS
#define MG_CC_HOOKS mg::CCHooks
#endif
