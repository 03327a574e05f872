// Adapters between the engine's kernels and an algorithm's hooks (user_function.hpp).
// They call the hooks with the arguments the algorithm's type bindings ask for: a
// second vertex value (V2: BC's path count) and an edge weight (E1: SSSP), so one
// kernel template serves every algorithm.  Acc is one vertex's accumulator of a pull
// level or sweep (the V1 value and, with V2, the second value).
#pragma once
#include "user_function.hpp"

namespace mg {

template <class H>
struct Acc {
    typename H::V1T v;
    double s;
};

// identity (initialize)
template <class H>
__device__ __forceinline__ void acc_init(Acc<H> &a) {
    a.s = 0.0;
    if constexpr (H::kV2) {
        typename H::V2T s2;
        H::initialize(&a.v, &s2);
        a.s = (double)s2;
    } else {
        H::initialize(&a.v);
    }
}

template <class H>
__device__ __forceinline__ bool acc_is_identity(const Acc<H> &a) {
    Acc<H> z;
    acc_init(z);
    return a.v == z.v;
}

// message of one edge from a source whose state is (sval, s2), edge weight w
// (get_partial_result), combined into a (accumulate_nonAtomics)
template <class H>
__device__ __forceinline__ void acc_edge(Acc<H> &a, typename H::V1T sval, double s2, int w, int iter) {
    typename H::V1T m;
    if constexpr (H::kV2) {
        typename H::V2T m2;
        H::get_partial_result(&m, sval, &m2, (typename H::V2T)s2);
        H::accumulate_nonAtomics(&a.v, m, iter, &a.s, m2);
    } else if constexpr (H::kE1) {
        H::get_partial_result(&m, sval, w);
        H::accumulate_nonAtomics(&a.v, m, iter);
    } else {
        H::get_partial_result(&m, sval);
        H::accumulate_nonAtomics(&a.v, m, iter);
    }
}

// a precomputed message value (already through get_partial_result) into a
template <class H>
__device__ __forceinline__ void acc_value(Acc<H> &a, typename H::V1T m, double m2, int iter) {
    if constexpr (H::kV2) H::accumulate_nonAtomics(&a.v, m, iter, &a.s, m2);
    else H::accumulate_nonAtomics(&a.v, m, iter);
}

// b's accumulator merged into a (accumulate_nonAtomics of b's values)
template <class H>
__device__ __forceinline__ void acc_merge(Acc<H> &a, const Acc<H> &b, int iter) {
    acc_value<H>(a, b.v, b.s, iter);
}

// all lanes' accumulators merged, the result in every lane
template <class H>
__device__ __forceinline__ Acc<H> acc_wave(Acc<H> a, int iter) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Acc<H> b;
        b.v = __shfl_xor(a.v, o);
        b.s = __shfl_xor(a.s, o);
        acc_merge<H>(a, b, iter);
    }
    return a;
}

// a merged into a slot shared by several lanes or workgroups (accumulate)
template <class H>
__device__ __forceinline__ void acc_atomic(typename H::V1T *loc, double *loc2, const Acc<H> &a, int iter) {
    if constexpr (H::kV2) H::accumulate(loc, a.v, iter, (typename H::V2T *)loc2, (typename H::V2T)a.s);
    else H::accumulate(loc, a.v, iter);
}

// apply a to a vertex whose value is `old` (update_condition); returns 1 when it changed
template <class H>
__device__ __forceinline__ int acc_apply(typename H::V1T old, typename H::V1T *loc, double *loc2, int *frontier,
                                         const Acc<H> &a, int iter) {
    if constexpr (H::kV2) {
        typename H::V2T o2 = 0;
        return H::update_condition(&old, loc, a.v, iter, frontier, &o2, (typename H::V2T *)loc2, (typename H::V2T)a.s);
    } else {
        return H::update_condition(&old, loc, a.v, iter, frontier);
    }
}

// apply a to a vertex other workgroups may update at the same time
// (fused_update_condition); returns 1 when this call enqueued it
template <class H>
__device__ __forceinline__ int acc_apply_atomic(typename H::V1T *loc, double *loc2, int *frontier, const Acc<H> &a,
                                                int iter) {
    if constexpr (H::kV2) {
        return H::fused_update_condition(loc, loc, a.v, iter, frontier, (typename H::V2T *)loc2,
                                         (typename H::V2T *)loc2, (typename H::V2T)a.s);
    } else {
        return H::fused_update_condition(loc, loc, a.v, iter, frontier);
    }
}

// push along one edge src -> u (edge index p, int or 64-bit): message from the source's
// state, then the push-side apply; returns 1 when u joins the next frontier
template <class H, class PT>
__device__ __forceinline__ int relax(typename H::V1T sval, double s2, int u, PT p, const int *wgt,
                                     typename H::V1T *state, int *vfc, double *sigma, int iter) {
    typename H::V1T msg;
    if constexpr (H::kV2) {
        typename H::V2T m2;
        H::get_partial_result(&msg, sval, &m2, (typename H::V2T)s2);
        return H::fused_update_condition(&state[u], &state[u], msg, iter, nullptr, (typename H::V2T *)&sigma[u],
                                         (typename H::V2T *)&sigma[u], m2);
    } else if constexpr (H::kE1) {
        H::get_partial_result(&msg, sval, wgt[p]);
        return H::fused_update_condition(&state[u], &state[u], msg, iter, &vfc[u]);
    } else {
        H::get_partial_result(&msg, sval);
        return H::fused_update_condition(&state[u], &state[u], msg, iter, nullptr);
    }
}

}  // namespace mg
