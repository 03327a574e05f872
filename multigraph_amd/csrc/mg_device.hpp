// Device primitives for gfx950 (CDNA4, wave64): ballot/mbcnt aggregation, wave and
// block scans and reductions.  Everything is written for 64-lane wavefronts; the
// reference's 32-lane tricks (code/common.h:239-307) are re-derived here, not ported.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef MG_INF
#define MG_INF 0x64646464
#endif

#define MG_WAVE 64

namespace mg {

__device__ __forceinline__ int lane() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p ? 1 : 0); }

// order this wave's LDS stores before its loads of other lanes' slots, and its loads
// before the next stores
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// number of set bits of `m` in lanes below this one
__device__ __forceinline__ int mbcnt(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
// lane l's 64-bit value in every lane (l wave-uniform)
__device__ __forceinline__ uint64_t bcast64(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)readlane((int)(uint32_t)v, l), hi = (uint32_t)readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Wave-aggregated append: one atomicAdd per wave for all lanes with `pred`;
// returns this lane's slot (or -1).  64-lane re-derivation of atomicAggInc
// (code/common.h:243-252).  Call with the whole wave converged.
__device__ __forceinline__ int wave_append(int *ctr, bool pred) {
    uint64_t m = ballot(pred);
    if (m == 0) return -1;
    int leader = __builtin_ctzll(m);
    int base = 0;
    if (lane() == leader) base = atomicAdd(ctr, (int)__builtin_popcountll(m));
    base = readlane(base, leader);
    return pred ? base + mbcnt(m) : -1;
}

// Workgroup-aggregated append: ONE atomicAdd per workgroup of B threads, all of which
// call this together (it holds three barriers); s_wc is B / 64 + 1 ints of LDS.
// Returns this thread's slot (or -1).  Where many workgroups append to one counter at
// once, the per-wave atomics of wave_append queue up on its cache line.
template <int B>
__device__ __forceinline__ int block_append_slot(int *ctr, bool pred, int *s_wc) {
    constexpr int NW = B / 64;
    const uint64_t m = ballot(pred);
    const int wv = threadIdx.x / 64;
    if (lane() == 0) s_wc[wv] = (int)__builtin_popcountll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < NW; ++i) {
            const int c = s_wc[i];
            s_wc[i] = tot;
            tot += c;
        }
        s_wc[NW] = tot ? atomicAdd(ctr, tot) : 0;
    }
    __syncthreads();
    const int pos = s_wc[NW] + s_wc[wv] + mbcnt(m);
    __syncthreads();  // s_wc is refilled by the next call
    return pred ? pos : -1;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
    return v;
}

// inclusive scan over the wave
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    int l = lane();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T t = __shfl_up(v, o);
        if (l >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan for blocks of NW waves; `tmp` holds NW+1 entries in LDS.
// Returns the exclusive prefix, writes the block total to *total.
template <int NW, class T>
__device__ __forceinline__ T block_excl_scan(T v, T *tmp, T *total) {
    const int w = threadIdx.x / MG_WAVE;
    T inc = wave_incl_scan(v);
    if (lane() == MG_WAVE - 1) tmp[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) { T t = tmp[i]; tmp[i] = run; run += t; }
        tmp[NW] = run;
    }
    __syncthreads();
    T r = tmp[w] + inc - v;
    *total = tmp[NW];
    return r;
}

template <int NW, class T>
__device__ __forceinline__ T block_sum(T v, T *tmp) {
    v = wave_sum(v);
    const int w = threadIdx.x / MG_WAVE;
    if (lane() == 0) tmp[w] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NW; ++i) r += tmp[i];
        tmp[0] = r;
    }
    __syncthreads();
    r = tmp[0];
    __syncthreads();
    return r;
}

template <int NW, class T>
__device__ __forceinline__ T block_max(T v, T *tmp) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { const T t = __shfl_xor(v, o); v = t > v ? t : v; }
    const int w = threadIdx.x / MG_WAVE;
    if (lane() == 0) tmp[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T r = tmp[0];
#pragma unroll
        for (int i = 1; i < NW; ++i) r = tmp[i] > r ? tmp[i] : r;
        tmp[0] = r;
    }
    __syncthreads();
    const T r = tmp[0];
    __syncthreads();
    return r;
}

// Neumaier-compensated FP32 summation: (s, c) with s + c the running sum.
struct KSum {
    float s, c;
    __device__ __forceinline__ KSum() : s(0.f), c(0.f) {}
    __device__ __forceinline__ void add(float x) {
        float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ void add(const KSum &o) { add(o.s); add(o.c); }
    __device__ __forceinline__ float value() const { return s + c; }
};

__device__ __forceinline__ KSum wave_ksum(KSum k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        KSum t;
        t.s = __shfl_xor(k.s, o);
        t.c = __shfl_xor(k.c, o);
        k.add(t);
    }
    return k;
}

// streaming load that does not keep the line in the caches
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) { return __builtin_nontemporal_load(p); }

typedef unsigned int mg_u32x4 __attribute__((ext_vector_type(4)));
typedef float mg_f32x4 __attribute__((ext_vector_type(4)));

typedef unsigned int mg_u32x2 __attribute__((ext_vector_type(2)));
// 8-byte streaming load (p must be 8-byte aligned)
__device__ __forceinline__ uint2 ld_stream_u2(const void *p) {
    const mg_u32x2 v = __builtin_nontemporal_load((const mg_u32x2 *)p);
    return make_uint2(v.x, v.y);
}

// 16-byte streaming loads (p must be 16-byte aligned)
__device__ __forceinline__ uint4 ld_stream_u4(const void *p) {
    const mg_u32x4 v = __builtin_nontemporal_load((const mg_u32x4 *)p);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 ld_stream_f4(const void *p) {
    const mg_f32x4 v = __builtin_nontemporal_load((const mg_f32x4 *)p);
    return make_float4(v.x, v.y, v.z, v.w);
}

// 16-byte streaming stores (p must be 16-byte aligned)
__device__ __forceinline__ void st_stream_f4(float *p, float a, float b, float c, float d) {
    const mg_f32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, (mg_f32x4 *)p);
}
typedef int mg_i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_stream_f4(int *p, int a, int b, int c, int d) {
    const mg_i32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, (mg_i32x4 *)p);
}

}  // namespace mg
