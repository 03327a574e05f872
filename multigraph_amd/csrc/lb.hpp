// Edge-balanced partitioning of a CSR edge range across workgroups: chunk c holds
// edges [c*kEPB, (c+1)*kEPB); its owners (row vertices) are found once per chunk by
// marking each owner's first in-chunk edge in LDS and max-scanning, so rows of any
// length, including empty ones, cost O(1) per edge.
#pragma once
#include "engine.hpp"
#include "mg_device.hpp"

namespace mg {

// cstart[c] = owner of the first edge of chunk c; cstart[nch] = owner of the last edge
// (OT: int or 64-bit row offsets)
template <class OT = int>
static __global__ void k_chunk_part(int nv, long long ne, const OT *__restrict__ off, int nch, int *__restrict__ cstart) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= nch; c += gridDim.x * blockDim.x) {
        const long long e = c < nch ? (long long)c * kEPB : ne - 1;
        int lo = 0, hi = nv - 1;  // largest u with off[u] <= e
        while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (off[mid] <= e) lo = mid; else hi = mid - 1; }
        cstart[c] = lo;
    }
}

// s_own[j] = owner of edge e0 + j for j < e1 - e0 (owners lie in [u0, u1])
__device__ __forceinline__ void chunk_owners(const int *__restrict__ off, int u0, int u1, int e0, int e1, int *s_own) {
    __shared__ int s_tmp[kBlock / MG_WAVE + 1];
    for (int j = threadIdx.x; j < kEPB; j += kBlock) s_own[j] = -1;
    __syncthreads();
    for (int u = u0 + threadIdx.x; u <= u1; u += kBlock) {
        const int a = off[u], b = off[u + 1];
        if (b > a && a < e1 && b > e0) atomicMax(&s_own[max(a, e0) - e0], u);
    }
    __syncthreads();
    constexpr int kPer = kEPB / kBlock;  // consecutive entries per thread
    int loc[kPer];
    int m = -1;
#pragma unroll
    for (int k = 0; k < kPer; ++k) { m = max(m, s_own[threadIdx.x * kPer + k]); loc[k] = m; }
    int x = m;
#pragma unroll
    for (int o = 1; o < MG_WAVE; o <<= 1) { const int t = __shfl_up(x, o); if (lane() >= o) x = max(x, t); }
    if (lane() == MG_WAVE - 1) s_tmp[threadIdx.x / MG_WAVE] = x;
    __syncthreads();
    int before = __shfl_up(x, 1);
    if (lane() == 0) before = -1;
    for (int w = 0; w < (int)(threadIdx.x / MG_WAVE); ++w) before = max(before, s_tmp[w]);
#pragma unroll
    for (int k = 0; k < kPer; ++k) s_own[threadIdx.x * kPer + k] = max(before, loc[k]);
    __syncthreads();
}

}  // namespace mg
