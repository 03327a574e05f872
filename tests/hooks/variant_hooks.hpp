// Hooks changed on purpose, to show that every path of the engine calls them: SSSP
// edges weigh twice their weight, BC path-count messages are doubled (so sigma at
// level L is 2^L times the true count), PageRank messages are halved.  Built only into
// multigraph_amd/libmgamd_hookvar.so (make USER_FUNCTION=...), which
// tests/test_gpu_hooks.py runs; the product library keeps the reference's hooks.
#pragma once

namespace mgtest {

struct SSSPDoubleWeight : mg::SSSPHooks {
    __device__ static void get_partial_result(int *loc, int data, int weight) { *loc = data + 2 * weight; }
};

struct BCDoubleSigma : mg::BCHooks {
    __device__ static void get_partial_result(int *loc, int data, double *loc2, double data2) {
        *loc = data + 1;
        *loc2 = 2.0 * data2;
    }
};

struct PRHalfMessage : mg::PRHooks {
    __device__ static void get_partial_result(float *loc, float rank, int deg) {
        *loc = deg > 0 ? 0.5f * (rank / (float)deg) : 0.f;
    }
};

}  // namespace mgtest

#define MG_SSSP_HOOKS mgtest::SSSPDoubleWeight
#define MG_BC_HOOKS mgtest::BCDoubleSigma
#define MG_PR_HOOKS mgtest::PRHalfMessage
