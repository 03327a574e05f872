// hbm_copy: a plain streaming copy of a buffer of known size, timed with HIP events.
// It calibrates what the HBM traffic counters (FETCH_SIZE, WRITE_SIZE) read for
// known byte counts, and reports the copy bandwidth the chip sustains, the
// practical ceiling beside the 8 TB/s peak.
//
//   hbm_copy [MiB per buffer, default 1024] [repetitions, default 10]
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) { fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); return 1; } \
    } while (0)

int main(int argc, char **argv) {
    const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
    const int reps = argc > 2 ? atoi(argv[2]) : 10;
    if (mib < 1 || reps < 1) { fprintf(stderr, "usage: hbm_copy [MiB] [reps]\n"); return 2; }
    const long long bytes = mib << 20;
    const long long n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int grid = 256 * 16;  // CUs x 16 workgroups of 256 threads, grid-stride
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);  // warm-up
    CHECK(hipDeviceSynchronize());
    float best = 1e30f, total = 0.f;
    for (int r = 0; r < reps; ++r) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
        total += ms;
    }
    printf("k_copy: %lld bytes read + %lld bytes written per launch, %d launches, avg %.4f ms, best %.4f ms, "
           "%.1f GB/s (avg), %.1f GB/s (best)\n",
           bytes, bytes, reps + 1, total / reps, best, 2.0 * bytes / (total / reps * 1e-3) / 1e9, 2.0 * bytes / (best * 1e-3) / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
