// hbm_copy: streaming kernels over buffers of known size, timed with HIP events.
//
//   hbm_copy [MiB per buffer, default 1024] [repetitions, default 10] [sweep: 0, 1 or 2, default 1]
//
// k_copy (a plain grid-stride float4 copy, 4096 x 256 threads) runs first, `reps`
// times: it calibrates what the HBM traffic counters read for known byte counts.
// With sweep = 1 the tuned kernels follow, each over a sweep of launch shapes, and the
// best of each is printed:
//   k_copy_nt<U>   U float4 loads per thread issued before any store (several loads in
//                  flight), nontemporal 16-byte loads and stores, each workgroup on a
//                  contiguous tile of blockDim * U float4 (coalesced), grid-stride over tiles;
//   k_read_nt<U>   the same loads, summed (read bandwidth alone; one word per
//                  workgroup is written so nothing is optimised away);
//   k_write_nt<U>  nontemporal 16-byte stores only (write bandwidth alone).
// These bound what a streaming kernel of the engine can reach on this chip: the
// engine's gathers are mostly reads, so the read ceiling matters as much as the copy's.
// With sweep = 2, the per-workgroup chunk kernels below run instead of the tuned sweep.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int U>
__global__ void k_copy_nt(const f32x4 *__restrict__ src, f32x4 *__restrict__ dst, long long n) {
    const long long tile = (long long)blockDim.x * U;
    for (long long t0 = (long long)blockIdx.x * tile; t0 < n; t0 += (long long)gridDim.x * tile) {
        f32x4 v[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const long long i = t0 + (long long)j * blockDim.x + threadIdx.x;
            v[j] = i < n ? __builtin_nontemporal_load(&src[i]) : f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const long long i = t0 + (long long)j * blockDim.x + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(v[j], &dst[i]);
        }
    }
}

template <int U>
__global__ void k_read_nt(const f32x4 *__restrict__ src, float *__restrict__ out, long long n) {
    const long long tile = (long long)blockDim.x * U;
    float acc = 0.f;
    for (long long t0 = (long long)blockIdx.x * tile; t0 < n; t0 += (long long)gridDim.x * tile) {
        f32x4 v[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const long long i = t0 + (long long)j * blockDim.x + threadIdx.x;
            v[j] = i < n ? __builtin_nontemporal_load(&src[i]) : f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int j = 0; j < U; ++j) acc += v[j].x + v[j].y + v[j].z + v[j].w;
    }
    if (acc == 12345.f) out[blockIdx.x] = acc;  // never true for the zero-filled input; keeps the loads
}

template <int U>
__global__ void k_write_nt(f32x4 *__restrict__ dst, long long n, float x) {
    const long long tile = (long long)blockDim.x * U;
    for (long long t0 = (long long)blockIdx.x * tile; t0 < n; t0 += (long long)gridDim.x * tile) {
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const long long i = t0 + (long long)j * blockDim.x + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(f32x4{x, x, x, x}, &dst[i]);
        }
    }
}

// k_read_chunks<U> / k_write_chunks<U>: workgroup b streams its own contiguous chunk
// (the buffer cut into gridDim.x * rounds chunks; chunk r * gridDim.x + b in round r),
// the access pattern of the partitioned-layout gathers and scatters (one workgroup per
// partition, each on its own region), against the grid-stride kernels above
template <int U>
__global__ void k_read_chunks(const f32x4 *__restrict__ src, float *__restrict__ out, long long n, int rounds) {
    const long long csz = n / ((long long)gridDim.x * rounds);
    float acc = 0.f;
    for (int r = 0; r < rounds; ++r) {
        const long long b0 = ((long long)r * gridDim.x + blockIdx.x) * csz, b1 = b0 + csz;
        for (long long t0 = b0; t0 < b1; t0 += (long long)blockDim.x * U) {
            f32x4 v[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const long long i = t0 + (long long)j * blockDim.x + threadIdx.x;
                v[j] = i < b1 ? __builtin_nontemporal_load(&src[i]) : f32x4{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
            for (int j = 0; j < U; ++j) acc += v[j].x + v[j].y + v[j].z + v[j].w;
        }
    }
    if (acc == 12345.f) out[blockIdx.x] = acc;
}

template <int U>
__global__ void k_write_chunks(f32x4 *__restrict__ dst, long long n, int rounds, float x) {
    const long long csz = n / ((long long)gridDim.x * rounds);
    for (int r = 0; r < rounds; ++r) {
        const long long b0 = ((long long)r * gridDim.x + blockIdx.x) * csz, b1 = b0 + csz;
        for (long long t0 = b0; t0 < b1; t0 += (long long)blockDim.x * U) {
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const long long i = t0 + (long long)j * blockDim.x + threadIdx.x;
                if (i < b1) __builtin_nontemporal_store(f32x4{x, x, x, x}, &dst[i]);
            }
        }
    }
}

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) { fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); return 1; } \
    } while (0)

struct Best { float ms = 1e30f; int grid = 0, block = 0; };

// time `launch(grid, block)` (median of reps after one warm-up) over a sweep of shapes
template <class F>
static int sweep(const char *name, int U, double bytes, int reps, hipEvent_t e0, hipEvent_t e1, F launch) {
    const int grids[] = {512, 1024, 2048, 4096, 8192, 16384, 32768};
    const int blocks[] = {256, 512, 1024};
    Best b;
    for (int bl : blocks)
        for (int g : grids) {
            launch(g, bl);
            CHECK(hipDeviceSynchronize());
            float t[64];
            const int r = reps < 64 ? reps : 64;
            for (int k = 0; k < r; ++k) {
                CHECK(hipEventRecord(e0, 0));
                launch(g, bl);
                CHECK(hipEventRecord(e1, 0));
                CHECK(hipEventSynchronize(e1));
                CHECK(hipEventElapsedTime(&t[k], e0, e1));
            }
            for (int i = 1; i < r; ++i)
                for (int j = i; j > 0 && t[j] < t[j - 1]; --j) { float x = t[j]; t[j] = t[j - 1]; t[j - 1] = x; }
            const float med = t[r / 2];
            if (med < b.ms) { b.ms = med; b.grid = g; b.block = bl; }
        }
    printf("%s<%d>: best %.4f ms (median of %d) at %d x %d threads: %.1f GB/s\n", name, U, b.ms, reps, b.grid, b.block,
           bytes / (b.ms * 1e-3) / 1e9);
    return 0;
}

int main(int argc, char **argv) {
    const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
    const int reps = argc > 2 ? atoi(argv[2]) : 10;
    const int do_sweep = argc > 3 ? atoi(argv[3]) : 1;
    if (mib < 1 || reps < 1) { fprintf(stderr, "usage: hbm_copy [MiB] [reps] [sweep 0|1|2]\n"); return 2; }
    const long long bytes = mib << 20;
    const long long n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    float *o = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMalloc(&o, sizeof(float) * 32768));
    CHECK(hipMemset(a, 0, bytes));
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
    if (do_sweep == 1) {
        const f32x4 *s4 = (const f32x4 *)a;
        f32x4 *d4 = (f32x4 *)b;
#define SWEEP_U(U)                                                                                                    \
        if (sweep("k_copy_nt", U, 2.0 * bytes, reps, e0, e1, [&](int g, int bl) {                                     \
                hipLaunchKernelGGL(k_copy_nt<U>, dim3(g), dim3(bl), 0, 0, s4, d4, n); })) return 1;                    \
        if (sweep("k_read_nt", U, 1.0 * bytes, reps, e0, e1, [&](int g, int bl) {                                     \
                hipLaunchKernelGGL(k_read_nt<U>, dim3(g), dim3(bl), 0, 0, s4, o, n); })) return 1;                     \
        if (sweep("k_write_nt", U, 1.0 * bytes, reps, e0, e1, [&](int g, int bl) {                                    \
                hipLaunchKernelGGL(k_write_nt<U>, dim3(g), dim3(bl), 0, 0, d4, n, 0.f); })) return 1;
        SWEEP_U(1)
        SWEEP_U(2)
        SWEEP_U(4)
        SWEEP_U(8)
#undef SWEEP_U
    }
    if (do_sweep == 2) {  // one chunk per workgroup: 256 / 512 / 768 workgroups of 1024 threads, 1 or 3 rounds
        const f32x4 *s4 = (const f32x4 *)a;
        f32x4 *d4 = (f32x4 *)b;
        for (int g : {256, 512, 768})
            for (int rounds : {1, 3}) {
                float best_r = 1e30f, best_w = 1e30f;
                for (int k = 0; k <= reps; ++k) {  // the first launch warms up
                    float ms = 0.f;
                    CHECK(hipEventRecord(e0, 0));
                    hipLaunchKernelGGL(k_read_chunks<4>, dim3(g), dim3(1024), 0, 0, s4, o, n, rounds);
                    CHECK(hipEventRecord(e1, 0));
                    CHECK(hipEventSynchronize(e1));
                    CHECK(hipEventElapsedTime(&ms, e0, e1));
                    if (k > 0 && ms < best_r) best_r = ms;
                    CHECK(hipEventRecord(e0, 0));
                    hipLaunchKernelGGL(k_write_chunks<4>, dim3(g), dim3(1024), 0, 0, d4, n, rounds, 0.f);
                    CHECK(hipEventRecord(e1, 0));
                    CHECK(hipEventSynchronize(e1));
                    CHECK(hipEventElapsedTime(&ms, e0, e1));
                    if (k > 0 && ms < best_w) best_w = ms;
                }
                printf("chunks: %d workgroups x 1024 threads, %d round(s): read %.1f GB/s, write %.1f GB/s (best of %d)\n", g, rounds,
                       bytes / (best_r * 1e-3) / 1e9, bytes / (best_w * 1e-3) / 1e9, reps);
            }
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(o));
    return 0;
}
