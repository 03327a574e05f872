// Host-side internals shared by the graph reader/generators and the device engine.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>
#include <atomic>
#include <thread>
#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <cstdarg>

#include "multigraph_amd.h"

// Host CSR after reading: the role of struct csc_package (code/common.h:170-182).
struct mg_graph {
    int32_t nv = 0;            // vertices after compaction
    int64_t ne = 0;            // directed edges stored
    int32_t orig_nv = 0;       // vertex count in the input header
    int64_t entries = 0;       // entries read (after symmetrization), reference `edges`
    int32_t start_p = 0;       // remapped source (read_input.cu:130)
    int flags = 0;
    bool weighted = true;
    bool symmetric = true;     // every stored edge has its reverse stored too
    std::vector<int32_t> row;  // nv + 1 (empty on a wide graph)
    std::vector<int64_t> row64;  // nv + 1 instead of row when ne > 2^31 - 1 (a wide graph)
    std::vector<int32_t> col;  // ne
    std::vector<int32_t> w;    // ne
    std::vector<int32_t> orig; // compacted id -> original id
    std::vector<int32_t> valid;// original id -> compacted prefix (length orig_nv + 1)
    bool wide() const { return !row64.empty(); }
    int64_t off(int64_t v) const { return row64.empty() ? (int64_t)row[(size_t)v] : row64[(size_t)v]; }
};

void mg_set_error(const char *fmt, ...);
int mg_fail(int code, const char *fmt, ...);

static inline uint64_t mg_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Weight of the undirected edge {u, v}: uniform in [1, 64], identical in both directions.
static inline int32_t mg_edge_weight(uint64_t seed, int32_t u, int32_t v) {
    uint64_t lo = (uint64_t)(uint32_t)(u < v ? u : v), hi = (uint64_t)(uint32_t)(u < v ? v : u);
    return 1 + (int32_t)(mg_splitmix64(mg_splitmix64(seed ^ 0x5bd1e995ULL) ^ ((lo << 32) | hi)) % 64);
}

int mg_host_threads();

// Run f(begin, end) over [0, n) on the host thread pool.
template <class F>
void mg_parallel_for(int64_t n, F f, int64_t grain = 1 << 16) {
    int T = mg_host_threads();
    int64_t chunks = (n + grain - 1) / grain;
    if (T > chunks) T = (int)chunks;
    if (T <= 1) { if (n > 0) f((int64_t)0, n); return; }
    std::vector<std::thread> th;
    int64_t per = (n + T - 1) / T;
    for (int t = 0; t < T; ++t) {
        int64_t b = per * t, e = std::min<int64_t>(n, b + per);
        if (b >= e) break;
        th.emplace_back([=, &f]() { f(b, e); });
    }
    for (auto &x : th) x.join();
}

// Same contract, but [0, n) is handed out in pieces of `grain` from a shared counter:
// for loops over vertices whose cost follows a skewed degree distribution.
template <class F>
void mg_parallel_dyn(int64_t n, F f, int64_t grain = 4096) {
    int T = mg_host_threads();
    const int64_t chunks = (n + grain - 1) / grain;
    if (T > chunks) T = (int)chunks;
    if (T <= 1) { if (n > 0) f((int64_t)0, n); return; }
    std::atomic<int64_t> next{0};
    auto body = [&]() {
        for (int64_t c; (c = next.fetch_add(1)) < chunks;) f(c * grain, std::min<int64_t>(n, (c + 1) * grain));
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(body);
    body();
    for (auto &x : th) x.join();
}

// generate_CSC semantics on entries already parsed (one per MTX line, 0-based ids).
int mg_build_from_entries(int32_t nv, int64_t nnz, const int32_t *src, const int32_t *dst,
                          const int32_t *val, bool sym, bool cc, int source, mg_graph *g);
