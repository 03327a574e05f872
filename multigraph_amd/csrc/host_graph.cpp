// Graph input on the host: Matrix Market reader with the reference's generate_CSC
// semantics (code/read_input.cu:11-254), seeded RMAT and grid generators, union of
// graphs, and the MTX writer.  All of it is host C++; nothing here touches the GPU.
#include <cstring>
#include <cmath>
#include <climits>
#include <mutex>
#include <atomic>

#include "mg_internal.hpp"

static thread_local std::string g_err;

void mg_set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int mg_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *mg_last_error(void) { return g_err.c_str(); }
extern "C" const char *mg_version(void) { return "multigraph-amd 0.1 (gfx950)"; }

int mg_host_threads() {
    static int n = [] {
        const char *s = getenv("MG_HOST_THREADS");
        if (!s) s = getenv("OMP_NUM_THREADS");
        int v = s ? atoi(s) : 0;
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        if (v <= 0) v = 1;
        if (v > 16) v = 16;
        return v;
    }();
    return n;
}

// Entries of one build (after symmetrization): int32 row offsets up to 2^31 - 1 kept
// edges, 64-bit ones past that (a "wide" graph, row64), up to 2^32 entries (RMAT-27
// with edge factor 16).
constexpr int64_t kMaxEntries = (int64_t)1 << 32;

// ---------------------------------------------------------------------------
// generate_CSC core (read_input.cu:88-171), restated for large inputs:
//  * entry i of the input becomes (src,dst,val) and, when symmetrizing, (dst,src,val)
//    directly after it (read_input.cu:88-95);
//  * the sort by (src,dst) (read_input.cu:97) is made stable, so of duplicate entries
//    the first in input order survives (the reference's qsort leaves this unspecified);
//  * duplicates dropped, CC keeps src < dst only; entry 0 of the sorted order is always
//    kept and never marks its endpoints valid (read_input.cu:105-114);
//  * isolated vertices compacted away and the source remapped (read_input.cu:115-139).
// ---------------------------------------------------------------------------
int mg_build_from_entries(int32_t nv, int64_t nnz, const int32_t *src, const int32_t *dst,
                          const int32_t *val, bool sym, bool cc, int source, mg_graph *g) {
    if (nv < 0 || nnz < 0) return mg_fail(MG_ERR_ARG, "negative size");
    const int64_t n = sym ? 2 * nnz : nnz;
    if (n > kMaxEntries) return mg_fail(MG_ERR_LIMIT, "%lld entries exceed the CSR limit of %lld", (long long)n, (long long)kMaxEntries);
    for (int64_t i = 0; i < nnz; ++i) {
        if (src[i] < 0 || src[i] >= nv || dst[i] < 0 || dst[i] >= nv)
            return mg_fail(MG_ERR_RANGE, "A vertex id is out of range %d %d", src[i], dst[i]);
    }
    // Counting sort by src, stable in expanded-entry order, in two parallel passes:
    // entries are first split by src into buckets of consecutive vertices (each thread
    // takes a contiguous slice of the input and writes its bucket runs in order), then
    // every bucket is counting-sorted on its own vertex range.
    const int T = std::max(1, std::min<int>(mg_host_threads(), (int)std::min<int64_t>(64, (nnz >> 16) + 1)));
    int shift = 0;
    while (((int64_t)std::max(nv - 1, 0) >> shift) >= 4096) ++shift;
    const int nb = (int)(((int64_t)std::max(nv - 1, 0) >> shift) + 1);
    std::vector<int64_t> bcnt((size_t)T * nb + 1, 0);  // [bucket][thread]
    auto slice = [&](int t, int64_t *lo, int64_t *hi) {
        const int64_t per = (nnz + T - 1) / T;
        *lo = std::min<int64_t>(nnz, per * t);
        *hi = std::min<int64_t>(nnz, *lo + per);
    };
    auto run_threads = [&](auto fn) {
        std::vector<std::thread> th;
        for (int t = 1; t < T; ++t) th.emplace_back([&, t]() { fn(t); });
        fn(0);
        for (auto &x : th) x.join();
    };
    run_threads([&](int t) {
        int64_t lo, hi;
        slice(t, &lo, &hi);
        for (int64_t i = lo; i < hi; ++i) {
            bcnt[(size_t)(src[i] >> shift) * T + t]++;
            if (sym) bcnt[(size_t)(dst[i] >> shift) * T + t]++;
        }
    });
    {
        int64_t run = 0;
        for (size_t k = 0; k < (size_t)T * nb; ++k) { const int64_t c = bcnt[k]; bcnt[k] = run; run += c; }
        bcnt[(size_t)T * nb] = run;
    }
    std::vector<int32_t> xs((size_t)n), xd((size_t)n), xw((size_t)n);
    run_threads([&](int t) {
        int64_t lo, hi;
        slice(t, &lo, &hi);
        std::vector<int64_t> cur((size_t)nb);
        for (int k = 0; k < nb; ++k) cur[k] = bcnt[(size_t)k * T + t];
        for (int64_t i = lo; i < hi; ++i) {
            const int32_t wv = val ? val[i] : 1;
            int64_t q = cur[src[i] >> shift]++;
            xs[q] = src[i]; xd[q] = dst[i]; xw[q] = wv;
            if (sym) {
                q = cur[dst[i] >> shift]++;
                xs[q] = dst[i]; xd[q] = src[i]; xw[q] = wv;
            }
        }
    });
    std::vector<int64_t> start((size_t)nv + 1, 0);
    // buckets of low RMAT ids hold far more entries: hand them out dynamically
    auto each_bucket = [&](auto fn) {
        std::atomic<int64_t> next{0};
        run_threads([&](int) { for (int64_t bk; (bk = next.fetch_add(1)) < nb;) fn(bk); });
    };
    each_bucket([&](int64_t bk) {
        for (int64_t q = bcnt[(size_t)bk * T]; q < bcnt[(size_t)(bk + 1) * T]; ++q) start[(size_t)xs[q] + 1]++;
    });
    for (int32_t v = 0; v < nv; ++v) start[v + 1] += start[v];
    std::vector<int32_t> sd((size_t)n), sw((size_t)n);
    each_bucket([&](int64_t bk) {
        const int64_t v0 = bk << shift, v1 = std::min<int64_t>(nv, (bk + 1) << shift);
        std::vector<int64_t> pos(start.begin() + v0, start.begin() + v1);
        for (int64_t q = bcnt[(size_t)bk * T]; q < bcnt[(size_t)(bk + 1) * T]; ++q) {
            const int64_t p = pos[xs[q] - v0]++;
            sd[p] = xd[q]; sw[p] = xw[q];
        }
    });
    std::vector<int32_t>().swap(xs);
    std::vector<int32_t>().swap(xd);
    std::vector<int32_t>().swap(xw);
    // stable sort of each row by dst: sort (dst, position) pairs
    std::vector<uint8_t> keep((size_t)n, 0);
    std::vector<uint8_t> mark((size_t)nv + 1, 0);
    int64_t first_nonempty = -1;
    int32_t first_row = -1;
    for (int32_t v = 0; v < nv; ++v) if (start[v + 1] > start[v]) { first_nonempty = start[v]; first_row = v; break; }
    mg_parallel_dyn(nv, [&](int64_t b, int64_t e) {
        std::vector<uint64_t> key;
        std::vector<int32_t> td, tw;
        for (int64_t v = b; v < e; ++v) {
            int64_t s = start[v], t = start[v + 1], len = t - s;
            if (len == 0) continue;
            if (len > 1) {
                key.resize(len);
                for (int64_t k = 0; k < len; ++k) key[k] = ((uint64_t)(uint32_t)sd[s + k] << 32) | (uint64_t)k;
                std::sort(key.begin(), key.end());
                td.resize(len); tw.resize(len);
                for (int64_t k = 0; k < len; ++k) { td[k] = (int32_t)(key[k] >> 32); tw[k] = sw[s + (int64_t)(uint32_t)key[k]]; }
                std::copy(td.begin(), td.end(), sd.begin() + s);
                std::copy(tw.begin(), tw.end(), sw.begin() + s);
            }
        }
    }, 1024);
    // keep / valid marks (read_input.cu:105-114)
    mg_parallel_dyn(nv, [&](int64_t b, int64_t e) {
        for (int64_t v = b; v < e; ++v) {
            for (int64_t p = start[v]; p < start[v + 1]; ++p) {
                if (p == first_nonempty) { keep[p] = 1; continue; }
                int32_t d = sd[p];
                bool k = !(p > start[v] && sd[p - 1] == d);
                if (cc && v >= d) k = false;
                keep[p] = k;
                __atomic_store_n(&mark[v], (uint8_t)1, __ATOMIC_RELAXED);
                __atomic_store_n(&mark[d], (uint8_t)1, __ATOMIC_RELAXED);
            }
        }
    }, 1024);
    g->orig_nv = nv;
    g->valid.assign((size_t)nv + 1, 0);
    for (int32_t v = 0; v < nv; ++v) g->valid[v + 1] = g->valid[v] + (mark[v] ? 1 : 0);
    const int32_t cnv = g->valid[nv];
    g->orig.resize(cnv);
    for (int32_t v = 0; v < nv; ++v) if (mark[v]) g->orig[g->valid[v]] = v;
    if (source < 0 || (nv > 0 && source >= nv)) return mg_fail(MG_ERR_ARG, "source %d outside [0,%d)", source, nv);
    g->start_p = g->valid[source];
    // Entry 0 of the sorted order never marks its endpoints valid; when no other entry
    // names them, the reference maps them past the last compacted id and writes out of
    // its arrays.  Such input is rejected here instead.
    if (first_nonempty >= 0 && (g->valid[first_row] >= cnv || g->valid[sd[first_nonempty]] >= cnv))
        return mg_fail(MG_ERR_RANGE, "entry (%d,%d) names a vertex no other entry does; its compacted id would be out of range",
                       first_row + 1, sd[first_nonempty] + 1);
    // CSR of kept edges in compacted ids (read_input.cu:135-171)
    g->nv = cnv;
    std::vector<int64_t> rowcnt((size_t)nv, 0);
    mg_parallel_dyn(nv, [&](int64_t b, int64_t e) {
        for (int64_t v = b; v < e; ++v) {
            int64_t c = 0;
            for (int64_t p = start[v]; p < start[v + 1]; ++p) c += keep[p];
            rowcnt[v] = c;
        }
    }, 4096);
    // A kept entry whose src is not marked valid (possible only through the entry-0
    // quirk without symmetrization) is filed under valid[src], the next valid id,
    // exactly as read_input.cu:136 does; rows are merged in id order below.
    std::vector<int64_t> rstart((size_t)nv + 1, 0);
    for (int32_t v = 0; v < nv; ++v) rstart[v + 1] = rstart[v] + rowcnt[v];
    const int64_t ne = rstart[nv];
    g->ne = ne;
    g->entries = n;
    g->col.resize((size_t)ne);
    g->w.resize((size_t)ne);
    {
        std::vector<int64_t> r((size_t)cnv + 1, 0);
        for (int32_t v = 0; v < nv; ++v) {
            if (!rowcnt[v]) continue;
            r[(size_t)g->valid[v] + 1] = rstart[v + 1];
        }
        // fill rows of compacted vertices without kept edges (read_input.cu:167-170)
        for (int32_t v = 1; v <= cnv; ++v) if (r[v] < r[v - 1]) r[v] = r[v - 1];
        r[cnv] = ne;
        if (ne > (int64_t)INT32_MAX) {
            g->row64 = std::move(r);
        } else {
            g->row.resize((size_t)cnv + 1);
            for (size_t v = 0; v <= (size_t)cnv; ++v) g->row[v] = (int32_t)r[v];
        }
    }
    mg_parallel_dyn(nv, [&](int64_t b, int64_t e) {
        for (int64_t v = b; v < e; ++v) {
            int64_t o = rstart[v];
            for (int64_t p = start[v]; p < start[v + 1]; ++p) if (keep[p]) {
                g->col[o] = g->valid[sd[p]];
                g->w[o] = sw[p];
                ++o;
            }
        }
    }, 4096);
    g->flags = (sym ? MG_READ_SYM : 0) | (cc ? MG_READ_CC : 0);
    g->symmetric = sym && !cc;
    return MG_OK;
}

// ---------------------------------------------------------------------------
// Matrix Market reader (read_input.cu:11-96)
// ---------------------------------------------------------------------------
namespace {
struct Cursor {
    const char *p, *end;
    bool skip_ws() { while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p; return p < end; }
    bool read_long(long long *out) {
        if (!skip_ws()) return false;
        char *e = nullptr;
        long long v = strtoll(p, &e, 10);
        if (e == p) return false;
        p = e; *out = v; return true;
    }
    bool read_float(float *out) {
        if (!skip_ws()) return false;
        char *e = nullptr;
        float v = strtof(p, &e);
        if (e == p) return false;
        p = e; *out = v; return true;
    }
};
}  // namespace

extern "C" int mg_graph_load_mtx(const char *path, int source, int flags, uint64_t weight_seed, mg_graph **out) {
    if (!path || !out) return mg_fail(MG_ERR_ARG, "null argument");
    *out = nullptr;
    FILE *fp = fopen(path, "rb");
    if (!fp) return mg_fail(MG_ERR_IO, "cannot open %s", path);
    fseek(fp, 0, SEEK_END);
    long long size = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    std::vector<char> buf((size_t)size + 1);
    if (size > 0 && fread(buf.data(), 1, (size_t)size, fp) != (size_t)size) { fclose(fp); return mg_fail(MG_ERR_IO, "short read on %s", path); }
    fclose(fp);
    buf[size] = 0;
    const char *p = buf.data(), *end = buf.data() + size;
    // header line: symmetric / Hermitian, pattern / complex (read_input.cu:25-35)
    const char *nl = (const char *)memchr(p, '\n', end - p);
    std::string header(p, nl ? nl : end);
    if (header.compare(0, 14, "%%MatrixMarket") != 0) return mg_fail(MG_ERR_IO, "%s: missing %%%%MatrixMarket header", path);
    bool hsym = header.find("symmetric") != std::string::npos || header.find("Hermitian") != std::string::npos;
    int nflag = header.find("pattern") != std::string::npos ? 0 : (header.find("complex") != std::string::npos ? -1 : 1);
    bool sym = (flags & MG_READ_HEADER) ? hsym : ((flags & MG_READ_SYM) != 0);
    // skip following lines that contain '%' (read_input.cu:46-56)
    p = nl ? nl + 1 : end;
    while (p < end) {
        const char *e2 = (const char *)memchr(p, '\n', end - p);
        const char *le = e2 ? e2 : end;
        if (memchr(p, '%', le - p) == nullptr) break;
        p = e2 ? e2 + 1 : end;
    }
    Cursor c{p, end};
    long long nv, dummy, nnz;
    if (!c.read_long(&nv) || !c.read_long(&dummy) || !c.read_long(&nnz)) return mg_fail(MG_ERR_IO, "%s: malformed size line", path);
    if (nv < 0 || nnz < 0 || nv > INT32_MAX) return mg_fail(MG_ERR_IO, "%s: bad sizes", path);
    std::vector<int32_t> s((size_t)nnz), d((size_t)nnz), v((size_t)nnz);
    for (long long i = 0; i < nnz; ++i) {
        long long a, b;
        if (!c.read_long(&a) || !c.read_long(&b)) return mg_fail(MG_ERR_IO, "%s: entry %lld missing", path, i + 1);
        a -= 1; b -= 1;
        if (a < 0 || a >= nv || b < 0 || b >= nv)
            return mg_fail(MG_ERR_RANGE, "A vertex id is out of range %lld %lld", a, b);
        s[i] = (int32_t)a; d[i] = (int32_t)b;
        if (nflag == 0) {
            v[i] = 1 + (int32_t)(mg_splitmix64(weight_seed ^ mg_splitmix64((uint64_t)i)) % 64);
        } else {
            float f1 = 0.f, f2 = 0.f;
            if (!c.read_float(&f1)) return mg_fail(MG_ERR_IO, "%s: value of entry %lld missing", path, i + 1);
            if (nflag < 0 && !c.read_float(&f2)) return mg_fail(MG_ERR_IO, "%s: imaginary part of entry %lld missing", path, i + 1);
            double fv = f1;
            if (!(fv > -2147483648.0)) fv = -2147483648.0;
            if (fv > 2147483647.0) fv = 2147483647.0;
            v[i] = (int32_t)fv;  // (int)float truncation, read_input.cu:81,86
        }
    }
    if (source < 0 || source >= nv) return mg_fail(MG_ERR_ARG, "source vertex %d outside [0,%lld)", source, nv);
    mg_graph *g = new mg_graph();
    int rc = mg_build_from_entries((int32_t)nv, nnz, s.data(), d.data(), v.data(), sym, (flags & MG_READ_CC) != 0, source, g);
    if (rc != MG_OK) { delete g; return rc; }
    g->weighted = nflag != 0;
    *out = g;
    return MG_OK;
}

extern "C" int mg_graph_from_coo(int nv, int64_t nnz, const int32_t *src, const int32_t *dst, const int32_t *val,
                                 int source, int flags, mg_graph **out) {
    if (!out || (nnz > 0 && (!src || !dst))) return mg_fail(MG_ERR_ARG, "null argument");
    *out = nullptr;
    if (source < 0 || source >= (nv > 0 ? nv : 1)) return mg_fail(MG_ERR_ARG, "source vertex %d outside [0,%d)", source, nv);
    mg_graph *g = new mg_graph();
    int rc = mg_build_from_entries(nv, nnz, src, dst, val, (flags & (MG_READ_SYM | MG_READ_HEADER)) != 0,
                                   (flags & MG_READ_CC) != 0, source, g);
    if (rc != MG_OK) { delete g; return rc; }
    *out = g;
    return MG_OK;
}

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------
extern "C" int mg_graph_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                                 uint64_t weight_seed, int flags, mg_graph **out) {
    if (!out || scale < 1 || scale > 30 || edge_factor < 1) return mg_fail(MG_ERR_ARG, "bad RMAT parameters");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return mg_fail(MG_ERR_ARG, "bad RMAT probabilities");
    *out = nullptr;
    const int64_t nv = (int64_t)1 << scale;
    const int64_t draws = nv * (int64_t)edge_factor;
    if (draws * 2 > kMaxEntries) return mg_fail(MG_ERR_LIMIT, "RMAT of %lld edge draws exceeds the CSR limit", (long long)draws);
    std::vector<int32_t> s((size_t)draws), d((size_t)draws), w((size_t)draws);
    std::vector<uint8_t> self((size_t)draws, 0);
    // quadrant choice per level from 16 random bits (4 levels per 64-bit draw), with the
    // probabilities quantised to 1/65536
    const uint32_t ta = (uint32_t)std::llround(a * 65536.0), tab = (uint32_t)std::llround((a + b) * 65536.0),
                   tabc = (uint32_t)std::llround((a + b + c) * 65536.0);
    const uint64_t base = mg_splitmix64(seed ^ 0x2545F4914F6CDD1DULL);
    mg_parallel_for(draws, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            uint64_t st = mg_splitmix64(base ^ (uint64_t)k * 0x9E3779B97F4A7C15ULL);
            uint32_t u = 0, v = 0;
            uint64_t bits = 0;
            for (int bit = 0; bit < scale; ++bit) {
                if ((bit & 3) == 0) { st = mg_splitmix64(st); bits = st; }
                const uint32_t r = (uint32_t)(bits & 0xffff);
                bits >>= 16;
                const uint32_t qu = r >= tab, qv = (r >= ta) & ((r < tab) | (r >= tabc));
                u = (u << 1) | qu;
                v = (v << 1) | qv;
            }
            s[k] = (int32_t)u; d[k] = (int32_t)v;
            self[k] = (u == v);
            w[k] = mg_edge_weight(weight_seed, (int32_t)u, (int32_t)v);
        }
    });
    // drop self-loops (the workloads remove them before reading)
    int64_t m = 0;
    for (int64_t k = 0; k < draws; ++k) {
        if (self[k]) continue;
        s[m] = s[k]; d[m] = d[k]; w[m] = w[k]; ++m;
    }
    mg_graph *g = new mg_graph();
    int rc = mg_build_from_entries((int32_t)nv, m, s.data(), d.data(), w.data(), (flags & MG_READ_SYM) != 0,
                                   (flags & MG_READ_CC) != 0, 0, g);
    if (rc != MG_OK) { delete g; return rc; }
    *out = g;
    return MG_OK;
}

extern "C" int mg_graph_gen_grid(int width, int height, uint64_t weight_seed, int flags, mg_graph **out) {
    if (!out || width < 1 || height < 1) return mg_fail(MG_ERR_ARG, "bad grid size");
    *out = nullptr;
    int64_t nv = (int64_t)width * height;
    if (nv > INT32_MAX / 4) return mg_fail(MG_ERR_LIMIT, "grid too large");
    std::vector<int32_t> s, d, w;
    s.reserve(nv * 2); d.reserve(nv * 2); w.reserve(nv * 2);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            int32_t id = y * width + x;
            if (x + 1 < width) { s.push_back(id); d.push_back(id + 1); w.push_back(mg_edge_weight(weight_seed, id, id + 1)); }
            if (y + 1 < height) { s.push_back(id); d.push_back(id + width); w.push_back(mg_edge_weight(weight_seed, id, id + width)); }
        }
    mg_graph *g = new mg_graph();
    int rc = mg_build_from_entries((int32_t)nv, (int64_t)s.size(), s.data(), d.data(), w.data(), (flags & MG_READ_SYM) != 0,
                                   (flags & MG_READ_CC) != 0, 0, g);
    if (rc != MG_OK) { delete g; return rc; }
    *out = g;
    return MG_OK;
}

// A CSR built by the caller, taken as it is: vertex ids are kept (no compaction), each
// row is sorted by column (weights carried along) and a repeated column keeps its
// first entry, as the reader's stable sort does; the graph counts as symmetric when
// every stored edge has its reverse, with the same weight.  Weights default to 1.
extern "C" int mg_graph_from_csr(int nv, const int64_t *row, const int32_t *col, const int32_t *w, mg_graph **out) {
    if (!out || !row || (nv > 0 && row[nv] > 0 && !col)) return mg_fail(MG_ERR_ARG, "null argument");
    *out = nullptr;
    if (nv < 0) return mg_fail(MG_ERR_ARG, "negative vertex count");
    if (row[0] != 0) return mg_fail(MG_ERR_ARG, "row[0] must be 0");
    for (int v = 0; v < nv; ++v)
        if (row[v + 1] < row[v]) return mg_fail(MG_ERR_ARG, "row offsets decrease at vertex %d", v);
    const int64_t nnz = row[nv];
    if (nnz > INT32_MAX) return mg_fail(MG_ERR_LIMIT, "%lld edges exceed int32 offsets", (long long)nnz);
    for (int64_t p = 0; p < nnz; ++p)
        if (col[p] < 0 || col[p] >= nv) return mg_fail(MG_ERR_RANGE, "column %d of edge %lld outside [0,%d)", col[p], (long long)p, nv);
    mg_graph *g = new mg_graph();
    g->nv = nv;
    g->orig_nv = nv;
    g->row.assign((size_t)nv + 1, 0);
    g->col.reserve((size_t)nnz);
    g->w.reserve((size_t)nnz);
    std::vector<std::pair<int32_t, int32_t>> r;
    for (int v = 0; v < nv; ++v) {
        r.clear();
        for (int64_t p = row[v]; p < row[v + 1]; ++p) r.emplace_back(col[p], w ? w[p] : 1);
        std::stable_sort(r.begin(), r.end(), [](const std::pair<int32_t, int32_t> &a, const std::pair<int32_t, int32_t> &b) {
            return a.first < b.first;
        });
        for (size_t k = 0; k < r.size(); ++k) {
            if (k > 0 && r[k].first == r[k - 1].first) continue;
            g->col.push_back(r[k].first);
            g->w.push_back(r[k].second);
        }
        g->row[v + 1] = (int32_t)g->col.size();
    }
    g->ne = (int64_t)g->col.size();
    g->entries = g->ne;
    g->orig.resize(nv);
    g->valid.resize((size_t)nv + 1);
    for (int v = 0; v < nv; ++v) { g->orig[v] = v; g->valid[v] = v; }
    g->valid[nv] = nv;
    g->start_p = 0;
    g->flags = 0;
    // symmetric: every edge's reverse is stored with the same weight (pull levels read
    // a row as in-edges, SSSP with their weights)
    bool sym = true;
    for (int v = 0; v < nv && sym; ++v)
        for (int32_t p = g->row[v]; p < g->row[v + 1] && sym; ++p) {
            const int32_t u = g->col[p];
            const auto b = g->col.begin() + g->row[u], e = g->col.begin() + g->row[u + 1];
            const auto it = std::lower_bound(b, e, v);
            sym = it != e && *it == v && g->w[(size_t)(it - g->col.begin())] == g->w[(size_t)p];
        }
    g->symmetric = sym;
    if (sym) g->flags |= MG_READ_SYM;
    *out = g;
    return MG_OK;
}

extern "C" int mg_graph_union(int count, const mg_graph *const *graphs, mg_graph **out) {
    if (!out || count < 1 || !graphs) return mg_fail(MG_ERR_ARG, "bad union arguments");
    int64_t nv = 0, ne = 0;
    for (int i = 0; i < count; ++i) { if (!graphs[i]) return mg_fail(MG_ERR_ARG, "null graph"); nv += graphs[i]->nv; ne += graphs[i]->ne; }
    if (nv > INT32_MAX || ne > INT32_MAX) return mg_fail(MG_ERR_LIMIT, "union too large");
    mg_graph *g = new mg_graph();
    g->nv = (int32_t)nv; g->ne = ne; g->orig_nv = (int32_t)nv; g->entries = ne;
    g->row.resize(nv + 1); g->col.resize(ne); g->w.resize(ne); g->orig.resize(nv); g->valid.resize(nv + 1);
    int64_t vo = 0, eo = 0;
    for (int i = 0; i < count; ++i) {
        const mg_graph *h = graphs[i];
        for (int32_t v = 0; v < h->nv; ++v) g->row[vo + v] = (int32_t)(eo + h->row[v]);
        for (int64_t e = 0; e < h->ne; ++e) { g->col[eo + e] = (int32_t)(vo + h->col[e]); g->w[eo + e] = h->w[e]; }
        vo += h->nv; eo += h->ne;
    }
    g->row[nv] = (int32_t)ne;
    for (int64_t v = 0; v < nv; ++v) { g->orig[v] = (int32_t)v; g->valid[v] = (int32_t)v; }
    g->valid[nv] = (int32_t)nv;
    g->start_p = 0;
    g->flags = graphs[0]->flags;
    g->symmetric = true;
    for (int i = 0; i < count; ++i) g->symmetric = g->symmetric && graphs[i]->symmetric;
    *out = g;
    return MG_OK;
}

extern "C" int mg_graph_write_mtx(const mg_graph *g, const char *path) {
    if (!g || !path) return mg_fail(MG_ERR_ARG, "null argument");
    FILE *fp = fopen(path, "w");
    if (!fp) return mg_fail(MG_ERR_IO, "cannot write %s", path);
    int64_t nnz = 0;
    for (int32_t u = 0; u < g->nv; ++u)
        for (int64_t p = g->off(u); p < g->off(u + 1); ++p) if (g->col[p] <= u) ++nnz;
    fprintf(fp, "%%%%MatrixMarket matrix coordinate integer symmetric\n%% written by multigraph-amd\n");
    fprintf(fp, "%d %d %lld\n", g->nv, g->nv, (long long)nnz);
    std::string line;
    for (int32_t u = 0; u < g->nv; ++u)
        for (int64_t p = g->off(u); p < g->off(u + 1); ++p)
            if (g->col[p] <= u) fprintf(fp, "%d %d %d\n", u + 1, g->col[p] + 1, g->w[p]);
    if (fclose(fp) != 0) return mg_fail(MG_ERR_IO, "write failed on %s", path);
    return MG_OK;
}

extern "C" int mg_graph_info(const mg_graph *g, int32_t *nv, int64_t *ne, int32_t *start_p) {
    if (!g) return mg_fail(MG_ERR_ARG, "null graph");
    if (nv) *nv = g->nv;
    if (ne) *ne = g->ne;
    if (start_p) *start_p = g->start_p;
    return MG_OK;
}

extern "C" int mg_graph_symmetric(const mg_graph *g, int *symmetric) {
    if (!g || !symmetric) return mg_fail(MG_ERR_ARG, "null argument");
    *symmetric = g->symmetric ? 1 : 0;
    return MG_OK;
}

extern "C" int mg_graph_csr(const mg_graph *g, const int32_t **row, const int32_t **col, const int32_t **weight) {
    if (!g) return mg_fail(MG_ERR_ARG, "null graph");
    if (row && g->wide()) return mg_fail(MG_ERR_LIMIT, "%lld edges: the row offsets are 64-bit (mg_graph_csr64)", (long long)g->ne);
    if (row) *row = g->row.data();
    if (col) *col = g->col.data();
    if (weight) *weight = g->w.data();
    return MG_OK;
}

extern "C" int mg_graph_csr64(const mg_graph *g, int64_t *row) {
    if (!g || !row) return mg_fail(MG_ERR_ARG, "null argument");
    for (int64_t v = 0; v <= g->nv; ++v) row[v] = g->off(v);
    return MG_OK;
}

extern "C" int mg_graph_orig_ids(const mg_graph *g, const int32_t **orig) {
    if (!g || !orig) return mg_fail(MG_ERR_ARG, "null argument");
    *orig = g->orig.data();
    return MG_OK;
}

extern "C" int mg_graph_remap(const mg_graph *g, int orig_id, int32_t *compact_id) {
    if (!g || !compact_id) return mg_fail(MG_ERR_ARG, "null argument");
    if (orig_id < 0 || orig_id >= g->orig_nv) return mg_fail(MG_ERR_ARG, "vertex %d outside [0,%d)", orig_id, g->orig_nv);
    *compact_id = g->valid[orig_id];
    return MG_OK;
}

extern "C" int mg_graph_max_degree_vertex(const mg_graph *g, int32_t *v) {
    if (!g || !v) return mg_fail(MG_ERR_ARG, "null argument");
    int32_t best = 0;
    int64_t bd = -1;
    for (int32_t u = 0; u < g->nv; ++u) {
        const int64_t d = g->off(u + 1) - g->off(u);
        if (d > bd) { bd = d; best = u; }
    }
    *v = best;
    return MG_OK;
}

extern "C" void mg_graph_free(mg_graph *g) { delete g; }
