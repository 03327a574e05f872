/*
 * oracle.c — CPU restatement of the reference's own CPU path (code/validate.cu) and
 * of the algorithms' defining rules (code/user_function.h, code/common.h).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker (or as
 * the timed CPU baseline); the product path never calls it.
 *
 * Every function works on a CSR (row[nv+1], col[ne], optional weight[ne]) in the
 * reader's compacted id space, exactly what validate.cu receives.
 *
 * Pinning (what checks each function independently of this file):
 *   orc_bfs            scipy.sparse.csgraph.shortest_path(unweighted), grid Manhattan
 *                      distances, path/star closed forms, |L(u)-L(v)| <= 1 invariant
 *   orc_sssp_bf        orc_sssp_dijkstra (textbook), scipy dijkstra, unit weights == BFS,
 *                      triangle inequality + tight predecessor invariant
 *   orc_cc_labelprop   orc_cc_unionfind (textbook), scipy connected_components
 *   orc_canon          hand-computed first-appearance relabelling
 *   orc_bc_fp64        networkx single-source dependencies, grid sigma = C(dx+dy,dx),
 *                      path / star / complete-graph closed forms
 *   orc_pr_f32/_fp64   regular graphs converge to 1, float rule vs double rule
 *                      agree to float precision, hand-iterated 3-vertex example
 *   orc_*_multi,       OpenMP drivers of the functions above for the full-size parity
 *   orc_pr_fp64_pull   tests: bit-identical to the single-thread functions (checked on
 *                      small graphs), so they inherit those functions' pins
 *   orc_bfs64          networkx shortest-path lengths with the rows cast to int64
 *   orc_sssp_certificate64  holds for networkx Dijkstra distances; fails for a raised
 *                      or lowered distance, a dropped or invented reached vertex, a
 *                      nonzero source (planted on small graphs)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define ORC_INF 0x64646464 /* IUNUSED, common.h:37-39 */

/* validate.cu:106-127 — FIFO queue BFS from `src`; unreached = ORC_INF. */
int orc_bfs(int nv, const int32_t *row, const int32_t *col, int src, int32_t *level) {
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!q) return -1;
    for (int i = 0; i < nv; ++i) level[i] = ORC_INF;
    int head = 0, tail = 0;
    if (nv > 0) { level[src] = 0; q[tail++] = src; }
    while (head < tail) {
        int i = q[head++];
        for (int32_t j = row[i]; j < row[i + 1]; ++j) {
            int k = col[j];
            if (level[k] == ORC_INF) { level[k] = level[i] + 1; q[tail++] = k; }
        }
    }
    free(q);
    return tail;
}

/* validate.cu:131-146 — in-place Bellman-Ford sweeps over all edges to a fixpoint.
 * Returns the number of sweeps (including the final unchanged one). */
int orc_sssp_bf(int nv, const int32_t *row, const int32_t *col, const int32_t *w, int src, int32_t *dist) {
    for (int i = 0; i < nv; ++i) dist[i] = ORC_INF;
    if (nv == 0) return 0;
    dist[src] = 0;
    int sweeps = 0;
    for (;;) {
        int conv = 1;
        ++sweeps;
        for (int i = 0; i < nv; ++i) {
            int32_t di = dist[i];
            for (int32_t j = row[i]; j < row[i + 1]; ++j) {
                int32_t nd = di + w[j];
                if (nd < dist[col[j]]) { conv = 0; dist[col[j]] = nd; }
            }
            /* the reference re-reads c_vv[i] per edge; a self-loop cannot lower it
             * (weights are positive), so reading it once per row is the same rule */
        }
        if (conv) break;
    }
    return sweeps;
}

/* Textbook Dijkstra with a binary heap (independent cross-check of orc_sssp_bf). */
int orc_sssp_dijkstra(int nv, const int32_t *row, const int32_t *col, const int32_t *w, int src, int32_t *dist) {
    for (int i = 0; i < nv; ++i) dist[i] = ORC_INF;
    if (nv == 0) return 0;
    int64_t cap = (int64_t)nv + 1;  /* grows by doubling: pushes can reach ne */
    int64_t *hk = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
    if (!hk) return -1;
    int64_t hn = 0;
#define HPUSH(key)                                                   \
    do {                                                             \
        int64_t _i = hn++;                                           \
        hk[_i] = (key);                                              \
        while (_i > 0 && hk[(_i - 1) / 2] > hk[_i]) {                \
            int64_t _p = (_i - 1) / 2, _t = hk[_p];                  \
            hk[_p] = hk[_i]; hk[_i] = _t; _i = _p;                   \
        }                                                            \
    } while (0)
    dist[src] = 0;
    HPUSH(((int64_t)0 << 32) | (uint32_t)src);
    int settled = 0;
    while (hn > 0) {
        int64_t top = hk[0];
        hk[0] = hk[--hn];
        int64_t i = 0;
        for (;;) {
            int64_t l = 2 * i + 1, r = l + 1, m = i;
            if (l < hn && hk[l] < hk[m]) m = l;
            if (r < hn && hk[r] < hk[m]) m = r;
            if (m == i) break;
            int64_t t = hk[m]; hk[m] = hk[i]; hk[i] = t; i = m;
        }
        int32_t d = (int32_t)(top >> 32), u = (int32_t)(top & 0xffffffff);
        if (d != dist[u]) continue;
        ++settled;
        for (int32_t j = row[u]; j < row[u + 1]; ++j) {
            int32_t v = col[j], nd = d + w[j];
            if (nd < dist[v]) {
                dist[v] = nd;
                if (hn >= cap) {
                    cap *= 2;
                    int64_t *nk = (int64_t *)realloc(hk, sizeof(int64_t) * (size_t)cap);
                    if (!nk) { free(hk); return -1; }
                    hk = nk;
                }
                HPUSH(((int64_t)nd << 32) | (uint32_t)v);
            }
        }
    }
#undef HPUSH
    free(hk);
    return settled;
}

/* validate.cu:234-242 / :263-273 — relabel by first appearance in id order. */
int orc_canon(int nv, const int32_t *label, int32_t *out) {
    int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!adj) return -1;
    for (int i = 0; i < nv; ++i) adj[i] = -1;
    int k = 0;
    for (int i = 0; i < nv; ++i) {
        int32_t l = label[i];
        if (l < 0 || l >= nv) { free(adj); return -2; }
        if (adj[l] < 0) adj[l] = k++;
    }
    for (int i = 0; i < nv; ++i) out[i] = adj[label[i]];
    free(adj);
    return k;
}

/* validate.cu:245-261 — label propagation sweeps (both endpoints of every stored
 * edge take the min) to a fixpoint, then canonical relabelling (:263-273).
 * Returns the number of components; `sweeps` (optional) receives the sweep count. */
int orc_cc_labelprop(int nv, const int32_t *row, const int32_t *col, int32_t *label, int *sweeps) {
    int32_t *c = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!c) return -1;
    for (int i = 0; i < nv; ++i) c[i] = i;
    int s = 0;
    for (;;) {
        int pflag = 0;
        ++s;
        for (int i = 0; i < nv; ++i) {
            for (int32_t j = row[i]; j < row[i + 1]; ++j) {
                int32_t k = col[j];
                if (c[i] != c[k]) {
                    int32_t m = c[i] < c[k] ? c[i] : c[k];
                    c[i] = m; c[k] = m;
                    pflag = 1;
                }
            }
        }
        if (!pflag) break;
    }
    if (sweeps) *sweeps = s;
    int k = orc_canon(nv, c, label);
    free(c);
    return k;
}

/* Textbook union-find (path halving, union by smaller id).  label[v] = minimum
 * vertex id of v's component.  Independent cross-check of orc_cc_labelprop. */
int orc_cc_unionfind(int nv, const int32_t *row, const int32_t *col, int32_t *label) {
    for (int i = 0; i < nv; ++i) label[i] = i;
    for (int u = 0; u < nv; ++u) {
        for (int32_t j = row[u]; j < row[u + 1]; ++j) {
            int32_t a = u, b = col[j];
            while (label[a] != a) { label[a] = label[label[a]]; a = label[a]; }
            while (label[b] != b) { label[b] = label[label[b]]; b = label[b]; }
            if (a == b) continue;
            if (a < b) label[b] = a; else label[a] = b;
        }
    }
    int comps = 0;
    for (int v = 0; v < nv; ++v) {
        int32_t a = v;
        while (label[a] != a) a = label[a];
        label[v] = a;
        if (a == v) ++comps;
    }
    return comps;
}

/* validate.cu:149-189 — Brandes from `src`: forward FIFO BFS counting shortest
 * paths (sigma), backward over the queue in reverse accumulating dependencies
 * delta[i] = sum sigma[i]/sigma[k] * (1 + delta[k]) over successors k of i.
 * The reference computes sigma in int and delta in float; here both are FP64, the
 * precision the GPU result is judged against. Returns the number of reached vertices. */
int orc_bc_fp64(int nv, const int32_t *row, const int32_t *col, int src, int32_t *level, double *sigma,
                double *delta) {
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!q) return -1;
    for (int i = 0; i < nv; ++i) { level[i] = ORC_INF; sigma[i] = 0.0; delta[i] = 0.0; }
    if (nv == 0) { free(q); return 0; }
    int head = 0, tail = 0;
    level[src] = 0; sigma[src] = 1.0; q[tail++] = src;
    while (head < tail) {
        int i = q[head++];
        for (int32_t j = row[i]; j < row[i + 1]; ++j) {
            int k = col[j];
            if (level[k] == ORC_INF) { level[k] = level[i] + 1; q[tail++] = k; }
            if (level[i] + 1 == level[k]) sigma[k] += sigma[i];
        }
    }
    /* Dependencies in successor form: delta[i] sums over out-edges i->k one level deeper.
     * validate.cu:170-185 pushes from k to the predecessors found in k's own row, which
     * is the same sum when every edge is stored both ways; on a directed CSR only the
     * successor form is Brandes' recurrence, so the oracle uses it throughout. */
    for (int t = tail - 1; t >= 0; --t) {
        int i = q[t];
        double acc = 0.0;
        for (int32_t j = row[i]; j < row[i + 1]; ++j) {
            int k = col[j];
            if (level[k] == level[i] + 1) acc += (sigma[i] / sigma[k]) * (1.0 + delta[k]);
        }
        delta[i] = acc;
    }
    free(q);
    return tail;
}

/* validate.cu:34-91 — PageRank with the reference's unnormalised rule
 * r' = (1-d) + d * sum_{u->v} r_u/deg_u, r_0 = 0, pushed in row order in float.
 * variant 0 = PR_T (DATA_DRIVEN_PR): r = r' only where |r'-r| > tol; stop when none.
 * variant 1 = PR_D (TOPOLOGY_DRIVEN_PR): r = r' always; a vertex stays active while
 * |r'-r| > tol*r, and the run stops when no active vertex changed.
 * max_iter caps the iterations (0: unbounded, as the reference); a negative max_iter
 * runs exactly -max_iter iterations whatever the stopping rule says.  Returns iterations. */
int orc_pr_f32(int nv, const int32_t *row, const int32_t *col, int variant, float damping, float tol, int max_iter,
               float *r) {
    float *nr = (float *)malloc(sizeof(float) * (size_t)(nv > 0 ? nv : 1));
    int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!nr || !mark) { free(nr); free(mark); return -1; }
    for (int i = 0; i < nv; ++i) { r[i] = 0.0f; mark[i] = 1; }
    int it = 0;
    for (;;) {
        int flag = 0;
        ++it;
        memset(nr, 0, sizeof(float) * (size_t)nv);
        for (int i = 0; i < nv; ++i) {
            int deg = row[i + 1] - row[i];
            for (int32_t j = row[i]; j < row[i + 1]; ++j) nr[col[j]] += r[i] / (float)deg;
        }
        for (int i = 0; i < nv; ++i) {
            float f1 = (1.0f - damping) + nr[i] * damping;
            float f2 = r[i];
            if (variant == 0) {
                if (fabsf(f1 - f2) > tol) { r[i] = f1; flag = 1; }
            } else {
                if (mark[i] == 1 && fabsf(f1 - f2) > tol * f2) { r[i] = f1; flag = 1; }
                else { r[i] = f1; mark[i] = 0; }
            }
        }
        if (max_iter < 0) { if (it >= -max_iter) break; continue; }
        if (!flag) break;
        if (max_iter > 0 && it >= max_iter) break;
    }
    free(nr); free(mark);
    return it;
}

/* The same rules as orc_pr_f32 evaluated entirely in FP64: the precision oracle for
 * the GPU's FP32 ranks (error reported as L1 and max-abs against this). */
int orc_pr_fp64(int nv, const int32_t *row, const int32_t *col, int variant, double damping, double tol, int max_iter,
                double *r) {
    double *nr = (double *)malloc(sizeof(double) * (size_t)(nv > 0 ? nv : 1));
    double *c = (double *)malloc(sizeof(double) * (size_t)(nv > 0 ? nv : 1));
    int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!nr || !mark || !c) { free(nr); free(mark); free(c); return -1; }
    for (int i = 0; i < nv; ++i) { r[i] = 0.0; mark[i] = 1; }
    int it = 0;
    for (;;) {
        int flag = 0;
        ++it;
        for (int i = 0; i < nv; ++i) {
            int deg = row[i + 1] - row[i];
            c[i] = deg > 0 ? r[i] / (double)deg : 0.0;
            nr[i] = 0.0;
        }
        for (int i = 0; i < nv; ++i)
            for (int32_t j = row[i]; j < row[i + 1]; ++j) nr[col[j]] += c[i];
        for (int i = 0; i < nv; ++i) {
            double f1 = (1.0 - damping) + nr[i] * damping;
            double f2 = r[i];
            if (variant == 0) {
                if (fabs(f1 - f2) > tol) { r[i] = f1; flag = 1; }
            } else {
                if (mark[i] == 1 && fabs(f1 - f2) > tol * f2) { r[i] = f1; flag = 1; }
                else { r[i] = f1; mark[i] = 0; }
            }
        }
        if (max_iter < 0) { if (it >= -max_iter) break; continue; }
        if (!flag) break;
        if (max_iter > 0 && it >= max_iter) break;
    }
    free(nr); free(mark); free(c);
    return it;
}

/* ---- OpenMP drivers for the full-size parity tests --------------------------------
 * Each is the single-thread function above, run on several threads, with results that
 * are bit-identical to it (pinned so on small graphs by the tests). */

int orc_threads(void) { return omp_get_max_threads(); }

/* row s of level[n][nv] = orc_bfs(src[s]) */
int orc_bfs_multi(int nv, const int32_t *row, const int32_t *col, int n, const int32_t *src, int32_t *level) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int s = 0; s < n; ++s)
        if (orc_bfs(nv, row, col, src[s], level + (size_t)s * nv) < 0) bad = 1;
    return bad ? -1 : 0;
}

/* row s of dist[n][nv] = orc_sssp_dijkstra(src[s]) */
int orc_sssp_dijkstra_multi(int nv, const int32_t *row, const int32_t *col, const int32_t *w, int n,
                            const int32_t *src, int32_t *dist) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int s = 0; s < n; ++s)
        if (orc_sssp_dijkstra(nv, row, col, w, src[s], dist + (size_t)s * nv) < 0) bad = 1;
    return bad ? -1 : 0;
}

/* rows s of level / sigma [n][nv] = orc_bc_fp64(src[s]); bc[v] = sum over s of delta_s[v]
 * with each source's own entry (delta_s[src[s]]) left out, summed in source order */
int orc_bc_fp64_multi(int nv, const int32_t *row, const int32_t *col, int n, const int32_t *src, int32_t *level,
                      double *sigma, double *bc) {
    double *dl = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * (size_t)(nv > 0 ? nv : 1));
    if (!dl) return -1;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int s = 0; s < n; ++s) {
        double *d = dl + (size_t)s * nv;
        if (orc_bc_fp64(nv, row, col, src[s], level + (size_t)s * nv, sigma + (size_t)s * nv, d) < 0) bad = 1;
        else d[src[s]] = 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int v = 0; v < nv; ++v) {
        double a = 0.0;
        for (int s = 0; s < n; ++s) a += dl[(size_t)s * nv + v];
        bc[v] = a;
    }
    free(dl);
    return bad ? -1 : 0;
}

/* orc_pr_fp64 in pull form over the transposed CSR, vertices on OpenMP threads.  The
 * transpose is built by a counting sort that scans sources in increasing order, so every
 * destination adds its contributions in the same order as orc_pr_fp64's push loop does:
 * the ranks are bit-identical to it. */
int orc_pr_fp64_pull(int nv, const int32_t *row, const int32_t *col, int variant, double damping, double tol,
                     int max_iter, double *r) {
    const size_t n1 = (size_t)(nv > 0 ? nv : 1);
    const int64_t ne = nv > 0 ? row[nv] : 0;
    int64_t *trow = (int64_t *)calloc(n1 + 1, sizeof(int64_t));
    int32_t *tcol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * n1);
    double *nr = (double *)malloc(sizeof(double) * n1);
    double *c = (double *)malloc(sizeof(double) * n1);
    int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * n1);
    int rc = -1;
    if (!trow || !tcol || !cur || !nr || !c || !mark) goto out;
    for (int64_t j = 0; j < ne; ++j) ++trow[col[j] + 1];
    for (int v = 0; v < nv; ++v) trow[v + 1] += trow[v];
    for (int v = 0; v < nv; ++v) cur[v] = trow[v];
    for (int u = 0; u < nv; ++u)
        for (int32_t j = row[u]; j < row[u + 1]; ++j) tcol[cur[col[j]]++] = u;
    for (int i = 0; i < nv; ++i) { r[i] = 0.0; mark[i] = 1; }
    int it = 0;
    for (;;) {
        int flag = 0;
        ++it;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nv; ++i) {
            int deg = row[i + 1] - row[i];
            c[i] = deg > 0 ? r[i] / (double)deg : 0.0;
        }
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : flag)
        for (int i = 0; i < nv; ++i) {
            double a = 0.0;
            for (int64_t j = trow[i]; j < trow[i + 1]; ++j) a += c[tcol[j]];
            double f1 = (1.0 - damping) + a * damping;
            double f2 = r[i];
            if (variant == 0) {
                if (fabs(f1 - f2) > tol) { r[i] = f1; flag = 1; }
            } else {
                if (mark[i] == 1 && fabs(f1 - f2) > tol * f2) { r[i] = f1; flag = 1; }
                else { r[i] = f1; mark[i] = 0; }
            }
        }
        if (max_iter < 0) { if (it >= -max_iter) break; continue; }
        if (!flag) break;
        if (max_iter > 0 && it >= max_iter) break;
    }
    rc = it;
out:
    free(trow); free(tcol); free(cur); free(nr); free(c); free(mark);
    return rc;
}

/* ---- graphs past 2^31 - 1 edges (64-bit row offsets) ---------------------------- */

/* validate.cu:106-127 with 64-bit row offsets: the same FIFO queue BFS as orc_bfs. */
int orc_bfs64(int nv, const int64_t *row, const int32_t *col, int src, int32_t *level) {
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    if (!q) return -1;
    for (int i = 0; i < nv; ++i) level[i] = ORC_INF;
    int head = 0, tail = 0;
    if (nv > 0) { level[src] = 0; q[tail++] = src; }
    while (head < tail) {
        int i = q[head++];
        for (int64_t j = row[i]; j < row[i + 1]; ++j) {
            int k = col[j];
            if (level[k] == ORC_INF) { level[k] = level[i] + 1; q[tail++] = k; }
        }
    }
    free(q);
    return tail;
}

/* Certificate of single-source shortest distances (positive integer weights): they are
 * exact iff dist[src] = 0, no edge can lower them (dist[v] <= dist[u] + w(u,v) for
 * every edge with dist[u] finite), every other reached vertex has an in-edge that is
 * tight (dist[v] = dist[u] + w, u reached), and the reached vertices are exactly those
 * the BFS levels `level` reach.  With a symmetric CSR a row's edges are its in-edges.
 * Returns 0 when it holds; else 1 (source), 2 (an edge lowers a distance), 3 (a
 * reached vertex has no tight in-edge) or 4 (reached sets differ). */
int orc_sssp_certificate64(int nv, const int64_t *row, const int32_t *col, const int32_t *w, int src,
                           const int32_t *dist, const int32_t *level) {
    if (nv == 0) return 0;
    if (dist[src] != 0) return 1;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : bad)
    for (int v = 0; v < nv; ++v) {
        if ((dist[v] == ORC_INF) != (level[v] == ORC_INF)) { bad |= 8; continue; }
        int tight = v == src;
        for (int64_t j = row[v]; j < row[v + 1]; ++j) {
            const int u = col[j];
            if (dist[u] == ORC_INF) continue;
            const int64_t cand = (int64_t)dist[u] + w[j];
            if ((int64_t)dist[v] > cand) bad |= 2;
            if ((int64_t)dist[v] == cand) tight = 1;
        }
        if (dist[v] != ORC_INF && !tight) bad |= 4;
    }
    if (bad & 8) return 4;
    if (bad & 2) return 2;
    if (bad & 4) return 3;
    return 0;
}
