/* oracle.c -- CPU restatement of the reference's BFS and PageRank algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it; the product
 * library (graph_analytics_applesilicon_amd) never calls it.
 *
 * Plain C99, serial, no dependencies. Each function names the reference code
 * it restates and the checks that pin it:
 *   or_bfs_serial      src/cpu/bfs_cpu.cpp:27-45. Pinned by closed forms (path,
 *                      star, grid corner, two components, empty), by the
 *                      compiled reference (oracle/_ref) on every suite graph and
 *                      source, by the committed golden fixture, and by the BFS
 *                      invariants (|d(u)-d(v)| <= 1 on every edge, parent at d-1).
 *   or_bfs_do_levels   the direction decisions of src/cpu/bfs_cpu.cpp:227-283
 *                      (serial; decisions depend only on counts). Pinned against
 *                      the compiled reference's per-level (direction, frontier)
 *                      trace and the committed fixture of it.
 *   or_pagerank        src/cpu/pagerank_cpu.cpp:20-53 (FP64 Jacobi, dangling
 *                      mass spread over all vertices). Pinned by the closed forms
 *                      of tests/test_main.cpp:630-652 (2-vertex dangling
 *                      0.5/1.425, 3-cycle, complete, empty), by sum == 1, by the
 *                      compiled reference, and by networkx.pagerank.
 *   or_ppr             personalized PageRank, the FP64 definition used for the
 *                      GPU block: restart and dangling mass both go to the seed
 *                      (networkx.pagerank with personalization={s:1}). Pinned by
 *                      networkx, by the isolated-seed closed form r = e_s, and by
 *                      the 2-vertex closed form worked out in the tests.
 *   or_check_ranks     src/cpu/pagerank_cpu.hpp:37-60 (pure arithmetic). Pinned
 *                      against the compiled reference checkRanks (bit-equal L1,
 *                      sum and top-1000 error on tie-free inputs; within the
 *                      extremes the tie choice allows otherwise).
 *   or_mlp_*           FP64 forward/backward of the 5->32->32->3 ReLU MLP with
 *                      mean-squared-error loss (no reference counterpart). Pinned
 *                      by central finite differences in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_INF 0xFFFFFFFFu

/* ------------------------------------------------------------------ BFS */

void or_bfs_serial(uint64_t n, const uint32_t* rp, const uint32_t* ci, uint32_t src, uint32_t* dist) {
    uint32_t* q = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint64_t head = 0, tail = 0;
    for (uint64_t v = 0; v < n; ++v) dist[v] = OR_INF;
    dist[src] = 0;
    q[tail++] = src;
    while (head < tail) {
        uint32_t u = q[head++];
        uint32_t du = dist[u] + 1;
        for (uint32_t e = rp[u]; e < rp[u + 1]; ++e) {
            uint32_t w = ci[e];
            if (dist[w] == OR_INF) {
                dist[w] = du;
                q[tail++] = w;
            }
        }
    }
    free(q);
}

/* Level-synchronous BFS that records, per level, the direction the GAP rule
 * picks and the input frontier size, exactly as bfsDO's trace does:
 *   TD: to_check -= min(to_check, scout); scout = sum of discovered degrees.
 *   go BU when scout > to_check / alpha; stay BU while the frontier grows or
 *   is larger than n / beta; after a BU phase scout = 1.
 * dir_out[i] is 'T' or 'B'; frontier_out[i] the input frontier of level i;
 * edges_out[i] the adjacency entries examined (TD: frontier degrees; BU: entries
 * scanned by unvisited vertices up to their first parent, in increasing
 * vertex order). Returns the number of levels (may exceed cap; only cap are
 * written). */
uint64_t or_bfs_do_levels(uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci, uint32_t src, double alpha,
                          double beta, uint32_t* dist, char* dir_out, uint64_t* frontier_out, uint64_t* edges_out,
                          uint64_t cap) {
    uint32_t* front = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t* next = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint8_t* in_front = (uint8_t*)calloc(n ? n : 1, 1);
    uint64_t fsize = 1, levels = 0;
    uint32_t level = 0;
    uint64_t to_check = m, scout = rp[src + 1] - rp[src];
    int bottom_up = 0;
    for (uint64_t v = 0; v < n; ++v) dist[v] = OR_INF;
    dist[src] = 0;
    front[0] = src;
    while (fsize > 0) {
        uint64_t found = 0, examined = 0, disc_deg = 0;
        if (!bottom_up && (double)scout > (double)to_check / alpha) bottom_up = 1;
        if (!bottom_up) to_check -= (to_check < scout ? to_check : scout);
        if (bottom_up) {
            for (uint64_t i = 0; i < fsize; ++i) in_front[front[i]] = 1;
            for (uint64_t v = 0; v < n; ++v) {
                if (dist[v] != OR_INF) continue;
                for (uint32_t e = rp[v]; e < rp[v + 1]; ++e) {
                    ++examined;
                    if (in_front[ci[e]]) {
                        dist[v] = level + 1;
                        next[found++] = (uint32_t)v;
                        disc_deg += rp[v + 1] - rp[v];
                        break;
                    }
                }
            }
            for (uint64_t i = 0; i < fsize; ++i) in_front[front[i]] = 0;
        } else {
            for (uint64_t i = 0; i < fsize; ++i) {
                uint32_t u = front[i];
                examined += rp[u + 1] - rp[u];
                for (uint32_t e = rp[u]; e < rp[u + 1]; ++e) {
                    uint32_t w = ci[e];
                    if (dist[w] == OR_INF) {
                        dist[w] = level + 1;
                        next[found++] = w;
                        disc_deg += rp[w + 1] - rp[w];
                    }
                }
            }
        }
        if (levels < cap) {
            dir_out[levels] = bottom_up ? 'B' : 'T';
            frontier_out[levels] = fsize;
            if (edges_out) edges_out[levels] = examined;
        }
        ++levels;
        if (bottom_up) {
            if (!(found >= fsize || (double)found > (double)n / beta)) {
                bottom_up = 0;
                scout = 1;
            }
        } else {
            scout = disc_deg;
        }
        {
            uint32_t* t = front;
            front = next;
            next = t;
        }
        fsize = found;
        ++level;
    }
    free(front);
    free(next);
    free(in_front);
    return levels;
}

/* ------------------------------------------------------------- PageRank */

/* r_new[v] = (1-d)/n + d * (sum_{u in in(v)} r[u]/outdeg[u] + dangling/n). */
void or_pagerank(uint64_t n, const uint32_t* in_rp, const uint32_t* in_ci, const uint32_t* outdeg, double damping,
                 int iters, double tol, double* r, int* iters_out, double* delta_out) {
    double* contrib = (double*)malloc((n ? n : 1) * sizeof(double));
    double* next = (double*)malloc((n ? n : 1) * sizeof(double));
    int it, done = 0;
    double delta = 0;
    if (n == 0) {
        if (iters_out) *iters_out = 0;
        if (delta_out) *delta_out = 0;
        free(contrib);
        free(next);
        return;
    }
    for (uint64_t v = 0; v < n; ++v) r[v] = 1.0 / (double)n;
    for (it = 0; it < iters && !done; ++it) {
        double dangling = 0, add;
        for (uint64_t u = 0; u < n; ++u) {
            if (outdeg[u]) contrib[u] = r[u] / outdeg[u];
            else {
                contrib[u] = 0;
                dangling += r[u];
            }
        }
        add = (1.0 - damping) / (double)n + damping * dangling / (double)n;
        delta = 0;
        for (uint64_t v = 0; v < n; ++v) {
            double s = 0;
            for (uint32_t e = in_rp[v]; e < in_rp[v + 1]; ++e) s += contrib[in_ci[e]];
            next[v] = add + damping * s;
            delta += fabs(next[v] - r[v]);
        }
        memcpy(r, next, n * sizeof(double));
        if (tol > 0 && delta < tol) done = 1;
    }
    if (iters_out) *iters_out = it;
    if (delta_out) *delta_out = delta;
    free(contrib);
    free(next);
}

/* Personalized PageRank with a one-hot restart on `seed`:
 *   r_new[v] = (1-d)*[v==seed] + d*(sum_{u in in(v)} r[u]/outdeg[u] + dangling*[v==seed])
 * starting from r = e_seed. Total mass stays 1. */
void or_ppr(uint64_t n, const uint32_t* in_rp, const uint32_t* in_ci, const uint32_t* outdeg, double damping,
            int iters, double tol, uint32_t seed, double* r, int* iters_out, double* delta_out) {
    double* contrib = (double*)malloc((n ? n : 1) * sizeof(double));
    double* next = (double*)malloc((n ? n : 1) * sizeof(double));
    int it, done = 0;
    double delta = 0;
    for (uint64_t v = 0; v < n; ++v) r[v] = v == seed ? 1.0 : 0.0;
    for (it = 0; it < iters && !done && n > 0; ++it) {
        double dangling = 0;
        for (uint64_t u = 0; u < n; ++u) {
            if (outdeg[u]) contrib[u] = r[u] / outdeg[u];
            else {
                contrib[u] = 0;
                dangling += r[u];
            }
        }
        delta = 0;
        for (uint64_t v = 0; v < n; ++v) {
            double s = 0;
            for (uint32_t e = in_rp[v]; e < in_rp[v + 1]; ++e) s += contrib[in_ci[e]];
            next[v] = damping * s + (v == seed ? (1.0 - damping) + damping * dangling : 0.0);
            delta += fabs(next[v] - r[v]);
        }
        memcpy(r, next, n * sizeof(double));
        if (tol > 0 && delta < tol) done = 1;
    }
    if (iters_out) *iters_out = it;
    if (delta_out) *delta_out = delta;
    free(contrib);
    free(next);
}

static const double* g_sort_ref;
static int by_ref_desc(const void* a, const void* b) {
    const uint32_t i = *(const uint32_t*)a, j = *(const uint32_t*)b;
    double x = g_sort_ref[i], y = g_sort_ref[j];
    if (x > y) return -1;
    if (x < y) return 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}

/* L1 distance, max relative error over the reference's 1000 largest entries,
 * and the sum of `got` (FP32 values widened to double). Equal reference
 * values that straddle rank 1000 are taken in id order (std::partial_sort in
 * the reference leaves that choice unspecified). */
void or_check_ranks(uint64_t n, const double* ref, const float* got, double* l1, double* max_rel_top, double* sum) {
    double a = 0, s = 0, mx = 0;
    uint64_t k = n < 1000 ? n : 1000;
    uint32_t* idx = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint64_t v = 0; v < n; ++v) {
        a += fabs(ref[v] - (double)got[v]);
        s += (double)got[v];
        idx[v] = (uint32_t)v;
    }
    g_sort_ref = ref;
    qsort(idx, n, sizeof(uint32_t), by_ref_desc);
    for (uint64_t i = 0; i < k; ++i) {
        double e = fabs(ref[idx[i]] - (double)got[idx[i]]) / ref[idx[i]];
        if (e > mx) mx = e;
    }
    free(idx);
    *l1 = a;
    *max_rel_top = mx;
    *sum = s;
}

/* ------------------------------------------------------------------ MLP
 * Parameter layout (1,347 values), row-major:
 *   W1[32][5] b1[32] W2[32][32] b2[32] W3[3][32] b3[3]
 * h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3.
 * Loss = (1 / (N*3)) * sum (y - t)^2 over the entries with mask != 0
 * (mask NULL = all), normalised by the count of masked entries. */
#define OR_IN 5
#define OR_H 32
#define OR_OUT 3
#define OR_NPARAM (OR_H * OR_IN + OR_H + OR_H * OR_H + OR_H + OR_OUT * OR_H + OR_OUT)

int or_mlp_nparam(void) { return OR_NPARAM; }

void or_mlp_forward(const double* p, const double* x, uint64_t count, double* y) {
    const double *W1 = p, *b1 = W1 + OR_H * OR_IN, *W2 = b1 + OR_H, *b2 = W2 + OR_H * OR_H, *W3 = b2 + OR_H,
                 *b3 = W3 + OR_OUT * OR_H;
    for (uint64_t s = 0; s < count; ++s) {
        double h1[OR_H], h2[OR_H];
        for (int j = 0; j < OR_H; ++j) {
            double a = b1[j];
            for (int i = 0; i < OR_IN; ++i) a += W1[j * OR_IN + i] * x[s * OR_IN + i];
            h1[j] = a > 0 ? a : 0;
        }
        for (int j = 0; j < OR_H; ++j) {
            double a = b2[j];
            for (int i = 0; i < OR_H; ++i) a += W2[j * OR_H + i] * h1[i];
            h2[j] = a > 0 ? a : 0;
        }
        for (int k = 0; k < OR_OUT; ++k) {
            double a = b3[k];
            for (int i = 0; i < OR_H; ++i) a += W3[k * OR_H + i] * h2[i];
            y[s * OR_OUT + k] = a;
        }
    }
}

/* Returns the loss; grad (OR_NPARAM doubles) receives dLoss/dparams. */
double or_mlp_loss_grad(const double* p, const double* x, const double* t, const uint8_t* mask, uint64_t count,
                        double* grad) {
    const double *W1 = p, *b1 = W1 + OR_H * OR_IN, *W2 = b1 + OR_H, *b2 = W2 + OR_H * OR_H, *W3 = b2 + OR_H,
                 *b3 = W3 + OR_OUT * OR_H;
    double *gW1 = grad, *gb1 = gW1 + OR_H * OR_IN, *gW2 = gb1 + OR_H, *gb2 = gW2 + OR_H * OR_H, *gW3 = gb2 + OR_H,
           *gb3 = gW3 + OR_OUT * OR_H;
    double loss = 0, norm;
    uint64_t used = 0;
    memset(grad, 0, OR_NPARAM * sizeof(double));
    for (uint64_t s = 0; s < count * OR_OUT; ++s) used += mask ? (mask[s] != 0) : 1;
    norm = used ? 1.0 / (double)used : 0.0;
    for (uint64_t s = 0; s < count; ++s) {
        double a1[OR_H], h1[OR_H], a2[OR_H], h2[OR_H], y[OR_OUT], dy[OR_OUT], dh2[OR_H], dh1[OR_H];
        for (int j = 0; j < OR_H; ++j) {
            double a = b1[j];
            for (int i = 0; i < OR_IN; ++i) a += W1[j * OR_IN + i] * x[s * OR_IN + i];
            a1[j] = a;
            h1[j] = a > 0 ? a : 0;
        }
        for (int j = 0; j < OR_H; ++j) {
            double a = b2[j];
            for (int i = 0; i < OR_H; ++i) a += W2[j * OR_H + i] * h1[i];
            a2[j] = a;
            h2[j] = a > 0 ? a : 0;
        }
        for (int k = 0; k < OR_OUT; ++k) {
            double a = b3[k];
            int on = mask ? mask[s * OR_OUT + k] != 0 : 1;
            for (int i = 0; i < OR_H; ++i) a += W3[k * OR_H + i] * h2[i];
            y[k] = a;
            dy[k] = on ? 2.0 * (a - t[s * OR_OUT + k]) * norm : 0.0;
            if (on) loss += (a - t[s * OR_OUT + k]) * (a - t[s * OR_OUT + k]) * norm;
        }
        for (int i = 0; i < OR_H; ++i) dh2[i] = 0;
        for (int k = 0; k < OR_OUT; ++k) {
            gb3[k] += dy[k];
            for (int i = 0; i < OR_H; ++i) {
                gW3[k * OR_H + i] += dy[k] * h2[i];
                dh2[i] += dy[k] * W3[k * OR_H + i];
            }
        }
        for (int i = 0; i < OR_H; ++i) dh1[i] = 0;
        for (int j = 0; j < OR_H; ++j) {
            double da = a2[j] > 0 ? dh2[j] : 0.0;
            gb2[j] += da;
            for (int i = 0; i < OR_H; ++i) {
                gW2[j * OR_H + i] += da * h1[i];
                dh1[i] += da * W2[j * OR_H + i];
            }
        }
        for (int j = 0; j < OR_H; ++j) {
            double da = a1[j] > 0 ? dh1[j] : 0.0;
            gb1[j] += da;
            for (int i = 0; i < OR_IN; ++i) gW1[j * OR_IN + i] += da * x[s * OR_IN + i];
        }
        (void)y;
    }
    return loss;
}
