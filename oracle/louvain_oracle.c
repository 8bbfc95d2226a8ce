/*
 * louvain_oracle.c -- CPU restatement of the reference's GPU Louvain method.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline.  The product path never links, loads or calls this file.
 *
 * Semantics follow, in order (reference paths relative to the reference repo):
 *   modularityAlgorithms.h:14-81       level loop
 *   phaseOptimization.cuh:50-181       sweep loop, sweep-kind schedule
 *   phaseOptimizationSort.cu:28-158    fast sweep (sweep 0 of each level)
 *   phaseOptimizationSort.cu:161-379   sort sweep
 *   phaseOptimizationHash.cu:23-162 + hashmap.cuh:13-91  hash sweep (an
 *                                      ideal table: no drops, exact sums)
 *   phaseOptimizationUpdate.cu:14-222  move rule, totals, active marking
 *   operatorsDelta.h:7-45, operators.h:9-70, operatorsCommunity.h:7-79
 *   community.h:31-106                 modularity
 *   phaseAggregation.cuh:19-175        renumbering and contraction
 *   graph.cu:17-87, graph.cuh:46-99    .edge parsing and CSR build
 *
 * Parity pins (tests/test_oracle_*.py): closed-form singleton modularity of
 * the bundled fixture, brute-force delta-Q by full modularity recomputation
 * on small graphs, an independent numpy emulation of the reference's Thrust
 * pipeline (copy_if / sort / reduce_by_key / leftmost GetMaxValue), networkx
 * modularity, and invariants (totals, aggregation weight conservation).
 */
#include "louvain_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define NONE32 0xFFFFFFFFu

void lvo_default_params(lvo_params *p) {
    p->mode = LVO_HASH;
    p->gain = LVO_GAIN_EXACT;
    p->tie = LVO_TIE_LOWEST;
    p->chunk_emulation = 0;
    p->chunk_entries = 50000000ull;          /* STEP_ROUND, constants.h:18 */
    p->sweep_threshold = 0.001;              /* constants.h:15 */
    p->level_threshold = 0.001;
    p->max_sweeps = 1000000000ll;            /* EARLY_STOP_LIMIT, constants.h:16 */
    p->max_total_sweeps = 0;
    p->max_levels = 0;
    p->refine = 0;
}

uint64_t lvo_fnv1a_u32(const uint32_t *a, uint64_t n) {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t x = a[i];
        for (int b = 0; b < 4; ++b) {
            h ^= (x >> (8 * b)) & 0xFFu;
            h *= 1099511628211ull;
        }
    }
    return h;
}

/* ---------------------------------------------------------------- sorting */

/* Stable LSD radix sort of (key, val) pairs by key, 16-bit digits over the
 * bits that are set anywhere in the keys. */
static int radix_sort_pairs(uint64_t *key, uint32_t *val, uint64_t n) {
    if (n < 2) return LVO_OK;
    uint64_t all = 0;
    for (uint64_t i = 0; i < n; ++i) all |= key[i];
    uint64_t *k2 = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *v2 = val ? (uint32_t *)malloc(n * sizeof(uint32_t)) : NULL;
    uint64_t *cnt = (uint64_t *)malloc(65536 * sizeof(uint64_t));
    if (!k2 || (val && !v2) || !cnt) { free(k2); free(v2); free(cnt); return LVO_ERR_OOM; }
    uint64_t *ks = key, *kd = k2;
    uint32_t *vs = val, *vd = v2;
    for (int shift = 0; shift < 64; shift += 16) {
        if (((all >> shift) & 0xFFFFull) == 0) continue;
        memset(cnt, 0, 65536 * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; ++i) cnt[(ks[i] >> shift) & 0xFFFF]++;
        uint64_t s = 0;
        for (int d = 0; d < 65536; ++d) { uint64_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t pos = cnt[(ks[i] >> shift) & 0xFFFF]++;
            kd[pos] = ks[i];
            if (val) vd[pos] = vs[i];
        }
        uint64_t *tk = ks; ks = kd; kd = tk;
        uint32_t *tv = vs; vs = vd; vd = tv;
    }
    /* an odd number of passes leaves the result in the scratch arrays */
    if (ks != key) {
        memcpy(key, ks, n * sizeof(uint64_t));
        if (val) memcpy(val, vs, n * sizeof(uint32_t));
    }
    free(cnt);
    free(k2);
    free(v2);
    return LVO_OK;
}

/* ------------------------------------------------------------------ graph */

void lvo_graph_free(lvo_graph *g) {
    if (!g) return;
    free(g->row); free(g->col); free(g->w); free(g->k);
    free(g);
}

static lvo_graph *graph_alloc(uint32_t n, uint64_t E) {
    lvo_graph *g = (lvo_graph *)calloc(1, sizeof(lvo_graph));
    if (!g) return NULL;
    g->n = n; g->E = E;
    g->row = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    g->col = (uint32_t *)malloc((E ? E : 1) * sizeof(uint32_t));
    g->w = (uint32_t *)malloc((E ? E : 1) * sizeof(uint32_t));
    g->k = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    if (!g->row || !g->col || !g->w || !g->k) { lvo_graph_free(g); return NULL; }
    return g;
}

static void graph_finish(lvo_graph *g) {
    uint64_t m2 = 0;
    for (uint32_t v = 0; v < g->n; ++v) {
        uint64_t s = 0;
        for (uint64_t j = g->row[v]; j < g->row[v + 1]; ++j) s += g->w[j];
        g->k[v] = s;
        m2 += s;
    }
    g->m2 = m2;
}

/* graph.cu:17-87 and graph.cuh:46-99.  Pairs are normalised to (min,max)
 * (graph.cu:38-43) and deduplicated (std::set, graph.cu:26-43); both
 * directions are emitted (graph.cu:68-79); entries are sorted by (src,dst)
 * (graph.cuh:60-64) and shifted by min_id (graph.cuh:66-76); the degree is
 * the entry count (graph.cuh:79-98), generalised to the weight sum. */
int lvo_graph_from_edges(uint64_t n_edges, const int64_t *a, const int64_t *b,
                         const uint32_t *w, lvo_graph **out) {
    *out = NULL;
    if (n_edges == 0) return LVO_ERR_PARSE;     /* graph.cu:52-55 exits */
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (uint64_t i = 0; i < n_edges; ++i) {
        int64_t x = a[i] < b[i] ? a[i] : b[i];
        int64_t y = a[i] < b[i] ? b[i] : a[i];
        if (x < lo) lo = x;
        if (y > hi) hi = y;
    }
    if (hi - lo >= (int64_t)0xFFFFFFFFll) return LVO_ERR_RANGE;
    /* undirected keys (min-lo)<<32 | (max-lo), value = original index so the
     * first occurrence of a duplicate can be kept */
    uint64_t *key = (uint64_t *)malloc(n_edges * sizeof(uint64_t));
    uint32_t *idx = (uint32_t *)malloc(n_edges * sizeof(uint32_t));
    if (!key || !idx) { free(key); free(idx); return LVO_ERR_OOM; }
    if (n_edges > 0xFFFFFFFFull) { free(key); free(idx); return LVO_ERR_RANGE; }
    for (uint64_t i = 0; i < n_edges; ++i) {
        int64_t x = a[i] < b[i] ? a[i] : b[i];
        int64_t y = a[i] < b[i] ? b[i] : a[i];
        key[i] = ((uint64_t)(x - lo) << 32) | (uint64_t)(y - lo);
        idx[i] = (uint32_t)i;
    }
    int rc = radix_sort_pairs(key, idx, n_edges);   /* stable: first occurrence first */
    if (rc) { free(key); free(idx); return rc; }
    uint64_t u = 0;
    for (uint64_t i = 0; i < n_edges; ++i) {
        if (i > 0 && key[i] == key[i - 1]) continue;
        key[u] = key[i]; idx[u] = idx[i]; ++u;
    }
    /* node set must be dense: every id in [lo, hi] used (graph.cuh:66-76
     * indexes per-vertex arrays by id - min_id with n = |node set|) */
    uint64_t span = (uint64_t)(hi - lo) + 1;
    uint8_t *seen = (uint8_t *)calloc(span, 1);
    if (!seen) { free(key); free(idx); return LVO_ERR_OOM; }
    for (uint64_t i = 0; i < u; ++i) { seen[key[i] >> 32] = 1; seen[key[i] & 0xFFFFFFFFu] = 1; }
    for (uint64_t i = 0; i < span; ++i) if (!seen[i]) { free(seen); free(key); free(idx); return LVO_ERR_NOT_DENSE; }
    free(seen);
    uint32_t n = (uint32_t)span;
    uint64_t E = 2 * u;
    uint64_t *ek = (uint64_t *)malloc(E * sizeof(uint64_t));
    uint32_t *ew = (uint32_t *)malloc(E * sizeof(uint32_t));
    if (!ek || !ew) { free(ek); free(ew); free(key); free(idx); return LVO_ERR_OOM; }
    for (uint64_t i = 0; i < u; ++i) {
        uint64_t x = key[i] >> 32, y = key[i] & 0xFFFFFFFFu;
        uint32_t wi = w ? w[idx[i]] : 1u;
        ek[2 * i] = (x << 32) | y;     ew[2 * i] = wi;
        ek[2 * i + 1] = (y << 32) | x; ew[2 * i + 1] = wi;
    }
    free(key); free(idx);
    rc = radix_sort_pairs(ek, ew, E);
    if (rc) { free(ek); free(ew); return rc; }
    lvo_graph *g = graph_alloc(n, E);
    if (!g) { free(ek); free(ew); return LVO_ERR_OOM; }
    for (uint64_t j = 0; j < E; ++j) {
        uint32_t s = (uint32_t)(ek[j] >> 32);
        g->row[s + 1]++;
        g->col[j] = (uint32_t)(ek[j] & 0xFFFFFFFFu);
        g->w[j] = ew[j];
    }
    free(ek); free(ew);
    for (uint32_t v = 0; v < n; ++v) g->row[v + 1] += g->row[v];
    graph_finish(g);
    g->total_weight = (double)(g->m2 / 2);   /* graph.cu:81 total_weight = n_links */
    g->n_links = u;
    g->min_id = lo;
    *out = g;
    return LVO_OK;
}

int lvo_graph_from_csr(uint32_t n, uint64_t E, const uint64_t *row, const uint32_t *col,
                       const uint32_t *w, double total_weight, lvo_graph **out) {
    *out = NULL;
    if (row[n] != E) return LVO_ERR_ARG;
    lvo_graph *g = graph_alloc(n, E);
    if (!g) return LVO_ERR_OOM;
    memcpy(g->row, row, ((size_t)n + 1) * sizeof(uint64_t));
    memcpy(g->col, col, E * sizeof(uint32_t));
    for (uint64_t j = 0; j < E; ++j) {
        if (col[j] >= n) { lvo_graph_free(g); return LVO_ERR_ARG; }
        g->w[j] = w ? w[j] : 1u;
    }
    graph_finish(g);
    g->total_weight = total_weight > 0 ? total_weight : (double)(g->m2 / 2);
    g->n_links = E / 2;
    g->min_id = 0;
    *out = g;
    return LVO_OK;
}

/* Token reader with `std::istream >> int` behaviour: skip whitespace, optional
 * sign, digits; fails when no digit follows. */
static int read_int(const char **pp, const char *end, int64_t *out) {
    const char *p = *pp;
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    int neg = 0;
    if (p < end && (*p == '-' || *p == '+')) { neg = (*p == '-'); ++p; }
    if (p >= end || *p < '0' || *p > '9') return 0;
    int64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        v = v * 10 + (*p - '0');
        if (v > 0x7FFFFFFFll + 1) return 0;    /* beyond int: stream failure */
        ++p;
    }
    v = neg ? -v : v;
    if (v > 0x7FFFFFFFll || v < -0x80000000ll) return 0;
    *out = v;
    *pp = p;
    return 1;
}

/* graph.cu:17-87: skip `skip_line` lines, read pairs until the first failure. */
int lvo_graph_load_edge(const char *path, int weighted, int skip_line, lvo_graph **out) {
    *out = NULL;
    FILE *f = fopen(path, "rb");
    if (!f) return LVO_ERR_IO;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc(sz > 0 ? (size_t)sz : 1);
    if (!buf) { fclose(f); return LVO_ERR_OOM; }
    size_t got = fread(buf, 1, (size_t)sz, f);
    fclose(f);
    const char *p = buf, *end = buf + got;
    for (int i = 0; i < skip_line && p < end; ++i) {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
    uint64_t cap = 1024, m = 0;
    int64_t *a = (int64_t *)malloc(cap * sizeof(int64_t));
    int64_t *b = (int64_t *)malloc(cap * sizeof(int64_t));
    uint32_t *w = weighted ? (uint32_t *)malloc(cap * sizeof(uint32_t)) : NULL;
    if (!a || !b || (weighted && !w)) { free(buf); free(a); free(b); free(w); return LVO_ERR_OOM; }
    for (;;) {
        int64_t x, y, z = 1;
        const char *q = p;
        if (!read_int(&q, end, &x) || !read_int(&q, end, &y)) break;
        if (weighted && (!read_int(&q, end, &z) || z < 0)) break;
        p = q;
        if (m == cap) {
            cap *= 2;
            a = (int64_t *)realloc(a, cap * sizeof(int64_t));
            b = (int64_t *)realloc(b, cap * sizeof(int64_t));
            if (weighted) w = (uint32_t *)realloc(w, cap * sizeof(uint32_t));
            if (!a || !b || (weighted && !w)) { free(buf); free(a); free(b); free(w); return LVO_ERR_OOM; }
        }
        a[m] = x; b[m] = y;
        if (weighted) w[m] = (uint32_t)z;
        ++m;
    }
    free(buf);
    int rc = lvo_graph_from_edges(m, a, b, w, out);
    free(a); free(b); free(w);
    return rc;
}

uint32_t lvo_graph_n(const lvo_graph *g) { return g->n; }
uint64_t lvo_graph_E(const lvo_graph *g) { return g->E; }
double lvo_graph_total_weight(const lvo_graph *g) { return g->total_weight; }
int64_t lvo_graph_min_id(const lvo_graph *g) { return g->min_id; }
void lvo_graph_copy(const lvo_graph *g, uint64_t *row, uint32_t *col, uint32_t *w, uint64_t *k) {
    if (row) memcpy(row, g->row, ((size_t)g->n + 1) * sizeof(uint64_t));
    if (col) memcpy(col, g->col, g->E * sizeof(uint32_t));
    if (w) memcpy(w, g->w, g->E * sizeof(uint32_t));
    if (k) memcpy(k, g->k, (size_t)g->n * sizeof(uint64_t));
}

/* ------------------------------------------------------------- modularity */

/* community.h:68-106: Q = sigma_in/(2*tw) - sum_c tot_c^2/(4*tw*tw), with
 * sigma_in the weight of entries whose endpoints share a community
 * (ActualNeighboorhood, operatorsCommunity.h:67-79) and tw = total_weight. */
double lvo_modularity(const lvo_graph *g, const uint32_t *comm,
                      int64_t *sigma_in_out, uint64_t *sum_tot_sq_out) {
    int64_t sigma_in = 0;
    uint64_t *tot = (uint64_t *)calloc(g->n ? g->n : 1, sizeof(uint64_t));
    for (uint32_t v = 0; v < g->n; ++v) {
        uint32_t cv = comm[v];
        tot[cv] += g->k[v];
        for (uint64_t j = g->row[v]; j < g->row[v + 1]; ++j)
            if (comm[g->col[j]] == cv) sigma_in += g->w[j];
    }
    uint64_t s = 0;
    for (uint32_t c = 0; c < g->n; ++c) s += tot[c] * tot[c];
    free(tot);
    if (sigma_in_out) *sigma_in_out = sigma_in;
    if (sum_tot_sq_out) *sum_tot_sq_out = s;
    double tw = g->total_weight;
    return (double)sigma_in / (2 * tw) - (double)s / (4 * tw * tw);
}

/* ------------------------------------------------------------------ sweep */

/* DeltaModularitySort / DeltaModularityHash (operatorsDelta.h:7-45):
 * float value = ((w - self) / tw) + ((k * (tot_own - k - tot_c)) / (2*tw*tw));
 * w - self is a float subtraction; the rest is double; rounded to float. */
static float dq_ref_f32(uint64_t w, uint64_t self, uint64_t k, int64_t tot_own,
                        int64_t tot_c, double tw) {
    float fw = (float)w, fs = (float)self;
    float d = fw - fs;
    double t1 = (double)d / tw;
    double t2 = ((double)(uint32_t)k * (((double)tot_own - (double)(uint32_t)k) - (double)tot_c)) /
                (2 * tw * tw);
    return (float)(t1 + t2);
}

float lvo_dq_ref_f32(uint64_t w, uint64_t self, uint64_t k, int64_t tot_own, int64_t tot_c,
                     double tw) {
    return dq_ref_f32(w, self, k, tot_own, tot_c, tw);
}

/* Vertex-aligned STEP_ROUND chunk ends (phaseOptimizationSort.cu:46-60,
 * :186-204; phaseOptimizationHash.cu:63-74):
 *   limit = round + STEP; limit = limit >= E ? E : row[src(limit) - 1].
 * The reference loops forever when that bound does not advance (a single
 * vertex spanning a whole chunk); here the bound then advances to the next
 * vertex boundary past `round`. */
static uint64_t *chunk_bounds(const lvo_graph *g, uint64_t step, uint32_t *n_chunks) {
    uint64_t cap = 16, nc = 0;
    uint64_t *b = (uint64_t *)malloc((cap + 1) * sizeof(uint64_t));
    uint64_t round = 0;
    b[0] = 0;
    while (round < g->E) {
        uint64_t limit = round + step;
        if (limit >= g->E) {
            limit = g->E;
        } else {
            /* src(limit): vertex whose row contains entry `limit` */
            uint32_t lo = 0, hi = g->n;  /* find largest v with row[v] <= limit */
            while (hi - lo > 1) { uint32_t mid = lo + (hi - lo) / 2; if (g->row[mid] <= limit) lo = mid; else hi = mid; }
            uint32_t s = lo;
            uint64_t cand = s > 0 ? g->row[s - 1] : 0;
            if (cand <= round) {
                cand = g->row[s] > round ? g->row[s] : g->row[s + 1];
            }
            limit = cand;
        }
        if (nc + 1 >= cap) { cap *= 2; b = (uint64_t *)realloc(b, (cap + 1) * sizeof(uint64_t)); }
        b[++nc] = limit;
        round = limit;
    }
    *n_chunks = (uint32_t)nc;
    return b;
}

int64_t lvo_best_moves(const lvo_graph *g, const uint32_t *comm, const int64_t *tot,
                       const uint8_t *active, int kind, const lvo_params *p,
                       uint32_t *target, uint64_t *e_active_out, uint32_t *n_considered_out) {
    const uint32_t n = g->n;
    const int64_t m2 = (int64_t)(2.0 * g->total_weight);
    const double tw = g->total_weight;
    const int highest = (p->tie == LVO_TIE_REFERENCE && kind == LVO_SWEEP_HASH);
    uint64_t *acc = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    uint32_t *touched = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint8_t *in_list = (uint8_t *)calloc(n ? n : 1, 1);

    /* vertex limit from the fast/sort sweeps' empty-chunk early return
     * (phaseOptimizationSort.cu:81-86, :235-240); the hash sweep has none. */
    uint32_t v_end = n;
    if (p->chunk_emulation && kind != LVO_SWEEP_HASH && g->E > 0) {
        uint32_t nc;
        uint64_t *b = chunk_bounds(g, p->chunk_entries, &nc);
        uint32_t v = 0;
        for (uint32_t c = 0; c < nc; ++c) {
            uint64_t sel = 0;
            uint32_t v0 = v;
            while (v < n && g->row[v] < b[c + 1]) {
                if (active[v])
                    for (uint64_t j = g->row[v]; j < g->row[v + 1]; ++j) sel += (g->col[j] != v);
                ++v;
            }
            if (sel == 0) { v_end = v0; break; }
        }
        free(b);
    }

    int64_t moved = 0;
    uint64_t e_active = 0;
    uint32_t n_considered = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t own = comm[v];
        target[v] = own;
        if (v >= v_end || !active[v]) continue;
        uint32_t nt = 0;
        /* copy_if TestTupleValue (operatorsCommunity.h:7-19): active src, src != dst;
         * hashmap.cuh:62-66 for the hash sweep. */
        for (uint64_t j = g->row[v]; j < g->row[v + 1]; ++j) {
            uint32_t x = g->col[j];
            if (x == v) continue;
            uint32_t c = comm[x];
            if (!in_list[c]) { in_list[c] = 1; touched[nt++] = c; }
            acc[c] += g->w[j];
            ++e_active;
        }
        if (nt == 0) continue;
        ++n_considered;
        /* self: select_self_node (phaseOptimizationSort.cu:9-25) / the hash
         * sweep's atomicAdd(self) (hashmap.cuh:62-63); the fast sweep passes a
         * zero vector (phaseOptimizationSort.cu:98), which equals w(v,own) at
         * the identity partition. */
        const uint64_t self = in_list[own] ? acc[own] : 0;
        const uint64_t kv = g->k[v];
        const int64_t tot_own = tot[own];
        uint32_t best_c = NONE32;
        int64_t best_i = 0;
        float best_f = 0.0f;
        for (uint32_t t = 0; t < nt; ++t) {
            const uint32_t c = touched[t];
            if (c == own) continue;   /* never wins: its delta-Q is -k^2/(2m^2) < 0 */
            int better;
            if (p->gain == LVO_GAIN_EXACT) {
                int64_t s = m2 * (int64_t)acc[c] - (int64_t)kv * tot[c];
                better = best_c == NONE32 || s > best_i ||
                         (s == best_i && (highest ? c > best_c : c < best_c));
                if (better) { best_i = s; best_c = c; }
            } else {
                float s = dq_ref_f32(acc[c], self, kv, tot_own, tot[c], tw);
                better = best_c == NONE32 || s > best_f ||
                         (s == best_f && (highest ? c > best_c : c < best_c));
                if (better) { best_f = s; best_c = c; }
            }
        }
        for (uint32_t t = 0; t < nt; ++t) { acc[touched[t]] = 0; in_list[touched[t]] = 0; }
        if (best_c == NONE32) continue;
        /* update_value_kernel_* (phaseOptimizationUpdate.cu:30, :64, :99):
         * move iff comm != c, delta-Q > 0, and on the fast sweep v < c. */
        int pos;
        if (p->gain == LVO_GAIN_EXACT)
            pos = best_i - m2 * (int64_t)self + (int64_t)kv * (tot_own - (int64_t)kv) > 0;
        else
            pos = best_f > 0.0f;
        if (!pos) continue;
        if (kind == LVO_SWEEP_FAST && v > best_c) continue;
        target[v] = best_c;
        ++moved;
    }
    free(acc); free(touched); free(in_list);
    if (e_active_out) *e_active_out = e_active;
    if (n_considered_out) *n_considered_out = n_considered;
    return moved;
}

/* ------------------------------------------------------------ aggregation */

int lvo_aggregate(const lvo_graph *g, const uint32_t *comm, uint32_t *new_id,
                  uint32_t *n_new_out, lvo_graph **out) {
    const uint32_t n = g->n;
    if (out) *out = NULL;
    /* community_map = inclusive_scan(CountNotZero(tot)) (phaseAggregation.cuh:27-43) */
    uint64_t *tot = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    for (uint32_t v = 0; v < n; ++v) tot[comm[v]] += g->k[v];
    uint32_t run = 0;
    for (uint32_t c = 0; c < n; ++c) {
        if (tot[c] > 0) new_id[c] = run++;
        else new_id[c] = NONE32;
    }
    free(tot);
    *n_new_out = run;
    if (!out) return LVO_OK;
    /* MakeCommunityPair (operatorsCommunity.h:37-51), sort (src,dst),
     * reduce_by_key sum (phaseAggregation.cuh:51-105) */
    uint64_t E = g->E;
    uint64_t *key = (uint64_t *)malloc((E ? E : 1) * sizeof(uint64_t));
    uint32_t *val = (uint32_t *)malloc((E ? E : 1) * sizeof(uint32_t));
    if (!key || !val) { free(key); free(val); return LVO_ERR_OOM; }
    for (uint32_t v = 0; v < n; ++v) {
        uint64_t cs = new_id[comm[v]];
        for (uint64_t j = g->row[v]; j < g->row[v + 1]; ++j) {
            key[j] = (cs << 32) | new_id[comm[g->col[j]]];
            val[j] = g->w[j];
        }
    }
    int rc = radix_sort_pairs(key, val, E);
    if (rc) { free(key); free(val); return rc; }
    uint64_t u = 0;
    for (uint64_t j = 0; j < E; ++j) {
        if (u > 0 && key[u - 1] == key[j]) {
            uint64_t s = (uint64_t)val[u - 1] + val[j];
            if (s > 0xFFFFFFFFull) { free(key); free(val); return LVO_ERR_RANGE; }
            val[u - 1] = (uint32_t)s;
        } else {
            key[u] = key[j]; val[u] = val[j]; ++u;
        }
    }
    lvo_graph *h = graph_alloc(run, u);
    if (!h) { free(key); free(val); return LVO_ERR_OOM; }
    for (uint64_t j = 0; j < u; ++j) {
        h->row[(key[j] >> 32) + 1]++;
        h->col[j] = (uint32_t)(key[j] & 0xFFFFFFFFu);
        h->w[j] = val[j];
    }
    free(key); free(val);
    for (uint32_t c = 0; c < run; ++c) h->row[c + 1] += h->row[c];
    /* PairSum reduce_by_key -> k', degree (phaseAggregation.cuh:146-170);
     * total_weight unchanged (:172) */
    graph_finish(h);
    h->total_weight = g->total_weight;
    h->n_links = u / 2;
    h->min_id = 0;
    *out = h;
    return LVO_OK;
}

/* ------------------------------------------------------------------- loop */

/* OptimizationPhase::optimize schedule (phaseOptimization.cuh:125-145). */
static int sweep_kind(int mode, int64_t exec) {
    if (exec == 0) return LVO_SWEEP_FAST;
    if (mode == LVO_HASH || mode == LVO_ADAPTIVE_MEMORY) return LVO_SWEEP_HASH;
    if (mode == LVO_ADAPTIVE_SPEED && exec > 4) return LVO_SWEEP_HASH;
    return LVO_SWEEP_SORT;
}

/* internal_run (phaseOptimization.cuh:158-181) from sweep index exec0 with
 * partition comm (totals tot, active set active); returns the sweep count
 * reached (exec).  Records sweeps at `level`. */
static int64_t sweep_loop(const lvo_graph *g, const lvo_params *p, lvo_result *r, int32_t level,
                          int64_t exec0, uint32_t *comm, uint32_t *target, int64_t *tot,
                          uint8_t *active, uint8_t *active_next, double *qp,
                          int32_t *n_sweep_recs, int32_t *total_sweeps, int *stop_all) {
    const uint32_t n = g->n;
    double q = *qp;
    int64_t exec = exec0;
    double delta;
    do {
        const double old_q = q;
        const int kind = sweep_kind(p->mode, exec);
        uint32_t n_active = 0;
        for (uint32_t v = 0; v < n; ++v) n_active += active[v];
        uint64_t e_active = 0;
        uint32_t n_considered = 0;
        int64_t moved = lvo_best_moves(g, comm, tot, active, kind, p, target, &e_active, &n_considered);
        /* fast_move_update (phaseOptimizationUpdate.cu:135-209): apply all
         * moves at once, then active[src] |= moved[dst] && comm[dst]!=comm[src] */
        memset(active_next, 0, n);
        for (uint32_t v = 0; v < n; ++v) {
            uint32_t c = target[v];
            if (c != comm[v]) {
                tot[c] += (int64_t)g->k[v];
                tot[comm[v]] -= (int64_t)g->k[v];
            }
        }
        uint8_t *changed = active; /* reuse: its_changed */
        for (uint32_t v = 0; v < n; ++v) { changed[v] = target[v] != comm[v]; comm[v] = target[v]; }
        for (uint32_t v = 0; v < n; ++v)
            for (uint64_t j = g->row[v]; j < g->row[v + 1]; ++j) {
                uint32_t x = g->col[j];
                if (changed[x] && comm[x] != comm[v]) { active_next[v] = 1; break; }
            }
        memcpy(active, active_next, n);
        ++exec;
        int64_t sigma_in; uint64_t stsq;
        q = lvo_modularity(g, comm, &sigma_in, &stsq);
        delta = q - old_q;
        if (r && r->sweeps && *n_sweep_recs < r->sweeps_cap) {
            lvo_sweep_rec *s = &r->sweeps[*n_sweep_recs];
            memset(s, 0, sizeof(*s));
            s->level = level; s->sweep = (int32_t)(exec - 1); s->kind = kind;
            s->n = n; s->n_active = n_active; s->n_moved = (uint32_t)moved;
            s->n_considered = n_considered;
            s->E = g->E; s->e_active = e_active;
            s->sigma_in = sigma_in; s->sum_tot_sq = stsq; s->q = q;
            s->comm_hash = lvo_fnv1a_u32(comm, n);
        }
        ++*n_sweep_recs;
        ++*total_sweeps;
        if (p->max_total_sweeps > 0 && *total_sweeps >= p->max_total_sweeps) { *stop_all = 1; break; }
    } while (exec <= p->max_sweeps && delta > p->sweep_threshold);
    *qp = q;
    return exec;
}

int lvo_louvain(const lvo_graph *g0, const lvo_params *p, lvo_result *r) {
    const uint32_t n0 = g0->n;
    uint32_t *best = (uint32_t *)malloc((n0 ? n0 : 1) * sizeof(uint32_t));
    if (!best) return LVO_ERR_OOM;
    for (uint32_t v = 0; v < n0; ++v) best[v] = v;  /* community.h:40-41 */
    const lvo_graph *g = g0;
    lvo_graph *owned = NULL;
    int32_t level = 0, total_sweeps = 0, n_sweep_recs = 0, n_level_recs = 0;
    uint32_t n_best = n0;
    double q = 0;
    int rc = LVO_OK;
    int stop_all = 0;
    int cont;
    do {
        const uint32_t n = g->n;
        uint32_t *comm = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
        uint32_t *target = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
        int64_t *tot = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
        uint8_t *active = (uint8_t *)malloc(n ? n : 1);
        uint8_t *active_next = (uint8_t *)malloc(n ? n : 1);
        uint32_t *new_id = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
        if (!comm || !target || !tot || !active || !active_next || !new_id) {
            free(comm); free(target); free(tot); free(active); free(active_next); free(new_id);
            rc = LVO_ERR_OOM; break;
        }
        /* Community ctor / update (community.h:31-46, :57-65) */
        for (uint32_t v = 0; v < n; ++v) { comm[v] = v; tot[v] = (int64_t)g->k[v]; active[v] = 1; }
        q = lvo_modularity(g, comm, NULL, NULL);
        const double q_level_start = q;
        const int64_t exec = sweep_loop(g, p, r, level, 0, comm, target, tot, active, active_next,
                                        &q, &n_sweep_recs, &total_sweeps, &stop_all);
        /* modularity_routine (modularityAlgorithms.h:52-53) */
        cont = (q - q_level_start) > p->level_threshold;
        if (stop_all) cont = 0;
        if (p->max_levels > 0 && level + 1 >= p->max_levels) cont = 0;
        /* AggregationPhase::run: update_best always, contraction if continuing */
        lvo_graph *next = NULL;
        rc = lvo_aggregate(g, comm, new_id, &n_best, cont ? &next : NULL);
        if (rc == LVO_OK) {
            for (uint32_t v = 0; v < n0; ++v) best[v] = new_id[comm[best[v]]];  /* MakeCommunityBest */
            if (r && r->level_recs && n_level_recs < r->levels_cap) {
                lvo_level_rec *L = &r->level_recs[n_level_recs];
                memset(L, 0, sizeof(*L));
                L->level = level; L->sweeps = (int32_t)exec; L->n = n; L->n_comm = n_best;
                L->E = g->E; L->q_start = q_level_start; L->q_end = q;
                L->best_hash = lvo_fnv1a_u32(best, n0);
            }
            ++n_level_recs;
        }
        free(comm); free(target); free(tot); free(active); free(active_next); free(new_id);
        if (rc != LVO_OK) break;
        if (cont) {
            if (owned) lvo_graph_free(owned);
            owned = next;
            g = owned;
            /* community.update() recomputes Q on the contracted graph; it equals q */
        }
        ++level;
    } while (cont);
    if (owned) lvo_graph_free(owned);
    /* Refinement (opt-in; the reference's RefinePhase is empty, phaseRefine.cuh:1-3):
     * the final partition on the level-0 graph, swept again from sweep index 1
     * (no fast sweep: the partition is not the identity), then renumbered. */
    if (rc == LVO_OK && p->refine && !stop_all) {
        uint32_t *comm = (uint32_t *)malloc((n0 ? n0 : 1) * sizeof(uint32_t));
        uint32_t *target = (uint32_t *)malloc((n0 ? n0 : 1) * sizeof(uint32_t));
        int64_t *tot = (int64_t *)calloc(n0 ? n0 : 1, sizeof(int64_t));
        uint8_t *active = (uint8_t *)malloc(n0 ? n0 : 1);
        uint8_t *active_next = (uint8_t *)malloc(n0 ? n0 : 1);
        uint32_t *new_id = (uint32_t *)malloc((n0 ? n0 : 1) * sizeof(uint32_t));
        if (!comm || !target || !tot || !active || !active_next || !new_id) {
            rc = LVO_ERR_OOM;
        } else {
            for (uint32_t v = 0; v < n0; ++v) {
                comm[v] = best[v];
                tot[best[v]] += (int64_t)g0->k[v];
                active[v] = 1;
            }
            q = lvo_modularity(g0, comm, NULL, NULL);
            (void)sweep_loop(g0, p, r, level, 1, comm, target, tot, active, active_next, &q,
                             &n_sweep_recs, &total_sweeps, &stop_all);
            rc = lvo_aggregate(g0, comm, new_id, &n_best, NULL);
            if (rc == LVO_OK)
                for (uint32_t v = 0; v < n0; ++v) best[v] = new_id[comm[v]];
        }
        free(comm); free(target); free(tot); free(active); free(active_next); free(new_id);
    }
    if (r) {
        r->modularity = q;
        r->n_of_best_communities = n_best;
        r->levels = level;
        r->total_sweeps = total_sweeps;
        if (r->best_communities) memcpy(r->best_communities, best, (size_t)n0 * sizeof(uint32_t));
    }
    free(best);
    return rc;
}
