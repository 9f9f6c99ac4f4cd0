/*
 * bm_oracle.c -- CPU restatement of brain-match's move-delta and score
 * arithmetic.  TEST INFRASTRUCTURE ONLY (the checker, never the product):
 * loaded by the test suite, the smoke check and the benchmark's CPU-baseline
 * leg.  Each function cites the reference routine (zhihuidu/brain-match,
 * src/...) that it restates.
 *
 * Pinning (what checks each function, beyond retyping the formula):
 *   orc_score_opt / orc_score_ref : score-matching built from source
 *       (golden scores), invariants S(id; G, G) = sum of weights.
 *   orc_delta_b : Delta_B = exact rescore difference on self-loop-free,
 *       duplicate-free graphs (brute force, every pair, tiny n); the
 *       accepted-move log of graph-alignment-optimized built from source.
 *   orc_first_improvement : the same log (rows, score, improvements every
 *       20 rows) and the converged mapping file.
 *   orc_delta_g3 : exact rescore difference for self-loop-free triples.
 *   orc_delta_p : Delta_P = exact + sum over intra-set edges (brute force).
 *   orc_delta_a : before/after restatement checked against the exact
 *       out-edge-only score difference (brute force).
 *   orc_rng_next : the published splitmix64 test vector (seed 1234567).
 *   orc_vertex_scores : hand-computed cases (last line wins, negative and
 *       vertex-0 lines, a weight wrapping negative as short, self-loops).
 *   orc_perturb : uniform draws = first 2k distinct generator values mod n;
 *       weighted draw frequencies against the weights (20,000 draws).
 *   orc_row_best_restart : equals plain row-best passes when no restart
 *       triggers; best score = rescoring of the returned mapping.
 *   orc_quick : counter identities and rescoring of the returned mapping.
 *   orc_lex_triples : itertools.combinations order.
 *   orc_triples_pass : a brute-force restatement that rescores every
 *       permutation of every triple (simple graphs, where Delta_G3 is the
 *       exact change), with the same generator draws.
 */
#include "bm_oracle.h"

#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OMIN(a, b) ((a) < (b) ? (a) : (b))

struct orc_graphs {
    int n;
    int64_t e1, e2;
    /* G1 per-vertex lists in file order (EdgeMap edges / reverse_edges) */
    int64_t *out_ptr, *in_ptr;
    int32_t *out_dst, *out_w, *in_src, *in_w;
    /* dense short matrices, last write wins (add_edge :185) */
    int16_t *w1d, *w2d;
    int16_t **w2rows; /* row pointers, the reference's short** layout */
    /* de-duplicated views with full int weights (score-matching add_edge
     * :339-358 / hill-climb insert_edge :101-125), rows sorted by target */
    int64_t *g1dd_ptr, *g2dd_ptr;
    int32_t *g1dd_dst, *g1dd_w, *g2dd_dst, *g2dd_w;
};

int orc_n(const orc_graphs *g) { return g->n; }

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* stable counting sort of edges by key[] into CSR */
static void csr_stable(int n, int64_t m, const int32_t *key, const int32_t *other, const int32_t *w,
                       int64_t **ptr_out, int32_t **oth_out, int32_t **w_out) {
    int64_t *ptr = xcalloc((size_t)n + 2, sizeof(int64_t));
    int32_t *oth = xcalloc((size_t)m, sizeof(int32_t));
    int32_t *ww = xcalloc((size_t)m, sizeof(int32_t));
    for (int64_t i = 0; i < m; i++) ptr[key[i] + 1]++;
    for (int v = 0; v <= n; v++) ptr[v + 1] += ptr[v];
    int64_t *pos = xcalloc((size_t)n + 1, sizeof(int64_t));
    for (int v = 0; v <= n; v++) pos[v] = ptr[v];
    for (int64_t i = 0; i < m; i++) {
        int64_t p = pos[key[i]]++;
        oth[p] = other[i];
        ww[p] = w[i];
    }
    free(pos);
    *ptr_out = ptr;
    *oth_out = oth;
    *w_out = ww;
}

/* de-duplicate a stable CSR (last occurrence wins), rows sorted by target */
typedef struct { int32_t d, w; } dw_t;
static int cmp_dw(const void *x, const void *y) {
    const dw_t *a = x, *b = y;
    return (a->d > b->d) - (a->d < b->d);
}
static void dedup_csr(int n, const int64_t *ptr, const int32_t *dst, const int32_t *w,
                      int64_t **dptr, int32_t **ddst, int32_t **dw) {
    int64_t m = ptr[n + 1];
    int64_t *np = xcalloc((size_t)n + 2, sizeof(int64_t));
    int32_t *nd = xcalloc((size_t)m, sizeof(int32_t));
    int32_t *nw = xcalloc((size_t)m, sizeof(int32_t));
    int32_t *last = xcalloc((size_t)n + 1, sizeof(int32_t)); /* last position+1 of dst in row */
    int64_t k = 0;
    for (int v = 0; v <= n; v++) {
        np[v] = k;
        for (int64_t i = ptr[v]; i < ptr[v + 1]; i++) last[dst[i]] = (int32_t)(i - ptr[v] + 1);
        int64_t start = k;
        for (int64_t i = ptr[v]; i < ptr[v + 1]; i++) {
            if (last[dst[i]] == (int32_t)(i - ptr[v] + 1)) { /* last occurrence */
                nd[k] = dst[i];
                nw[k] = w[i];
                k++;
            }
        }
        for (int64_t i = ptr[v]; i < ptr[v + 1]; i++) last[dst[i]] = 0;
        dw_t *tmp = xcalloc((size_t)(k - start), sizeof(dw_t));
        for (int64_t i = start; i < k; i++) tmp[i - start] = (dw_t){nd[i], nw[i]};
        qsort(tmp, (size_t)(k - start), sizeof(dw_t), cmp_dw);
        for (int64_t i = start; i < k; i++) {
            nd[i] = tmp[i - start].d;
            nw[i] = tmp[i - start].w;
        }
        free(tmp);
    }
    np[n + 1] = k;
    free(last);
    *dptr = np;
    *ddst = nd;
    *dw = nw;
}

orc_graphs *orc_build(int n, const int32_t *s1, const int32_t *d1, const int32_t *w1, int64_t e1,
                      const int32_t *s2, const int32_t *d2, const int32_t *w2, int64_t e2) {
    for (int64_t i = 0; i < e1; i++)
        if (s1[i] < 0 || s1[i] > n || d1[i] < 0 || d1[i] > n) return NULL;
    for (int64_t i = 0; i < e2; i++)
        if (s2[i] < 0 || s2[i] > n || d2[i] < 0 || d2[i] > n) return NULL;
    orc_graphs *g = xcalloc(1, sizeof(*g));
    g->n = n;
    g->e1 = e1;
    g->e2 = e2;
    csr_stable(n, e1, s1, d1, w1, &g->out_ptr, &g->out_dst, &g->out_w);
    csr_stable(n, e1, d1, s1, w1, &g->in_ptr, &g->in_src, &g->in_w);
    size_t N1 = (size_t)n + 1;
    g->w1d = xcalloc(N1 * N1, sizeof(int16_t));
    g->w2d = xcalloc(N1 * N1, sizeof(int16_t));
    for (int64_t i = 0; i < e1; i++) g->w1d[(size_t)s1[i] * N1 + d1[i]] = (int16_t)w1[i];
    for (int64_t i = 0; i < e2; i++) g->w2d[(size_t)s2[i] * N1 + d2[i]] = (int16_t)w2[i];
    g->w2rows = xcalloc(N1, sizeof(int16_t *));
    for (size_t r = 0; r < N1; r++) g->w2rows[r] = g->w2d + r * N1;
    dedup_csr(n, g->out_ptr, g->out_dst, g->out_w, &g->g1dd_ptr, &g->g1dd_dst, &g->g1dd_w);
    int64_t *p2;
    int32_t *o2, *ww2;
    csr_stable(n, e2, s2, d2, w2, &p2, &o2, &ww2);
    dedup_csr(n, p2, o2, ww2, &g->g2dd_ptr, &g->g2dd_dst, &g->g2dd_w);
    free(p2);
    free(o2);
    free(ww2);
    return g;
}

void orc_free(orc_graphs *g) {
    if (!g) return;
    free(g->out_ptr); free(g->in_ptr); free(g->out_dst); free(g->out_w); free(g->in_src); free(g->in_w);
    free(g->w1d); free(g->w2d); free(g->w2rows);
    free(g->g1dd_ptr); free(g->g1dd_dst); free(g->g1dd_w);
    free(g->g2dd_ptr); free(g->g2dd_dst); free(g->g2dd_w);
    free(g);
}

#define W1(g, x, y) ((int)(g)->w1d[(size_t)(x) * ((size_t)(g)->n + 1) + (size_t)(y)])
#define W2(g, x, y) ((int)(g)->w2d[(size_t)(x) * ((size_t)(g)->n + 1) + (size_t)(y)])

/* graph-alignment-optimized.c:292-307: the loop starts at src_m = 1 (:295),
 * so lines leaving vertex 0 are never scored (Delta_B still counts them as
 * in-entries of their target, which is why the reference's verify step at
 * :570-576 can disagree with an accepted delta on such input) */
int64_t orc_score_opt(const orc_graphs *g, const int32_t *map) {
    int64_t s = 0;
    for (int u = 1; u <= g->n; u++)
        for (int64_t i = g->out_ptr[u]; i < g->out_ptr[u + 1]; i++) {
            int v = g->out_dst[i];
            s += OMIN(g->out_w[i], W2(g, map[u], map[v]));
        }
    return s;
}

/* find_edge in the de-duplicated G2 (score-matching.c:327-337) */
static int g2_lookup(const orc_graphs *g, int u, int v, int *w) {
    if (u < 0 || u > g->n || v < 0 || v > g->n) return 0;
    int64_t lo = g->g2dd_ptr[u], hi = g->g2dd_ptr[u + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (g->g2dd_dst[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    if (lo < g->g2dd_ptr[u + 1] && g->g2dd_dst[lo] == v) {
        *w = g->g2dd_w[lo];
        return 1;
    }
    return 0;
}

/* score-matching.c:558-598 */
int64_t orc_score_ref(const orc_graphs *g, const int32_t *map) {
    int64_t s = 0;
    for (int u = 0; u <= g->n; u++) {
        for (int64_t i = g->g1dd_ptr[u]; i < g->g1dd_ptr[u + 1]; i++) {
            int v = g->g1dd_dst[i];
            int su = map[u], tv = map[v];
            if (su < 0 || tv < 0) continue;
            int w2;
            if (g2_lookup(g, su, tv, &w2)) s += OMIN(g->g1dd_w[i], w2);
        }
    }
    return s;
}

/* graph-alignment-optimized.c:321-408 */
int32_t orc_delta_b(const orc_graphs *g, const int32_t *map, int a, int b) {
    int f1 = map[a], f2 = map[b];
    int delta = 0;
    for (int64_t i = g->out_ptr[a]; i < g->out_ptr[a + 1]; i++) {
        int d = g->out_dst[i];
        if (d == b) continue;
        int w = g->out_w[i], df = map[d];
        delta += OMIN(w, W2(g, f2, df)) - OMIN(w, W2(g, f1, df));
    }
    for (int64_t i = g->in_ptr[a]; i < g->in_ptr[a + 1]; i++) {
        int s = g->in_src[i];
        if (s == b) continue;
        int w = g->in_w[i], sf = map[s];
        delta += OMIN(w, W2(g, sf, f2)) - OMIN(w, W2(g, sf, f1));
    }
    for (int64_t i = g->out_ptr[b]; i < g->out_ptr[b + 1]; i++) {
        int d = g->out_dst[i];
        if (d == a) continue;
        int w = g->out_w[i], df = map[d];
        delta += OMIN(w, W2(g, f1, df)) - OMIN(w, W2(g, f2, df));
    }
    for (int64_t i = g->in_ptr[b]; i < g->in_ptr[b + 1]; i++) {
        int s = g->in_src[i];
        if (s == a) continue;
        int w = g->in_w[i], sf = map[s];
        delta += OMIN(w, W2(g, sf, f1)) - OMIN(w, W2(g, sf, f2));
    }
    int m12 = W1(g, a, b);
    if (m12 > 0) delta += OMIN(m12, W2(g, f2, f1)) - OMIN(m12, W2(g, f1, f2));
    int m21 = W1(g, b, a);
    if (m21 > 0) delta += OMIN(m21, W2(g, f1, f2)) - OMIN(m21, W2(g, f2, f1));
    return delta;
}

/* graph-re-alignment-gpu.cu:60-128 (dense scan; self-loops never visited) */
int32_t orc_delta_g2(const orc_graphs *g, const int32_t *map, int a, int b) {
    int f1 = map[a], f2 = map[b];
    int delta = 0;
    for (int i = 1; i <= g->n; i++) {
        if (i == a || i == b) continue;
        int w;
        if ((w = W1(g, a, i)) > 0) delta += OMIN(w, W2(g, f2, map[i])) - OMIN(w, W2(g, f1, map[i]));
        if ((w = W1(g, i, a)) > 0) delta += OMIN(w, W2(g, map[i], f2)) - OMIN(w, W2(g, map[i], f1));
        if ((w = W1(g, b, i)) > 0) delta += OMIN(w, W2(g, f1, map[i])) - OMIN(w, W2(g, f2, map[i]));
        if ((w = W1(g, i, b)) > 0) delta += OMIN(w, W2(g, map[i], f1)) - OMIN(w, W2(g, map[i], f2));
    }
    int m12 = W1(g, a, b);
    if (m12 > 0) delta += OMIN(m12, W2(g, f2, f1)) - OMIN(m12, W2(g, f1, f2));
    int m21 = W1(g, b, a);
    if (m21 > 0) delta += OMIN(m21, W2(g, f1, f2)) - OMIN(m21, W2(g, f2, f1));
    return delta;
}

/* permutation table of triple.cu:85-92 */
static void g3_new(int p, int f1, int f2, int f3, int *n1, int *n2, int *n3) {
    static const int tab[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int f[3] = {f1, f2, f3};
    *n1 = f[tab[p][0]];
    *n2 = f[tab[p][1]];
    *n3 = f[tab[p][2]];
}

/* triple.cu:98-208 */
int32_t orc_delta_g3(const orc_graphs *g, const int32_t *map, int m1, int m2, int m3, int p) {
    int m[3] = {m1, m2, m3};
    int f[3] = {map[m1], map[m2], map[m3]};
    int nf[3];
    g3_new(p, f[0], f[1], f[2], &nf[0], &nf[1], &nf[2]);
    int delta = 0;
    for (int k = 0; k < 3; k++) {
        for (int i = 1; i <= g->n; i++) {
            if (i == m1 || i == m2 || i == m3) continue;
            int w;
            if ((w = W1(g, m[k], i)) > 0)
                delta += OMIN(w, W2(g, nf[k], map[i])) - OMIN(w, W2(g, f[k], map[i]));
            if ((w = W1(g, i, m[k])) > 0)
                delta += OMIN(w, W2(g, map[i], nf[k])) - OMIN(w, W2(g, map[i], f[k]));
        }
    }
    /* the six intra-triple cells, :164-207 */
    static const int pairs[6][2] = {{0, 1}, {1, 0}, {1, 2}, {2, 1}, {2, 0}, {0, 2}};
    for (int q = 0; q < 6; q++) {
        int x = pairs[q][0], y = pairs[q][1];
        int w = W1(g, m[x], m[y]);
        if (w > 0) delta += OMIN(w, W2(g, nf[x], nf[y])) - OMIN(w, W2(g, f[x], f[y]));
    }
    return delta;
}

int32_t orc_best_g3(const orc_graphs *g, const int32_t *map, int m1, int m2, int m3, int mode, int *best_p) {
    int best = 0, bp = 0;
    for (int p = 1; p < 6; p++) {
        if (mode == 0 && p != 3 && p != 4) continue;
        int d = orc_delta_g3(g, map, m1, m2, m3, p);
        if (d > best) {
            best = d;
            bp = p;
        }
    }
    *best_p = bp;
    return best;
}

/* triple.cu:903-934 */
void orc_apply_g3(int32_t *map, int m1, int m2, int m3, int p) {
    int n1, n2, n3;
    g3_new(p, map[m1], map[m2], map[m3], &n1, &n2, &n3);
    map[m1] = n1;
    map[m2] = n2;
    map[m3] = n3;
}

/* graph-alignment-permset-parallel.c:304-339 */
int32_t orc_delta_p(const orc_graphs *g, const int32_t *map, const int32_t *set, const int32_t *perm, int k) {
    /* new_mapping = copy of current with the set permuted (:307-313); only
     * the set's entries differ, so look them up in a small table. */
    int32_t newf[8];
    for (int i = 0; i < k; i++) newf[i] = map[set[perm[i]]];
    int delta = 0;
    for (int i = 0; i < k; i++) {
        int v = set[i];
        for (int64_t j = g->out_ptr[v]; j < g->out_ptr[v + 1]; j++) {
            int w = g->out_dst[j], wt = g->out_w[j];
            int nv = newf[i], nw = map[w];
            for (int q = 0; q < k; q++)
                if (set[q] == w) nw = newf[q];
            delta -= OMIN(wt, W2(g, map[v], map[w]));
            delta += OMIN(wt, W2(g, nv, nw));
        }
        for (int64_t j = g->in_ptr[v]; j < g->in_ptr[v + 1]; j++) {
            int u = g->in_src[j], wt = g->in_w[j];
            int nv = newf[i], nu = map[u];
            for (int q = 0; q < k; q++)
                if (set[q] == u) nu = newf[q];
            delta -= OMIN(wt, W2(g, map[u], map[v]));
            delta += OMIN(wt, W2(g, nu, nv));
        }
    }
    return delta;
}

/* graph-alignment-permset-parallel.c:342-363 */
int orc_next_permutation(int32_t *arr, int n) {
    int i = n - 2;
    while (i >= 0 && arr[i] >= arr[i + 1]) i--;
    if (i < 0) return 0;
    int j = n - 1;
    while (arr[j] <= arr[i]) j--;
    int32_t t = arr[i];
    arr[i] = arr[j];
    arr[j] = t;
    for (int x = i + 1, y = n - 1; x < y; x++, y--) {
        t = arr[x];
        arr[x] = arr[y];
        arr[y] = t;
    }
    return 1;
}

/* Lehmer-code decode: the rank-th permutation of 0..k-1 in lexicographic
 * order, which is the order next_permutation visits from the identity. */
void orc_unrank_perm(int k, int64_t rank, int32_t *perm) {
    int32_t pool[8];
    int64_t fact[9] = {1};
    for (int i = 1; i <= 8; i++) fact[i] = fact[i - 1] * i;
    for (int i = 0; i < k; i++) pool[i] = i;
    int left = k;
    for (int i = 0; i < k; i++) {
        int64_t f = fact[k - 1 - i];
        int idx = (int)(rank / f);
        rank %= f;
        perm[i] = pool[idx];
        for (int j = idx; j < left - 1; j++) pool[j] = pool[j + 1];
        left--;
    }
}

/* evaluate all k! permutations in next_permutation order, strict > from 0
 * (graph-alignment-permset-parallel.c:547-573, :665-674) */
int32_t orc_best_perm(const orc_graphs *g, const int32_t *map, const int32_t *set, int k, int64_t *rank) {
    int32_t perm[8];
    for (int i = 0; i < k; i++) perm[i] = i;
    int32_t best = 0;
    int64_t br = -1, r = 0;
    do {
        int32_t d = orc_delta_p(g, map, set, perm, k);
        if (d > best) {
            best = d;
            br = r;
        }
        r++;
    } while (orc_next_permutation(perm, k));
    *rank = br;
    return best;
}

/* :508-518 */
void orc_apply_perm(int32_t *map, const int32_t *set, const int32_t *perm, int k) {
    int32_t tmp[8];
    for (int i = 0; i < k; i++) tmp[i] = map[set[perm[i]]];
    for (int i = 0; i < k; i++) map[set[i]] = tmp[i];
}

/* calculate_edge_alignment, refine-matching-hillclimb.c:210-220 */
static int align_a(const orc_graphs *g, int w1, int f, int h) {
    int w2;
    if (f <= 0 || h <= 0) return 0;
    if (!g2_lookup(g, f, h, &w2)) return 0;
    return OMIN(w1, w2);
}

/* refine-matching-hillclimb.c:244-296 */
int32_t orc_delta_a(const orc_graphs *g, const int32_t *map_in, int a, int b) {
    int32_t fa = map_in[a], fb = map_in[b];
    int before = 0, after = 0;
    for (int pass = 0; pass < 2; pass++) {
        int acc = 0;
        for (int side = 0; side < 2; side++) {
            int x = side ? b : a, other = side ? a : b;
            for (int64_t i = g->g1dd_ptr[x]; i < g->g1dd_ptr[x + 1]; i++) {
                int t = g->g1dd_dst[i];
                if (t == other) continue;
                int mx = (x == a) ? (pass ? fb : fa) : (pass ? fa : fb);
                int mt = map_in[t];
                if (t == a) mt = pass ? fb : fa;
                else if (t == b) mt = pass ? fa : fb;
                acc += align_a(g, g->g1dd_w[i], mx, mt);
            }
        }
        if (pass == 0) before = acc;
        else after = acc;
    }
    return after - before;
}

int64_t orc_delta_swap_exact(const orc_graphs *g, const int32_t *map, int a, int b) {
    int32_t *m2 = xcalloc((size_t)g->n + 1, sizeof(int32_t));
    memcpy(m2, map, ((size_t)g->n + 1) * sizeof(int32_t));
    int64_t s0 = orc_score_opt(g, m2);
    int32_t t = m2[a];
    m2[a] = m2[b];
    m2[b] = t;
    int64_t s1 = orc_score_opt(g, m2);
    free(m2);
    return s1 - s0;
}

/* graph-alignment-optimized.c:495-581 */
int orc_first_improvement(const orc_graphs *g, int32_t *map, int rows_per_sweep, int max_sweeps,
                          int32_t *acc, int64_t acc_cap, int64_t *n_acc,
                          int64_t *logv, int64_t log_cap, int64_t *n_log, int log_every,
                          int64_t *final_score) {
    int n = g->n;
    if (rows_per_sweep < n) rows_per_sweep = n;
    int64_t score = orc_score_opt(g, map);
    int64_t rows = 0, improvements = 0, na = 0, nl = 0;
    int sweeps = 0;
    for (;;) {
        if (max_sweeps > 0 && sweeps >= max_sweeps) break;
        int64_t before = improvements;
        for (int m1 = 1; m1 <= rows_per_sweep; m1++) {
            rows++;
            if (log_every > 0 && rows % log_every == 0) {
                if (logv && nl < log_cap) {
                    logv[3 * nl] = rows;
                    logv[3 * nl + 1] = score;
                    logv[3 * nl + 2] = improvements;
                }
                nl++;
            }
            if (m1 > n) continue; /* padding rows never accept (weights >= 0) */
            for (int m2 = 1; m2 <= n; m2++) {
                if (m1 == m2) continue;
                int d = orc_delta_b(g, map, m1, m2);
                if (d > 0) {
                    int32_t t = map[m1];
                    map[m1] = map[m2];
                    map[m2] = t;
                    score += d;
                    improvements++;
                    if (acc && na < acc_cap) {
                        acc[3 * na] = m1;
                        acc[3 * na + 1] = m2;
                        acc[3 * na + 2] = d;
                    }
                    na++;
                    int64_t verify = orc_score_opt(g, map);
                    if (verify != score) score = verify; /* :570-576 */
                }
            }
        }
        sweeps++;
        if (improvements == before) break;
    }
    if (n_acc) *n_acc = na;
    if (n_log) *n_log = nl;
    if (final_score) *final_score = score;
    return sweeps;
}

/* alignment-omp.c:598-658 at one thread */
int64_t orc_row_best_pass(const orc_graphs *g, int32_t *map, int32_t *acc, int64_t acc_cap) {
    int n = g->n;
    int64_t na = 0;
    for (int m1 = 1; m1 <= n; m1++) {
        int maxdelta = -1, maxv = -1;
        for (int m2 = m1 + 1; m2 <= n; m2++) {
            int d = orc_delta_b(g, map, m1, m2);
            if (d > maxdelta) {
                maxdelta = d;
                maxv = m2;
            }
        }
        if (maxdelta > 0) {
            int32_t t = map[m1];
            map[m1] = map[maxv];
            map[maxv] = t;
            if (acc && na < acc_cap) {
                acc[3 * na] = m1;
                acc[3 * na + 1] = maxv;
                acc[3 * na + 2] = maxdelta;
            }
            na++;
        }
    }
    return na;
}

/* graph-re-alignment-gpu.cu:859-868 arg-max over the lexicographic pair list */
int32_t orc_all_pairs(const orc_graphs *g, const int32_t *map, int32_t *best_a, int32_t *best_b,
                      int32_t *row_best_b, int32_t *row_best_d, int32_t *delta_matrix) {
    int n = g->n;
    int32_t best = 0, ba = 0, bb = 0;
    size_t N1 = (size_t)n + 1;
    if (delta_matrix) memset(delta_matrix, 0, N1 * N1 * sizeof(int32_t));
    for (int a = 0; a <= n; a++) {
        if (row_best_b) row_best_b[a] = 0;
        if (row_best_d) row_best_d[a] = INT32_MIN;
    }
    for (int a = 1; a <= n; a++) {
        int32_t rb = INT32_MIN, rbb = 0;
        for (int b = a + 1; b <= n; b++) {
            int32_t d = orc_delta_b(g, map, a, b);
            if (delta_matrix) {
                delta_matrix[(size_t)a * N1 + b] = d;
                delta_matrix[(size_t)b * N1 + a] = d;
            }
            if (d > rb) {
                rb = d;
                rbb = b;
            }
            if (d > best) {
                best = d;
                ba = a;
                bb = b;
            }
        }
        if (row_best_b) row_best_b[a] = rbb;
        if (row_best_d) row_best_d[a] = rb;
    }
    if (best_a) *best_a = ba;
    if (best_b) *best_b = bb;
    return best;
}

/* ---- reference-layout OpenMP row-best (CPU baseline timing) ---------- */

/* calculate_swap_delta exactly as graph-alignment-optimized.c:321-408,
 * reading W2 through short** row pointers as the reference does. */
static int ref_layout_delta(const orc_graphs *g, int16_t *const *adj_f, const int32_t *mapping, int m1, int m2) {
    int f1 = mapping[m1], f2 = mapping[m2];
    int delta = 0;
    const int64_t *op = g->out_ptr, *ip = g->in_ptr;
    for (int64_t i = op[m1]; i < op[m1 + 1]; i++) {
        int d = g->out_dst[i];
        if (d == m2) continue;
        int w = g->out_w[i], df = mapping[d];
        delta += OMIN(w, adj_f[f2][df]) - OMIN(w, adj_f[f1][df]);
    }
    for (int64_t i = ip[m1]; i < ip[m1 + 1]; i++) {
        int s = g->in_src[i];
        if (s == m2) continue;
        int w = g->in_w[i], sf = mapping[s];
        delta += OMIN(w, adj_f[sf][f2]) - OMIN(w, adj_f[sf][f1]);
    }
    for (int64_t i = op[m2]; i < op[m2 + 1]; i++) {
        int d = g->out_dst[i];
        if (d == m1) continue;
        int w = g->out_w[i], df = mapping[d];
        delta += OMIN(w, adj_f[f1][df]) - OMIN(w, adj_f[f2][df]);
    }
    for (int64_t i = ip[m2]; i < ip[m2 + 1]; i++) {
        int s = g->in_src[i];
        if (s == m1) continue;
        int w = g->in_w[i], sf = mapping[s];
        delta += OMIN(w, adj_f[sf][f1]) - OMIN(w, adj_f[sf][f2]);
    }
    int m12 = W1(g, m1, m2);
    if (m12 > 0) delta += OMIN(m12, adj_f[f2][f1]) - OMIN(m12, adj_f[f1][f2]);
    int m21 = W1(g, m2, m1);
    if (m21 > 0) delta += OMIN(m21, adj_f[f1][f2]) - OMIN(m21, adj_f[f2][f1]);
    return delta;
}

/* alignment-omp.c:598-622: one parallel-for per row over m2 > m1 with a
 * critical-section arg-max per evaluation. */
int64_t orc_omp_rowbest_rows(const orc_graphs *g, const int32_t *map, int row_lo, int row_hi, int threads,
                             int64_t *checksum) {
    int n = g->n;
    int64_t evals = 0, cs = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    int16_t *const *adj_f = g->w2rows;
    for (int m1 = row_lo; m1 < row_hi && m1 <= n; m1++) {
        if (m1 < 1) continue;
        int maxdelta = -1, maxv = -1;
#pragma omp parallel for shared(maxdelta, maxv) schedule(dynamic)
        for (int m2 = m1 + 1; m2 <= n; m2++) {
            int delta = ref_layout_delta(g, adj_f, map, m1, m2);
#pragma omp critical
            {
                if (delta > maxdelta) {
                    maxdelta = delta;
                    maxv = m2;
                }
            }
        }
        evals += n - m1;
        cs += maxdelta; /* maxv depends on thread arrival order on ties */
        (void)maxv;
    }
    if (checksum) *checksum = cs;
    return evals;
}

/* ---- the same arithmetic over many rows / moves ------------------------ */

static void set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

void orc_row_best_rows(const orc_graphs *g, const int32_t *map, const int32_t *rows, int64_t count,
                       int32_t *best_b, int32_t *best_d, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < count; i++) {
        int a = rows[i];
        int32_t rb = INT32_MIN, rbb = 0;
        for (int b = a + 1; b <= g->n; b++) {
            int32_t d = orc_delta_b(g, map, a, b);
            if (d > rb) {
                rb = d;
                rbb = b;
            }
        }
        best_b[i] = rbb;
        best_d[i] = rb;
    }
}

void orc_delta_rows(const orc_graphs *g, const int32_t *map, const int32_t *rows, int64_t count, int32_t *out,
                    int threads) {
    const int64_t N1 = (int64_t)g->n + 1;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < count * N1; t++) {
        int64_t i = t / N1;
        int b = (int)(t % N1), a = rows[i];
        out[t] = (b == 0 || b == a) ? 0 : orc_delta_b(g, map, a, b);
    }
}

int64_t orc_first_improvement_rows(const orc_graphs *g, int32_t *map, int max_rows, int32_t *acc, int64_t acc_cap,
                                   int threads) {
    const int n = g->n, BLOCK = 1024;
    int32_t *d = xcalloc(BLOCK, sizeof(int32_t));
    int64_t na = 0;
    set_threads(threads);
    for (int m1 = 1; m1 <= max_rows && m1 <= n; m1++) {
        int cursor = 1;
        while (cursor <= n) {
            int hi = cursor + BLOCK - 1 > n ? n : cursor + BLOCK - 1;
#pragma omp parallel for schedule(static)
            for (int m2 = cursor; m2 <= hi; m2++) d[m2 - cursor] = m2 == m1 ? 0 : orc_delta_b(g, map, m1, m2);
            int hit = 0;
            for (int m2 = cursor; m2 <= hi; m2++)
                if (d[m2 - cursor] > 0) {
                    int32_t t = map[m1];
                    map[m1] = map[m2];
                    map[m2] = t;
                    if (acc && na < acc_cap) {
                        acc[3 * na] = m1;
                        acc[3 * na + 1] = m2;
                        acc[3 * na + 2] = d[m2 - cursor];
                    }
                    na++;
                    cursor = m2 + 1;
                    hit = 1;
                    break;
                }
            if (!hit) cursor = hi + 1;
        }
    }
    free(d);
    return na;
}

void orc_best_g3_many(const orc_graphs *g, const int32_t *map, const int32_t *triples, int64_t count, int mode,
                      int32_t *best_d, int32_t *best_p, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < count; i++) {
        int p;
        best_d[i] = orc_best_g3(g, map, triples[3 * i], triples[3 * i + 1], triples[3 * i + 2], mode, &p);
        best_p[i] = p;
    }
}

void orc_best_perm_many(const orc_graphs *g, const int32_t *map, const int32_t *sets, int k, int64_t count,
                        int32_t *best_d, int64_t *best_rank, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < count; i++) best_d[i] = orc_best_perm(g, map, sets + (size_t)k * i, k, &best_rank[i]);
}

/* ------------------------------------------------------------------------
 * Perturbation, restart and quick search (alignment-omp.c:495-541, 543-678;
 * quick.cu:766-887, 1137-1324).  The reference draws from a time-seeded
 * rand(); these take a splitmix64 state so that a run is repeatable.
 */
uint64_t orc_rng_next(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* quick.cu:766-793 */
void orc_vertex_scores(const orc_graphs *g, float *scores) {
    int n = g->n;
    scores[0] = 0.f;
    for (int i = 1; i <= n; i++) {
        float out_weight = 0.0f, in_weight = 0.0f;
        for (int j = 1; j <= n; j++)
            if (W1(g, i, j) > 0) out_weight += W1(g, i, j);
        for (int j = 1; j <= n; j++)
            if (W1(g, j, i) > 0) in_weight += W1(g, j, i);
        float a = 0.7f * out_weight;
        float b = 0.3f * in_weight;
        scores[i] = a + b;
    }
}

/* quick.cu:796-814 */
static int orc_select_by_score(const float *scores, const uint8_t *used, int n, float total, uint64_t *rng) {
    float u = (float)(orc_rng_next(rng) >> 40) * (1.0f / 16777216.0f);
    float r = u * total;
    float cumsum = 0.0f;
    for (int i = 1; i <= n; i++) {
        if (!used[i]) {
            cumsum += scores[i];
            if (cumsum >= r) return i;
        }
    }
    for (int i = 1; i <= n; i++)
        if (!used[i]) return i;
    return 1;
}

int orc_perturb(const orc_graphs *g, int32_t *map, int k, int mode, uint64_t *rng, int32_t *sel) {
    int n = g->n;
    if (k < 1 || 2 * k > n) return -1;
    uint8_t *used = xcalloc((size_t)n + 1, 1);
    int count = 0;
    if (mode == 0) {
        while (count < 2 * k) {
            int idx = (int)(orc_rng_next(rng) % (uint64_t)n);
            if (!used[idx]) {
                sel[count++] = idx;
                used[idx] = 1;
            }
        }
    } else {
        float *scores = xcalloc((size_t)n + 1, sizeof(float));
        orc_vertex_scores(g, scores);
        float total = 0;
        for (int i = 1; i <= n; i++) total += scores[i];
        while (count < k) {
            int idx = orc_select_by_score(scores, used, n, total, rng);
            if (!used[idx]) {
                sel[count++] = idx;
                used[idx] = 1;
                total -= scores[idx];
            }
        }
        while (count < 2 * k) {
            int idx = (int)(orc_rng_next(rng) % (uint64_t)n) + 1;
            if (!used[idx]) {
                sel[count++] = idx;
                used[idx] = 1;
            }
        }
        free(scores);
    }
    int32_t orig[16];
    for (int i = 0; i < 2 * k && i < 16; i++) orig[i] = map[sel[i]];
    if (2 * k > 16) {
        int32_t *o = xcalloc((size_t)2 * k, sizeof(int32_t));
        for (int i = 0; i < 2 * k; i++) o[i] = map[sel[i]];
        for (int i = 0; i < k; i++) {
            map[sel[i]] = o[i + k];
            map[sel[i + k]] = o[i];
        }
        free(o);
    } else {
        for (int i = 0; i < k; i++) {
            map[sel[i]] = orig[i + k];
            map[sel[i + k]] = orig[i];
        }
    }
    free(used);
    return 0;
}

/* alignment-omp.c:543-678 at one thread */
int orc_row_best_restart(const orc_graphs *g, int32_t *map, uint64_t seed, int initial_swaps, int restart_swaps,
                         int idle_passes, int64_t max_passes, int64_t *out) {
    int n = g->n;
    size_t bytes = ((size_t)n + 1) * sizeof(int32_t);
    int32_t *best = xcalloc((size_t)n + 1, sizeof(int32_t));
    int32_t *sel = xcalloc((size_t)2 * (initial_swaps > restart_swaps ? initial_swaps : restart_swaps) + 2,
                           sizeof(int32_t));
    uint64_t rng = seed;
    memcpy(best, map, bytes);
    int64_t initial = orc_score_opt(g, map), best_score = initial;
    if (initial_swaps > 0 && orc_perturb(g, map, initial_swaps, 0, &rng, sel)) return -1;
    int64_t cur = orc_score_opt(g, map);
    int64_t pass = 0, improvements = 0, last = 0, passes = 0, restarts = 0, accepted = 0;
    while (passes < max_passes) {
        for (int m1 = 1; m1 <= n; m1++) {
            int maxdelta = -1, maxv = -1;
            for (int m2 = m1 + 1; m2 <= n; m2++) {
                int d = orc_delta_b(g, map, m1, m2);
                if (d > maxdelta) {
                    maxdelta = d;
                    maxv = m2;
                }
            }
            if (maxdelta > 0) {
                int32_t t = map[m1];
                map[m1] = map[maxv];
                map[maxv] = t;
                cur += maxdelta;
                improvements++;
                accepted++;
                int64_t verify = orc_score_opt(g, map);
                if (verify != cur) cur = verify;
                if (cur > best_score) {
                    best_score = cur;
                    memcpy(best, map, bytes);
                }
            }
        }
        pass++;
        passes++;
        if (pass > idle_passes && improvements - last == 0) {
            memcpy(map, best, bytes);
            if (orc_perturb(g, map, restart_swaps, 0, &rng, sel)) return -1;
            cur = orc_score_opt(g, map);
            improvements = last = pass = 0;
            restarts++;
        }
        if (last < improvements) last = improvements;
    }
    if (out) {
        out[0] = passes;
        out[1] = restarts;
        out[2] = accepted;
        out[3] = initial;
        out[4] = cur;
        out[5] = best_score;
    }
    memcpy(map, best, bytes);
    free(best);
    free(sel);
    return 0;
}

/* quick.cu:151-196 pair numbering and the scan of :1172-1191 */
typedef struct {
    int32_t *pos, *zero;
    int64_t npos, nzero, pairs;
    int32_t best_a, best_b, max_delta;
} orc_qscan;

static void orc_qscan_pairs(const orc_graphs *g, const int32_t *map, const int32_t *changed, int nc, int kind,
                            orc_qscan *q) {
    int n = g->n;
    q->npos = q->nzero = 0;
    q->max_delta = 0;
    q->best_a = q->best_b = 0;
    q->pairs += (int64_t)nc * (nc - 1) / 2 + (int64_t)nc * (n - nc); /* quick.cu:154 */
    uint8_t *in_set = xcalloc((size_t)n + 1, 1);
    for (int i = 0; i < nc; i++) in_set[changed[i]] = 1;
    for (int pass = 0; pass < 2; pass++)
        for (int i = 0; i < nc; i++)
            for (int v = 1; v <= n; v++) {
                int a = changed[i], b;
                if (pass == 0) {
                    /* pairs inside the changed set: i < j in sorted order */
                    if (!in_set[v] || v <= a) continue;
                    b = v;
                } else {
                    if (in_set[v]) continue;
                    b = v;
                }
                int d = kind ? orc_delta_g2(g, map, a, b) : orc_delta_b(g, map, a, b);
                if (d > q->max_delta) {
                    q->max_delta = d;
                    q->best_a = a;
                    q->best_b = b;
                }
                if (d > 0) {
                    q->pos[2 * q->npos] = a;
                    q->pos[2 * q->npos + 1] = b;
                    q->npos++;
                } else if (d == 0) {
                    q->zero[2 * q->nzero] = a;
                    q->zero[2 * q->nzero + 1] = b;
                    q->nzero++;
                }
            }
    free(in_set);
}

static void orc_sort_ints(int32_t *a, int m) {
    for (int i = 1; i < m; i++)
        for (int j = i; j > 0 && a[j - 1] > a[j]; j--) {
            int32_t t = a[j];
            a[j] = a[j - 1];
            a[j - 1] = t;
        }
}

/* quick.cu:1137-1324, one process */
int orc_quick(const orc_graphs *g, int32_t *map, uint64_t seed, int64_t max_evals, int zero_limit,
              int best_window, double gap, double gap_norm, int kind, int64_t *out) {
    int n = g->n;
    if (n < 8) return -1;
    size_t bytes = ((size_t)n + 1) * sizeof(int32_t);
    int32_t *best = xcalloc((size_t)n + 1, sizeof(int32_t));
    orc_qscan q;
    q.pairs = 0;
    q.pos = xcalloc((size_t)16 * n + 2, sizeof(int32_t));
    q.zero = xcalloc((size_t)16 * n + 2, sizeof(int32_t));
    uint64_t rng = seed;
    memcpy(best, map, bytes);
    int64_t cur = orc_score_opt(g, map), initial = cur, best_score = cur;
    int64_t last = 0, iter = 0, perturbations = 0, moves = 0, pos_moves = 0, zero_moves = 0;
    while (iter < max_evals) {
        int k = 3 + (int)(orc_rng_next(&rng) % 2);
        if (iter - last < best_window || (double)(best_score - cur) / gap_norm > gap) memcpy(map, best, bytes);
        int32_t changed[8];
        orc_perturb(g, map, k, 1, &rng, changed);
        orc_sort_ints(changed, 2 * k);
        cur = orc_score_opt(g, map);
        perturbations++;
        orc_qscan_pairs(g, map, changed, 2 * k, kind, &q);
        iter++;
        int zerocheck = 0;
        while ((q.max_delta > 0 || q.nzero > 0) && zerocheck < zero_limit && iter < max_evals) {
            int m1, m2;
            if (q.max_delta > 0) {
                if (orc_rng_next(&rng) % 2 == 0) {
                    int64_t pick = (int64_t)(orc_rng_next(&rng) % (uint64_t)q.npos);
                    m1 = q.pos[2 * pick];
                    m2 = q.pos[2 * pick + 1];
                } else {
                    m1 = q.best_a;
                    m2 = q.best_b;
                }
                zerocheck = 0;
                pos_moves++;
            } else {
                int64_t pick = (int64_t)(orc_rng_next(&rng) % (uint64_t)q.nzero);
                m1 = q.zero[2 * pick];
                m2 = q.zero[2 * pick + 1];
                zerocheck++;
                zero_moves++;
            }
            int32_t t = map[m1];
            map[m1] = map[m2];
            map[m2] = t;
            cur = orc_score_opt(g, map);
            moves++;
            if (cur > best_score) {
                best_score = cur;
                memcpy(best, map, bytes);
                last = iter;
            }
            changed[0] = m1 < m2 ? m1 : m2;
            changed[1] = m1 < m2 ? m2 : m1;
            orc_qscan_pairs(g, map, changed, 2, kind, &q);
            iter++;
        }
    }
    if (out) {
        out[0] = iter;
        out[1] = perturbations;
        out[2] = moves;
        out[3] = pos_moves;
        out[4] = zero_moves;
        out[5] = initial;
        out[6] = cur;
        out[7] = best_score;
        out[8] = q.pairs;
    }
    memcpy(map, best, bytes);
    free(best);
    free(q.pos);
    free(q.zero);
    return 0;
}

/* ------------------------------------------------------------------------
 * q3.cu changed-vertex triple search
 */

/* q3.cu:59-150 with fillArray :982-996 */
int64_t orc_q3_triples(int n, const int32_t *chg, int k, int32_t *out) {
    int32_t *rest = xcalloc((size_t)n + 1, sizeof(int32_t));
    uint8_t *in = xcalloc((size_t)n + 2, 1);
    for (int i = 0; i < k; i++) in[chg[i]] = 1;
    int nr = 0;
    for (int v = 1; v <= n; v++)
        if (!in[v]) rest[nr++] = v;
    int64_t c = 0;
#define ORC_Q3_PUT(x, y, z)                  \
    do {                                     \
        if (out) {                           \
            out[3 * c] = (x);                \
            out[3 * c + 1] = (y);            \
            out[3 * c + 2] = (z);            \
        }                                    \
        c++;                                 \
    } while (0)
    for (int i = 0; i < k; i++) /* first section */
        for (int j = i + 1; j < k; j++)
            for (int l = j + 1; l < k; l++) ORC_Q3_PUT(chg[i], chg[j], chg[l]);
    for (int r = 0; r < nr; r++) /* second: idx / pairs = unchanged index, idx % pairs = the pair */
        for (int i = 0; i < k; i++)
            for (int j = i + 1; j < k; j++) ORC_Q3_PUT(chg[i], chg[j], rest[r]);
    for (int i = 0; i < k; i++) /* third: idx / (nr - 1) = changed index, idx % (nr - 1) = t */
        for (int t = 0; t + 1 < nr; t++) ORC_Q3_PUT(chg[i], rest[t], rest[t + 1]);
#undef ORC_Q3_PUT
    free(rest);
    free(in);
    return c;
}

typedef struct {
    int64_t count, npos, nzero, best_idx, triples;
    int32_t best_delta;
    int32_t *tri, *d, *p;
    int64_t *pos, *zero;
} orc_q3scan;

/* evaluate every enumerated triple; best = first maximum (from -1,
 * :1268-1273), positives and zeros in index order (:1276-1288) */
static void orc_q3_scan(const orc_graphs *g, const int32_t *map, const int32_t *chg, int k, orc_q3scan *q) {
    q->count = orc_q3_triples(g->n, chg, k, q->tri);
    q->best_delta = -1;
    q->best_idx = -1;
    q->npos = q->nzero = 0;
    for (int64_t i = 0; i < q->count; i++) {
        int bp;
        int d = orc_best_g3(g, map, q->tri[3 * i], q->tri[3 * i + 1], q->tri[3 * i + 2], 1, &bp);
        q->d[i] = d;
        q->p[i] = bp;
        if (d > q->best_delta) {
            q->best_delta = d;
            q->best_idx = i;
        }
        if (d > 0) q->pos[q->npos++] = i;
        if (d == 0) q->zero[q->nzero++] = i;
    }
    q->triples += q->count;
}

/* q3.cu:1176-1440 */
int orc_q3(const orc_graphs *g, int32_t *map, uint64_t seed, int64_t max_evals, int zero_limit, int best_window,
           double gap, double gap_norm, int64_t *out) {
    const int n = g->n;
    if (n < 8) return -1;
    const size_t bytes = ((size_t)n + 1) * sizeof(int32_t);
    const int64_t cap = orc_q3_triples(n, (const int32_t[]){1, 2, 3, 4, 5, 6, 7, 8}, 8, NULL) + 8;
    orc_q3scan q;
    memset(&q, 0, sizeof q);
    q.tri = xcalloc((size_t)cap * 3, sizeof(int32_t));
    q.d = xcalloc((size_t)cap, sizeof(int32_t));
    q.p = xcalloc((size_t)cap, sizeof(int32_t));
    q.pos = xcalloc((size_t)cap, sizeof(int64_t));
    q.zero = xcalloc((size_t)cap, sizeof(int64_t));
    int32_t *best = xcalloc((size_t)n + 1, sizeof(int32_t));
    uint64_t rng = seed;
    memcpy(best, map, bytes);
    int64_t cur = orc_score_opt(g, map), initial = cur, best_score = cur;
    const int64_t last = 0; /* never updated in q3.cu */
    int64_t iter = 0, perturbations = 0, moves = 0, pos_moves = 0, zero_moves = 0;
    while (iter < max_evals) {
        const int k = 3 + (int)(orc_rng_next(&rng) % 2); /* :1189 */
        if (iter - last < best_window || (double)(best_score - cur) / gap_norm > gap) memcpy(map, best, bytes);
        int32_t chg[8];
        orc_perturb(g, map, k, 1, &rng, chg);
        orc_sort_ints(chg, 2 * k);
        cur = orc_score_opt(g, map);
        perturbations++;
        orc_q3_scan(g, map, chg, 2 * k, &q);
        iter++;
        int zerocheck = 0;
        while ((q.best_delta > 0 || q.nzero > 0) && zerocheck < zero_limit && iter < max_evals) {
            int64_t t;
            if (q.best_delta > 0) {
                if (orc_rng_next(&rng) % 2 == 0) t = q.pos[orc_rng_next(&rng) % (uint64_t)q.npos];
                else t = q.best_idx;
                orc_apply_g3(map, q.tri[3 * t], q.tri[3 * t + 1], q.tri[3 * t + 2], q.p[t]);
                cur = orc_score_opt(g, map); /* :1293 */
                zerocheck = 0;
                pos_moves++;
            } else {
                t = q.zero[orc_rng_next(&rng) % (uint64_t)q.nzero];
                orc_apply_g3(map, q.tri[3 * t], q.tri[3 * t + 1], q.tri[3 * t + 2], q.p[t]);
                zerocheck++;
                zero_moves++;
            }
            moves++;
            if (cur > best_score) {
                best_score = cur;
                memcpy(best, map, bytes);
            }
            int32_t m3[3] = {q.tri[3 * t], q.tri[3 * t + 1], q.tri[3 * t + 2]};
            orc_sort_ints(m3, 3);
            orc_q3_scan(g, map, m3, 3, &q);
            iter++;
        }
    }
    if (out) {
        out[0] = iter;
        out[1] = perturbations;
        out[2] = moves;
        out[3] = pos_moves;
        out[4] = zero_moves;
        out[5] = initial;
        out[6] = cur;
        out[7] = best_score;
        out[8] = q.triples;
    }
    memcpy(map, best, bytes);
    free(best);
    free(q.tri);
    free(q.d);
    free(q.p);
    free(q.pos);
    free(q.zero);
    return 0;
}

/* rank -> triple by counting (O(n) per triple; tests use small n) */
int64_t orc_lex_triples(int n, int64_t start, int64_t count, int32_t *out) {
    int64_t got = 0;
    for (int64_t t = start; got < count; t++) {
        int64_t r = t;
        int m1 = 1, m2 = 0, m3 = 0, ok = 0;
        for (; m1 <= n - 2 && !ok; m1++) {
            const int64_t with_m1 = (int64_t)(n - m1) * (n - m1 - 1) / 2;
            if (r >= with_m1) {
                r -= with_m1;
                continue;
            }
            for (m2 = m1 + 1; m2 <= n - 1; m2++) {
                if (r >= n - m2) {
                    r -= n - m2;
                    continue;
                }
                m3 = m2 + 1 + (int)r;
                ok = 1;
                break;
            }
            if (ok) break;
        }
        if (!ok) break;
        if (out) {
            out[3 * got] = m1;
            out[3 * got + 1] = m2;
            out[3 * got + 2] = m3;
        }
        got++;
    }
    return got;
}

/* triple.cu:941-1196 */
int orc_triples_pass(const orc_graphs *g, int32_t *map, uint64_t seed, int64_t max_evals, int batch, int64_t *out) {
    const int n = g->n;
    if (n < 10 || batch < 1 || max_evals < 1) return -1;
    const size_t bytes = ((size_t)n + 1) * sizeof(int32_t);
    const int64_t total = (int64_t)n * (n - 1) * (n - 2) / 6;
    const int64_t B = batch < total ? batch : total;
    int32_t *tri = xcalloc((size_t)B * 3, sizeof(int32_t));
    int32_t *best = xcalloc((size_t)n + 1, sizeof(int32_t));
    uint64_t rng = seed;
    memcpy(best, map, bytes);
    int64_t cur = orc_score_opt(g, map), initial = cur, best_score = cur;
    int64_t evals = 0, passes = 0, perturbations = 0, moves = 0, triples = 0;
    while (evals < max_evals) {
        int found = 0;
        int64_t start = 0;
        passes++;
        for (; start < total && evals < max_evals; start += B) {
            const int64_t m = orc_lex_triples(n, start, B, tri);
            evals++;
            triples += m;
            int best_delta = 0, best_p = 0;
            int64_t best_idx = -1;
            for (int64_t i = 0; i < m; i++) { /* :1136-1143: the first index of the largest delta */
                int p;
                const int d = orc_best_g3(g, map, tri[3 * i], tri[3 * i + 1], tri[3 * i + 2], 1, &p);
                if (d > best_delta) {
                    best_delta = d;
                    best_idx = i;
                    best_p = p;
                }
            }
            if (best_delta > 0) { /* :1148-1168 */
                orc_apply_g3(map, tri[3 * best_idx], tri[3 * best_idx + 1], tri[3 * best_idx + 2], best_p);
                cur += best_delta;
                moves++;
                if (cur > best_score) {
                    memcpy(best, map, bytes);
                    best_score = cur;
                }
                found = 1;
                break;
            }
        }
        if (!found && start >= total) { /* :1174-1181 */
            memcpy(map, best, bytes);
            const int k = 3 + (int)(orc_rng_next(&rng) % 3);
            int32_t sel[16];
            orc_perturb(g, map, k, 0, &rng, sel);
            cur = orc_score_opt(g, map);
            perturbations++;
        }
    }
    if (out) {
        out[0] = evals;
        out[1] = passes;
        out[2] = perturbations;
        out[3] = moves;
        out[4] = initial;
        out[5] = cur;
        out[6] = best_score;
        out[7] = triples;
    }
    memcpy(map, best, bytes);
    free(best);
    free(tri);
    return 0;
}
