/*
 * bm_gen.c -- seeded synthetic graph pairs and move lists.
 *
 * The reference's real inputs (FlyWire VNC gm/gf) are not distributed, so
 * every workload is synthetic.  A project PRNG (splitmix64-seeded
 * xoshiro256**) makes each instance a pure function of its seed; libc rand()
 * is never used.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bm_host.h"

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform in [0, 1) */
static double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
/* uniform in [0, n) */
static uint32_t rng_below(rng_t *r, uint32_t n) { return (uint32_t)(((rng_next(r) >> 32) * (uint64_t)n) >> 32); }

static double rng_normal(rng_t *r) {
    double u1 = 1.0 - rng_unit(r), u2 = rng_unit(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Pareto(alpha) sample with x_min = 1 */
static double powerlaw(rng_t *r, double alpha) { return pow(1.0 - rng_unit(r), -1.0 / (alpha - 1.0)); }

/* geometric(p) failures before the first success, plus 1, capped */
static int32_t geom_weight(rng_t *r, double p, int32_t cap) {
    double u = 1.0 - rng_unit(r); /* (0, 1] */
    double g = floor(log(u) / log(1.0 - p));
    if (g > cap) g = cap;
    int32_t w = 1 + (int32_t)g;
    return w > cap ? cap : w;
}

static int32_t lognormal_weight(rng_t *r, double mu, double sigma, int32_t lo, int32_t hi) {
    double x = exp(mu + sigma * rng_normal(r));
    double w = floor(x + 0.5);
    if (w < lo) w = lo;
    if (w > hi) w = hi;
    return (int32_t)w;
}

/* ------------------------------------------------------------ hash set */

typedef struct {
    uint64_t *k;
    uint64_t mask;
} hset_t;

static int hset_init(hset_t *h, int64_t expect) {
    uint64_t cap = 1024;
    while (cap < (uint64_t)expect * 2 + 16) cap <<= 1;
    h->k = malloc(cap * sizeof(uint64_t));
    if (!h->k) return BM_ENOMEM;
    memset(h->k, 0xFF, cap * sizeof(uint64_t));
    h->mask = cap - 1;
    return BM_OK;
}

static uint64_t hmix(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    return x;
}

/* returns 1 if inserted, 0 if present */
static int hset_add(hset_t *h, uint64_t key) {
    uint64_t i = hmix(key) & h->mask;
    while (h->k[i] != UINT64_MAX) {
        if (h->k[i] == key) return 0;
        i = (i + 1) & h->mask;
    }
    h->k[i] = key;
    return 1;
}

static void hset_free(hset_t *h) { free(h->k); }

/* ------------------------------------------------------------ helpers */

void bm_instance_free(bm_instance *in) {
    free(in->s1); free(in->d1); free(in->w1);
    free(in->s2); free(in->d2); free(in->w2);
    free(in->pi); free(in->p0);
    memset(in, 0, sizeof *in);
}

static int alloc_inst(bm_instance *in, int32_t n, int64_t m1, int64_t m2cap) {
    memset(in, 0, sizeof *in);
    in->n = n;
    in->s1 = malloc((size_t)m1 * 4); in->d1 = malloc((size_t)m1 * 4); in->w1 = malloc((size_t)m1 * 4);
    in->s2 = malloc((size_t)m2cap * 4); in->d2 = malloc((size_t)m2cap * 4); in->w2 = malloc((size_t)m2cap * 4);
    in->pi = calloc((size_t)n + 1, 4);
    in->p0 = calloc((size_t)n + 1, 4);
    if (!in->s1 || !in->d1 || !in->w1 || !in->s2 || !in->d2 || !in->w2 || !in->pi || !in->p0) {
        bm_instance_free(in);
        return BM_ENOMEM;
    }
    return BM_OK;
}

/* pi: a uniform bijection of 1..n; p0: pi with a shuffled fraction */
static void make_mappings(rng_t *r, bm_instance *in, double shuffle_frac) {
    int32_t n = in->n;
    in->pi[0] = 0;
    for (int32_t v = 1; v <= n; v++) in->pi[v] = v;
    for (int32_t v = n; v > 1; v--) {
        int32_t j = 1 + (int32_t)rng_below(r, (uint32_t)v);
        int32_t t = in->pi[v];
        in->pi[v] = in->pi[j];
        in->pi[j] = t;
    }
    memcpy(in->p0, in->pi, ((size_t)n + 1) * 4);
    int32_t k = (int32_t)floor(shuffle_frac * n + 0.5);
    if (k < 2) return;
    int32_t *idx = malloc((size_t)n * 4);
    if (!idx) return;
    for (int32_t i = 0; i < n; i++) idx[i] = i + 1;
    for (int32_t i = 0; i < k; i++) { /* partial Fisher-Yates: the chosen vertices */
        int32_t j = i + (int32_t)rng_below(r, (uint32_t)(n - i));
        int32_t t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
    }
    for (int32_t i = k - 1; i > 0; i--) { /* permute their images */
        int32_t j = (int32_t)rng_below(r, (uint32_t)(i + 1));
        int32_t t = in->p0[idx[i]];
        in->p0[idx[i]] = in->p0[idx[j]];
        in->p0[idx[j]] = t;
    }
    free(idx);
}

/* delete `ndel` random lines of G2 (order of survivors kept) */
static int64_t delete_random(rng_t *r, int32_t *s, int32_t *d, int32_t *w, int64_t m, int64_t ndel, hset_t *keys,
                             int32_t n) {
    if (ndel <= 0) return m;
    char *dead = calloc((size_t)m, 1);
    int64_t *idx = malloc((size_t)m * sizeof(int64_t));
    if (!dead || !idx) {
        free(dead);
        free(idx);
        return m;
    }
    for (int64_t i = 0; i < m; i++) idx[i] = i;
    for (int64_t i = 0; i < ndel; i++) {
        int64_t j = i + (int64_t)(rng_unit(r) * (double)(m - i));
        if (j >= m) j = m - 1;
        int64_t t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
        dead[idx[i]] = 1;
    }
    int64_t k = 0;
    for (int64_t i = 0; i < m; i++) {
        if (dead[i]) continue;
        s[k] = s[i];
        d[k] = d[i];
        w[k] = w[i];
        k++;
    }
    free(dead);
    free(idx);
    /* rebuild the key set of the survivors */
    memset(keys->k, 0xFF, (keys->mask + 1) * sizeof(uint64_t));
    for (int64_t i = 0; i < k; i++) hset_add(keys, (uint64_t)s[i] * (uint64_t)(n + 1) + (uint64_t)d[i]);
    return k;
}

/* ------------------------------------------------------------ power law */

int bm_gen_powerlaw(uint64_t seed, int32_t n, int64_t m, double del_frac, double shuffle_frac, bm_instance *out) {
    if (n < 3 || m < n || m > (int64_t)n * (n - 1) / 2) return BM_EINVAL;
    rng_t r;
    rng_seed(&r, seed);
    int rc = alloc_inst(out, n, m, m);
    if (rc) return rc;
    /* out-degrees: discrete power law, exponent 2.1, min 1, cap n-1, scaled to m */
    int64_t *deg = malloc(((size_t)n + 1) * sizeof(int64_t));
    double *raw = malloc(((size_t)n + 1) * sizeof(double));
    int32_t *stamp = calloc((size_t)n + 1, 4);
    if (!deg || !raw || !stamp) {
        free(deg); free(raw); free(stamp);
        bm_instance_free(out);
        return BM_ENOMEM;
    }
    double sum = 0;
    for (int32_t u = 1; u <= n; u++) {
        double x = floor(powerlaw(&r, 2.1));
        if (x > n - 1) x = n - 1;
        raw[u] = x;
        sum += x;
    }
    int64_t tot = 0;
    for (int32_t u = 1; u <= n; u++) {
        int64_t dd = (int64_t)floor(raw[u] * (double)m / sum + 0.5);
        if (dd < 1) dd = 1;
        if (dd > n - 1) dd = n - 1;
        deg[u] = dd;
        tot += dd;
    }
    while (tot < m) {
        int32_t u = 1 + (int32_t)rng_below(&r, (uint32_t)n);
        if (deg[u] < n - 1) { deg[u]++; tot++; }
    }
    while (tot > m) {
        int32_t u = 1 + (int32_t)rng_below(&r, (uint32_t)n);
        if (deg[u] > 1) { deg[u]--; tot--; }
    }
    /* targets uniform over V\{u}, no repeats; weights geometric(0.3)+1 <= 255 */
    int64_t e = 0;
    for (int32_t u = 1; u <= n; u++) {
        stamp[u] = u;
        for (int64_t k = 0; k < deg[u]; k++) {
            int32_t v;
            do v = 1 + (int32_t)rng_below(&r, (uint32_t)n);
            while (stamp[v] == u);
            stamp[v] = u;
            out->s1[e] = u;
            out->d1[e] = v;
            out->w1[e] = geom_weight(&r, 0.3, 255);
            e++;
        }
    }
    out->e1 = e;
    free(deg);
    free(raw);
    free(stamp);
    make_mappings(&r, out, shuffle_frac);
    /* G2 = pi(G1), then delete a fraction and add as many fresh lines */
    hset_t keys;
    if ((rc = hset_init(&keys, m))) {
        bm_instance_free(out);
        return rc;
    }
    for (int64_t i = 0; i < e; i++) {
        out->s2[i] = out->pi[out->s1[i]];
        out->d2[i] = out->pi[out->d1[i]];
        out->w2[i] = out->w1[i];
        hset_add(&keys, (uint64_t)out->s2[i] * (uint64_t)(n + 1) + (uint64_t)out->d2[i]);
    }
    int64_t ndel = (int64_t)floor(del_frac * (double)e + 0.5);
    int64_t k = delete_random(&r, out->s2, out->d2, out->w2, e, ndel, &keys, n);
    while (k < e) {
        int32_t u = 1 + (int32_t)rng_below(&r, (uint32_t)n), v = 1 + (int32_t)rng_below(&r, (uint32_t)n);
        if (u == v) continue;
        if (!hset_add(&keys, (uint64_t)u * (uint64_t)(n + 1) + (uint64_t)v)) continue;
        out->s2[k] = u;
        out->d2[k] = v;
        out->w2[k] = geom_weight(&r, 0.3, 255);
        k++;
    }
    out->e2 = k;
    hset_free(&keys);
    return BM_OK;
}

/* ------------------------------------------------------------ Chung-Lu */

typedef struct {
    double *prob;
    int32_t *alias;
    int32_t n;
} alias_t;

/* Vose alias table over weights wt[1..n] */
static int alias_build(alias_t *a, const double *wt, int32_t n) {
    a->n = n;
    a->prob = malloc((size_t)n * sizeof(double));
    a->alias = malloc((size_t)n * 4);
    int32_t *small = malloc((size_t)n * 4), *large = malloc((size_t)n * 4);
    double *p = malloc((size_t)n * sizeof(double));
    if (!a->prob || !a->alias || !small || !large || !p) {
        free(small); free(large); free(p);
        return BM_ENOMEM;
    }
    double sum = 0;
    for (int32_t i = 0; i < n; i++) sum += wt[i + 1];
    int32_t ns = 0, nl = 0;
    for (int32_t i = 0; i < n; i++) {
        p[i] = wt[i + 1] * n / sum;
        if (p[i] < 1.0) small[ns++] = i;
        else large[nl++] = i;
    }
    while (ns && nl) {
        int32_t s = small[--ns], l = large[--nl];
        a->prob[s] = p[s];
        a->alias[s] = l;
        p[l] = (p[l] + p[s]) - 1.0;
        if (p[l] < 1.0) small[ns++] = l;
        else large[nl++] = l;
    }
    while (nl) { int32_t l = large[--nl]; a->prob[l] = 1.0; a->alias[l] = l; }
    while (ns) { int32_t s = small[--ns]; a->prob[s] = 1.0; a->alias[s] = s; }
    free(small); free(large); free(p);
    return BM_OK;
}

static int32_t alias_draw(const alias_t *a, rng_t *r) {
    int32_t i = (int32_t)rng_below(r, (uint32_t)a->n);
    return 1 + (rng_unit(r) < a->prob[i] ? i : a->alias[i]);
}

static void alias_free(alias_t *a) { free(a->prob); free(a->alias); }

/* power-law vertex weights, capped so no expected degree exceeds cap */
static void chunglu_weights(rng_t *r, double *th, int32_t n, int64_t m, double cap) {
    for (int32_t u = 1; u <= n; u++) th[u] = powerlaw(r, 2.1);
    for (int it = 0; it < 8; it++) {
        double sum = 0;
        for (int32_t u = 1; u <= n; u++) sum += th[u];
        double lim = cap * sum / (double)m;
        int changed = 0;
        for (int32_t u = 1; u <= n; u++)
            if (th[u] > lim) { th[u] = lim; changed = 1; }
        if (!changed) break;
    }
}

int bm_gen_chunglu(uint64_t seed, int32_t n, int64_t m, double rewire_frac, double reweight_frac,
                   double shuffle_frac, bm_instance *out) {
    if (n < 3 || m < 1 || m > (int64_t)n * (n - 1) / 4) return BM_EINVAL;
    rng_t r;
    rng_seed(&r, seed);
    int rc = alloc_inst(out, n, m, m);
    if (rc) return rc;
    double *tho = malloc(((size_t)n + 1) * sizeof(double)), *thi = malloc(((size_t)n + 1) * sizeof(double));
    alias_t ao = {0}, ai = {0};
    hset_t k1 = {0}, k2 = {0};
    if (!tho || !thi) { rc = BM_ENOMEM; goto done; }
    double cap = (n - 1) / 4.0;
    chunglu_weights(&r, tho, n, m, cap);
    chunglu_weights(&r, thi, n, m, cap);
    if ((rc = alias_build(&ao, tho, n)) || (rc = alias_build(&ai, thi, n))) goto done;
    if ((rc = hset_init(&k1, m)) || (rc = hset_init(&k2, m))) goto done;
    int64_t e = 0;
    while (e < m) {
        int32_t u = alias_draw(&ao, &r), v = alias_draw(&ai, &r);
        if (u == v) continue;
        if (!hset_add(&k1, (uint64_t)u * (uint64_t)(n + 1) + (uint64_t)v)) continue;
        out->s1[e] = u;
        out->d1[e] = v;
        out->w1[e] = lognormal_weight(&r, 1.0, 1.2, 1, 2000);
        e++;
    }
    out->e1 = e;
    make_mappings(&r, out, shuffle_frac);
    for (int64_t i = 0; i < e; i++) {
        out->s2[i] = out->pi[out->s1[i]];
        out->d2[i] = out->pi[out->d1[i]];
        out->w2[i] = out->w1[i];
        hset_add(&k2, (uint64_t)out->s2[i] * (uint64_t)(n + 1) + (uint64_t)out->d2[i]);
    }
    int64_t ndel = (int64_t)floor(rewire_frac * (double)e + 0.5);
    int64_t k = delete_random(&r, out->s2, out->d2, out->w2, e, ndel, &k2, n);
    int64_t kept = k;
    /* weight noise on a fraction of the surviving lines */
    int64_t nrew = (int64_t)floor(reweight_frac * (double)kept + 0.5);
    for (int64_t i = 0; i < nrew; i++) {
        int64_t j = (int64_t)(rng_unit(&r) * (double)kept);
        if (j >= kept) j = kept - 1;
        out->w2[j] = lognormal_weight(&r, 1.0, 1.2, 1, 2000);
    }
    /* rewired lines: fresh draws of the same model, relabelled by pi */
    while (k < e) {
        int32_t u = alias_draw(&ao, &r), v = alias_draw(&ai, &r);
        if (u == v) continue;
        int32_t pu = out->pi[u], pv = out->pi[v];
        if (!hset_add(&k2, (uint64_t)pu * (uint64_t)(n + 1) + (uint64_t)pv)) continue;
        out->s2[k] = pu;
        out->d2[k] = pv;
        out->w2[k] = lognormal_weight(&r, 1.0, 1.2, 1, 2000);
        k++;
    }
    out->e2 = k;
    rc = BM_OK;
done:
    free(tho);
    free(thi);
    alias_free(&ao);
    alias_free(&ai);
    hset_free(&k1);
    hset_free(&k2);
    if (rc) bm_instance_free(out);
    return rc;
}

/* ------------------------------------------------------------ move lists */

int bm_gen_triples_2hop(uint64_t seed, const bm_instance *in, int64_t count, int32_t *out) {
    int32_t n = in->n;
    if (n < 3 || count < 0 || !out) return BM_EINVAL;
    /* undirected neighbour lists over G1 */
    int64_t *ptr = calloc((size_t)n + 2, sizeof(int64_t));
    int32_t *adj = malloc((size_t)(2 * in->e1 + 1) * 4);
    if (!ptr || !adj) { free(ptr); free(adj); return BM_ENOMEM; }
    for (int64_t i = 0; i < in->e1; i++) { ptr[in->s1[i] + 1]++; ptr[in->d1[i] + 1]++; }
    for (int32_t v = 0; v <= n; v++) ptr[v + 1] += ptr[v];
    int64_t *pos = malloc(((size_t)n + 1) * sizeof(int64_t));
    if (!pos) { free(ptr); free(adj); return BM_ENOMEM; }
    for (int32_t v = 0; v <= n; v++) pos[v] = ptr[v];
    for (int64_t i = 0; i < in->e1; i++) {
        adj[pos[in->s1[i]]++] = in->d1[i];
        adj[pos[in->d1[i]]++] = in->s1[i];
    }
    free(pos);
    rng_t r;
    rng_seed(&r, seed);
    for (int64_t t = 0; t < count; t++) {
        for (;;) {
            int32_t u = 1 + (int32_t)rng_below(&r, (uint32_t)n);
            if (ptr[u + 1] == ptr[u]) continue;
            int32_t pick[2];
            int got = 0;
            for (int tries = 0; tries < 64 && got < 2; tries++) {
                /* a walk of one or two steps from u */
                int64_t du = ptr[u + 1] - ptr[u];
                int32_t x = adj[ptr[u] + (int64_t)rng_below(&r, (uint32_t)du)];
                if (rng_next(&r) & 1) {
                    int64_t dx = ptr[x + 1] - ptr[x];
                    if (dx > 0) x = adj[ptr[x] + (int64_t)rng_below(&r, (uint32_t)dx)];
                }
                if (x == u || x < 1 || (got == 1 && x == pick[0])) continue;
                pick[got++] = x;
            }
            if (got < 2) continue;
            out[3 * t] = u;
            out[3 * t + 1] = pick[0];
            out[3 * t + 2] = pick[1];
            break;
        }
    }
    free(ptr);
    free(adj);
    return BM_OK;
}

int bm_gen_sets(uint64_t seed, int32_t n, int32_t k, int64_t count, int32_t *out) {
    if (k < 1 || k > 8 || n < k || count < 0 || !out) return BM_EINVAL;
    rng_t r;
    rng_seed(&r, seed);
    for (int64_t t = 0; t < count; t++) {
        int32_t *s = out + (size_t)k * t;
        for (int i = 0; i < k; i++) {
            int dup;
            do {
                s[i] = 1 + (int32_t)rng_below(&r, (uint32_t)n);
                dup = 0;
                for (int j = 0; j < i; j++)
                    if (s[j] == s[i]) dup = 1;
            } while (dup);
        }
    }
    return BM_OK;
}
