/*
 * refine-matching-hillclimb -- drop-in for src/refine-matching-hillclimb.c
 * (zhihuidu/brain-match; README name refine-match-hillclimb) with the
 * pair-swap search on the MI355X engine.
 *
 *   refine-matching-hillclimb <g1.csv> <g2.csv> <initial_matching.csv> <output_matching.csv>
 *
 * Same arguments, formats and messages as the reference (:508-574):
 * vertex ids of each graph are indexed in first-appearance order
 * (:193-207, :427-428), edges are de-duplicated with the last weight
 * winning (insert_edge :101-125), the matching is "%d,%d" lines
 * (:471-488) and the output is "Male ID,Female ID" with the original ids in
 * G1 index order (:494-506).  The search is a pair-swap hill-climb run to
 * convergence on the GPU: first-improvement sweeps with the exact swap delta
 * (Delta_B over the de-duplicated graphs; the rescore after every accept
 * keeps the score exact when a vertex has a self-loop).  The reference's
 * 8-thread simulated annealing is time-seeded and, built from source, hangs
 * in its edge hash table (:113-115) or overruns its matching array
 * (:301-302) before optimising, so it is not reproduced.
 * BM_DELTA=hillclimb accepts on the reference's out-edge-only delta
 * (calculate_swap_delta, :244-296) instead.  Environment: BM_DEVICE,
 * BM_DELTA, BM_MAX_SWEEPS.
 */
#include "cli_common.h"

typedef struct {
    int32_t *id_to_idx; /* id -> 1-based index, 0 = unseen */
    int32_t *idx_to_id; /* 1-based */
    int32_t cap_ids, count, cap_idx;
} idmap_t;

static int32_t idmap_add(idmap_t *m, int32_t id) {
    if (id < 0) return 0;
    if (id >= m->cap_ids) {
        int32_t nc = m->cap_ids ? m->cap_ids : 1024;
        while (nc <= id) nc *= 2;
        m->id_to_idx = realloc(m->id_to_idx, (size_t)nc * 4);
        memset(m->id_to_idx + m->cap_ids, 0, (size_t)(nc - m->cap_ids) * 4);
        m->cap_ids = nc;
    }
    if (!m->id_to_idx[id]) {
        if (m->count + 2 > m->cap_idx) {
            m->cap_idx = m->cap_idx ? 2 * m->cap_idx : 1024;
            m->idx_to_id = realloc(m->idx_to_id, (size_t)m->cap_idx * 4);
        }
        m->id_to_idx[id] = ++m->count;
        m->idx_to_id[m->count] = id;
    }
    return m->id_to_idx[id];
}

static int32_t idmap_get(const idmap_t *m, int32_t id) {
    return (id >= 0 && id < m->cap_ids) ? m->id_to_idx[id] : 0;
}

static void handle_error(const char *msg) {
    fprintf(stderr, "Error: %s\n", msg);
    exit(EXIT_FAILURE);
}

/* load_graph, :405-453: ids re-indexed, the last weight of a pair wins */
static void load_graph(const char *path, idmap_t *ids, bm_edges *g) {
    FILE *f = fopen(path, "r");
    if (!f) handle_error("Cannot open graph file");
    char line[1024];
    int empty = !fgets(line, sizeof line, f);
    fclose(f);
    if (empty) handle_error("Empty graph file");
    bm_edges raw;
    if (bm_edges_read_csv(path, BM_CSV_WARN_LINE, &raw)) handle_error("Cannot open graph file");
    for (int64_t i = 0; i < raw.m; i++) {
        raw.src[i] = idmap_add(ids, raw.src[i]);
        raw.dst[i] = idmap_add(ids, raw.dst[i]);
    }
    *g = raw;
}

/* the hash table keeps one entry per (src, dst): last weight, first slot */
static void dedup(bm_edges *g, int32_t n) {
    int64_t *last = malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t *ptr = calloc((size_t)n + 2, sizeof(int64_t));
    int64_t *ord = malloc((size_t)(g->m ? g->m : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < g->m; i++) ptr[g->src[i] + 1]++;
    for (int32_t v = 0; v <= n; v++) ptr[v + 1] += ptr[v];
    int64_t *pos = malloc(((size_t)n + 1) * sizeof(int64_t));
    for (int32_t v = 0; v <= n; v++) pos[v] = ptr[v];
    for (int64_t i = 0; i < g->m; i++) ord[pos[g->src[i]]++] = i;
    for (int32_t v = 0; v <= n; v++) last[v] = -1;
    int64_t k = 0;
    int32_t *s = malloc((size_t)(g->m ? g->m : 1) * 4), *d = malloc((size_t)(g->m ? g->m : 1) * 4),
            *w = malloc((size_t)(g->m ? g->m : 1) * 4);
    for (int32_t u = 0; u <= n; u++) {
        for (int64_t j = ptr[u]; j < ptr[u + 1]; j++) last[g->dst[ord[j]]] = ord[j];
        for (int64_t j = ptr[u]; j < ptr[u + 1]; j++) {
            int64_t i = ord[j];
            if (last[g->dst[i]] < 0) continue;
            s[k] = u;
            d[k] = g->dst[i];
            w[k] = g->w[last[g->dst[i]]];
            last[g->dst[i]] = -1;
            k++;
        }
    }
    free(g->src);
    free(g->dst);
    free(g->w);
    g->src = s;
    g->dst = d;
    g->w = w;
    g->m = k;
    free(last);
    free(ptr);
    free(ord);
    free(pos);
}

/* first-improvement sweep accepting on Delta_A */
static int hillclimb_compat(bm_ctx *ctx, int32_t n, int64_t max_sweeps) {
    int32_t *a = malloc((size_t)n * 4), *b = malloc((size_t)n * 4), *d = malloc((size_t)n * 4);
    for (int64_t sweep = 0; max_sweeps <= 0 || sweep < max_sweeps; sweep++) {
        int64_t acc = 0;
        for (int32_t x = 1; x <= n; x++) {
            int32_t cursor = 1;
            while (cursor <= n) {
                int64_t m = 0;
                for (int32_t y = cursor; y <= n; y++)
                    if (y != x) {
                        a[m] = x;
                        b[m] = y;
                        m++;
                    }
                if (!m) break;
                int rc = bm_eval_swaps_hillclimb(ctx, a, b, m, d);
                if (rc) return rc;
                int64_t j = 0;
                while (j < m && d[j] <= 0) j++;
                if (j == m) break;
                if ((rc = bm_apply_swap(ctx, x, b[j]))) return rc;
                acc++;
                cursor = b[j] + 1;
            }
        }
        if (!acc) break;
    }
    free(a);
    free(b);
    free(d);
    return BM_OK;
}

int main(int argc, char *argv[]) {
    if (argc != 5) {
        fprintf(stderr, "Usage: %s <g1.csv> <g2.csv> <initial_matching.csv> <output_matching.csv>\n", argv[0]);
        return EXIT_FAILURE;
    }
    idmap_t m1 = {0}, m2 = {0};
    bm_edges g1, g2;
    printf("Loading graph 1...\n");
    load_graph(argv[1], &m1, &g1);
    printf("Loading graph 2...\n");
    load_graph(argv[2], &m2, &g2);
    if (m1.count != m2.count) {
        fprintf(stderr, "Error: Graphs have different numbers of nodes (G1: %d, G2: %d)\n", m1.count, m2.count);
        return EXIT_FAILURE;
    }
    int32_t n = m1.count > 1 ? m1.count : 1;
    printf("Loading initial matching...\n");
    FILE *f = fopen(argv[3], "r");
    if (!f) handle_error("Cannot open initial matching file");
    char line[1024];
    int empty = !fgets(line, sizeof line, f);
    fclose(f);
    if (empty) handle_error("Empty matching file");
    bm_pairs pairs;
    bm_pairs_read_plain(argv[3], &pairs);
    int32_t *mapping = calloc((size_t)n + 1, 4);
    for (int64_t i = 0; i < pairs.m; i++) {
        int32_t x = idmap_get(&m1, pairs.male[i]), y = idmap_get(&m2, pairs.female[i]);
        if (x > 0 && y > 0) mapping[x] = y;
    }
    bm_pairs_free(&pairs);
    dedup(&g1, n);

    bm_ctx *ctx = cli_engine(n, &g1, &g2);
    if (!ctx || bm_set_mapping(ctx, mapping)) return EXIT_FAILURE;
    int64_t initial = 0, final = 0;
    bm_score(ctx, &initial);
    printf("Initial matching score: %lld\n", (long long)initial);
    printf("Optimizing matching using GPU device %lld...\n", (long long)env_i64("BM_DEVICE", 0));
    fflush(stdout);
    const char *mode = getenv("BM_DELTA");
    int rc;
    if (mode && strcmp(mode, "hillclimb") == 0) {
        rc = hillclimb_compat(ctx, n, env_i64("BM_MAX_SWEEPS", 0));
    } else {
        bm_drive_stats st;
        rc = bm_drive_first_improvement(ctx, n, env_i64("BM_MAX_SWEEPS", 0), 1, NULL, NULL, NULL, 0, &st);
    }
    if (rc) {
        fprintf(stderr, "Error: engine: %s (%s)\n", bm_strerror(rc), bm_last_error());
        return EXIT_FAILURE;
    }
    bm_score(ctx, &final);
    printf("Final matching score: %lld\n", (long long)final);
    if (initial > 0) printf("Improvement: %.2f%%\n", 100.0 * (final - initial) / initial);
    printf("Writing results to %s...\n", argv[4]);
    bm_get_mapping(ctx, mapping);
    FILE *o = fopen(argv[4], "w");
    if (!o) handle_error("Cannot open output file for writing");
    fprintf(o, "Male ID,Female ID\n");
    for (int32_t i = 1; i <= m1.count; i++)
        if (mapping[i] >= 1 && mapping[i] <= m2.count) fprintf(o, "%d,%d\n", m1.idx_to_id[i], m2.idx_to_id[mapping[i]]);
    fclose(o);
    printf("Cleaning up...\n");
    bm_destroy(ctx);
    bm_edges_free(&g1);
    bm_edges_free(&g2);
    free(mapping);
    free(m1.id_to_idx);
    free(m1.idx_to_id);
    free(m2.id_to_idx);
    free(m2.idx_to_id);
    printf("Done.\n");
    return EXIT_SUCCESS;
}
