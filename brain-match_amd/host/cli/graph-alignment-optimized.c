/*
 * graph-alignment-optimized -- drop-in for src/graph-alignment-optimized.c
 * (zhihuidu/brain-match) with the swap-delta loop on the MI355X engine.
 *
 *   graph-alignment-optimized <male graph> <female graph> <male ordering>
 *                             <female ordering> <in mapping> <out mapping>
 *
 * Same arguments, file formats, log lines and cadence as the reference
 * (:682-738): first-improvement pair swaps (:513-581), a progress block
 * every 20 rows (:519-534), the mapping rewritten and re-read every 25 rows
 * (:536-540, :478-492), a device rescore after every accepted swap
 * with the reference's per-swap mismatch line (:570-576).  Differences, all deliberate:
 *   - n is the largest vertex id in the inputs instead of NUM_NODES = 18524
 *     (set BM_ROWS_PER_SWEEP=18524 to reproduce the reference's vertex range
 *     for smaller graphs: its rows, its row counter and the swaps with the
 *     empty vertices above the data, which pay when G2 has lines at vertex 0);
 *   - the run ends after a sweep without an accepted swap (the reference
 *     loops forever, rewriting the same converged mapping);
 *   - the ordering files must exist, as in the reference, but are unused
 *     there too (the similarity test at :549-553 is commented out).
 * Environment: BM_DEVICE, BM_ROWS_PER_SWEEP, BM_MAX_SWEEPS, BM_VERIFY (1).
 */
#include "cli_common.h"

#define SAVE_INTERVAL 25
#define UPDATE_INTERVAL 20

static int64_t count_lines(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[1024];
    int64_t k = 0;
    while (fgets(line, sizeof line, f)) k++;
    fclose(f);
    return k;
}

/* load_graph_from_csv, :635-679 */
static int load_graph(const char *path, bm_edges *g) {
    int64_t total = count_lines(path);
    if (total < 0) {
        LOG_ERROR("Failed to open file: %s", path);
        return -1;
    }
    LOG_INFO("Loading graph from %s (%s lines)", path, fmt_num(total - 1));
    time_t t0 = time(NULL);
    if (bm_edges_read_csv(path, BM_CSV_LOG_ERROR, g)) {
        LOG_ERROR("Failed to open file: %s", path);
        return -1;
    }
    LOG_INFO("Graph loaded successfully:");
    LOG_INFO("  - Nodes: %s", fmt_num(g->max_id < 0 ? 0 : g->max_id));
    LOG_INFO("  - Edges: %s", fmt_num(g->m));
    LOG_INFO("  - Time taken: %ld seconds", (long)(time(NULL) - t0));
    return 0;
}

/* read_ordering, :216-239: the file must open (exit(1) otherwise) */
static void check_ordering(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) {
        LOG_ERROR("Error opening file: %s", path);
        exit(1);
    }
    LOG_INFO("Reading ordering from %s", path);
    fclose(f);
}

typedef struct {
    bm_ctx *ctx;
    int32_t n, rows_per_sweep;
    const char *out_path;
    time_t start;
    int32_t *scratch, *reread;
} run_t;

/* save_intermediate_mapping, :478-492 */
static void save_intermediate(run_t *r, int64_t current_score) {
    bm_get_mapping(r->ctx, r->scratch);
    if (bm_mapping_write_mf(r->out_path, r->scratch, r->n)) {
        LOG_ERROR("Error creating file: %s", r->out_path);
        exit(1);
    }
    bm_pairs p;
    if (bm_pairs_read_optimized(r->out_path, &p)) return;
    bm_pairs_to_mapping(&p, r->n, r->reread);
    bm_pairs_free(&p);
    if (memcmp(r->reread, r->scratch, ((size_t)r->n + 1) * 4) != 0) {
        int64_t written = 0;
        bm_set_mapping(r->ctx, r->reread);
        bm_score(r->ctx, &written);
        bm_set_mapping(r->ctx, r->scratch);
        if (written != current_score)
            fprintf(stderr, "[ERROR] save_intermediate_mapping:487: Score mismatch - internal: %lld, written: %lld\n",
                    (long long)current_score, (long long)written);
    }
}

/* :570-576, logged at each accepted swap whose delta the rescore corrects
 * (lines leaving vertex 0, self-loops, repeated lines); the message names
 * the reference's function and line as its LOG_ERROR does */
static void mismatch_hook(void *user, int64_t calculated, int64_t verified) {
    (void)user;
    fprintf(stderr, "[ERROR] optimize_mapping:573: Score mismatch after swap - calculated: %lld, verified: %lld\n",
            (long long)calculated, (long long)verified);
}

static int row_hook(void *user, int64_t rows, int32_t row, int64_t score, int64_t improvements) {
    run_t *r = user;
    (void)row;
    if (rows % UPDATE_INTERVAL == 0) {
        double elapsed = difftime(time(NULL), r->start);
        double nodes_per_sec = rows / elapsed;
        LOG_INFO("Optimization progress:");
        LOG_INFO("  - Processed: %s/%s nodes (%.1f%%)", fmt_num(rows), fmt_num(r->rows_per_sweep),
                 (double)rows / r->rows_per_sweep * 100);
        LOG_INFO("  - Current score: %s", fmt_num(score));
        LOG_INFO("  - Improvements found: %s", fmt_num(improvements));
        LOG_INFO("  - Processing speed: %.1f nodes/sec", nodes_per_sec);
        LOG_INFO("  - Estimated time remaining: %.1f minutes", (r->rows_per_sweep - rows) / nodes_per_sec / 60);
        fflush(stdout);
    }
    if (rows % SAVE_INTERVAL == 0) save_intermediate(r, score);
    return 0;
}

int main(int argc, char *argv[]) {
    if (argc < 7) {
        LOG_ERROR("Usage: %s <male graph> <female graph> <male ordering> <female ordering> <in mapping> <out mapping>",
                  argv[0]);
        return 1;
    }
    time_t total_start = time(NULL);
    LOG_INFO("Graph Alignment Tool v1.0");
    LOG_INFO("Starting process with:");
    LOG_INFO("  - Male graph: %s", argv[1]);
    LOG_INFO("  - Female graph: %s", argv[2]);
    LOG_INFO("  - Output mapping: %s", argv[6]);

    bm_edges gm, gf;
    if (load_graph(argv[1], &gm)) {
        LOG_ERROR("Failed to load male graph");
        return 1;
    }
    if (load_graph(argv[2], &gf)) {
        LOG_ERROR("Failed to load female graph");
        bm_edges_free(&gm);
        return 1;
    }
    /* load_benchmark_mapping, :428-475 */
    const char *us = strrchr(argv[5], '_');
    if (us) LOG_INFO("Expected score from filename: %d", atoi(us + 1));
    bm_pairs pairs;
    if (bm_pairs_read_optimized(argv[5], &pairs)) {
        LOG_ERROR("Failed to open file: %s", argv[5]);
        LOG_ERROR("Failed to load benchmark mapping");
        bm_edges_free(&gm);
        bm_edges_free(&gf);
        return 1;
    }
    LOG_INFO("Loading benchmark mapping from %s", argv[5]);
    LOG_INFO("Loaded %s mappings", fmt_num(pairs.m));

    int32_t n = 1;
    if (gm.max_id > n) n = gm.max_id;
    if (gf.max_id > n) n = gf.max_id;
    if (pairs.max_id > n) n = pairs.max_id;
    /* BM_ROWS_PER_SWEEP: the reference's fixed vertex range (NUM_NODES, :15);
     * its rows above the data are empty vertices that a swap can still reach */
    const int64_t rows_env = env_i64("BM_ROWS_PER_SWEEP", 0);
    if (rows_env > n && rows_env <= INT32_MAX - 1) n = (int32_t)rows_env;
    int32_t *mapping = calloc((size_t)n + 1, 4);
    bm_pairs_to_mapping(&pairs, n, mapping);
    bm_pairs_free(&pairs);

    bm_ctx *ctx = cli_engine(n, &gm, &gf);
    bm_edges_free(&gm);
    bm_edges_free(&gf);
    if (!ctx) return 1;
    int rc = bm_set_mapping(ctx, mapping);
    if (rc) {
        LOG_ERROR("bm_set_mapping: %s (%s)", bm_strerror(rc), bm_last_error());
        return 1;
    }
    int64_t initial_score = 0;
    bm_score(ctx, &initial_score);
    LOG_INFO("Initial alignment score: %s", fmt_num(initial_score));

    /* optimize_mapping, :495-596 */
    LOG_INFO("Calculating node metrics");
    check_ordering(argv[3]);
    LOG_INFO("Calculating node metrics");
    check_ordering(argv[4]);
    run_t r;
    r.ctx = ctx;
    r.n = n;
    r.rows_per_sweep = (int32_t)env_i64("BM_ROWS_PER_SWEEP", n);
    if (r.rows_per_sweep < n) r.rows_per_sweep = n;
    r.out_path = argv[6];
    r.start = time(NULL);
    r.scratch = calloc((size_t)n + 1, 4);
    r.reread = calloc((size_t)n + 1, 4);
    LOG_INFO("Starting optimization with initial score: %s", fmt_num(initial_score));
    bm_drive_stats st;
    bm_drive_on_mismatch(mismatch_hook, NULL);
    rc = bm_drive_first_improvement(ctx, r.rows_per_sweep, env_i64("BM_MAX_SWEEPS", 0), (int)env_i64("BM_VERIFY", 1),
                                    row_hook, &r, NULL, 0, &st);
    if (rc) {
        LOG_ERROR("engine: %s (%s)", bm_strerror(rc), bm_last_error());
        return 1;
    }
    LOG_INFO("Optimization completed:");
    LOG_INFO("  - Final score: %s", fmt_num(st.final_score));
    LOG_INFO("  - Total improvements: %s", fmt_num(st.accepted));
    LOG_INFO("  - Time taken: %.1f minutes", difftime(time(NULL), r.start) / 60);
    save_intermediate(&r, st.final_score);

    int64_t optimized_score = 0;
    bm_score(ctx, &optimized_score);
    LOG_INFO("Process completed:");
    LOG_INFO("  - Initial score: %s", fmt_num(initial_score));
    LOG_INFO("  - Final score: %s", fmt_num(optimized_score));
    LOG_INFO("  - Improvement: %.2f%%", (double)(optimized_score - initial_score) / initial_score * 100.0);
    LOG_INFO("  - Total time: %.1f minutes", difftime(time(NULL), total_start) / 60);
    free(r.scratch);
    free(r.reread);
    free(mapping);
    bm_destroy(ctx);
    return 0;
}
