/*
 * graph-alignment-permset-parallel -- drop-in for
 * src/graph-alignment-permset-parallel.c (zhihuidu/brain-match) with the
 * permutation-delta loop on the MI355X engine.
 *
 *   graph-alignment-permset-parallel <male_graph> <female_graph> <initial_mapping>
 *
 * Same arguments and formats as the reference (:761-800): the initial
 * mapping's ids lose their first character (load_mapping, :457-460), every
 * accepted improvement is written to
 * improved_mapping_score_<S>_<YYYYmmdd_HHMMSS>.csv in the working directory
 * (:701-719).  Per iteration (:625-744): NUM_GROUPS random k-sets, every
 * permutation of each in next_permutation order, the strict-> best per set
 * and across sets, apply, full device rescore, keep if the score rose, else
 * apply the same permutation once more -- the reference's "revert"
 * (:722-728), kept as written.  The GPU evaluates all groups of an
 * iteration in one launch.  Differences: groups are drawn with the
 * project PRNG (seed BM_SEED, default time) instead of rand(); n comes from
 * the data.  Environment: BM_GROUP_SIZE (4, as GROUP_SIZE :11; 2..8),
 * BM_NUM_GROUPS (4096), BM_MAX_ITERS (0 = run until killed, as the
 * reference), BM_SEED, BM_DEVICE.
 */
#include "cli_common.h"

static int load_graph(const char *path, bm_edges *g) {
    printf("Loading graph from %s...\n", path);
    if (bm_edges_read_csv(path, BM_CSV_SILENT, g)) {
        fprintf(stderr, "Error opening file: %s\n", path);
        return -1;
    }
    printf("\nLoaded total of %lld edges\n", (long long)g->m);
    return 0;
}

/* save_mapping, :391-407 */
static void save_mapping(const char *path, const int32_t *mapping, int32_t n, int64_t score) {
    if (bm_mapping_write_mf(path, mapping, n)) {
        fprintf(stderr, "Error creating file: %s\n", path);
        return;
    }
    printf("Saved mapping with score %lld to %s\n", (long long)score, path);
}

int main(int argc, char *argv[]) {
    if (argc != 4) {
        fprintf(stderr, "Usage: %s <male_graph> <female_graph> <initial_mapping>\n", argv[0]);
        return 1;
    }
    const int k = (int)env_i64("BM_GROUP_SIZE", 4);
    const int64_t num_groups = env_i64("BM_NUM_GROUPS", 4096);
    const int64_t max_iters = env_i64("BM_MAX_ITERS", 0);
    uint64_t seed = (uint64_t)env_i64("BM_SEED", (int64_t)time(NULL));
    if (k < 2 || k > 8 || num_groups < 1) {
        fprintf(stderr, "BM_GROUP_SIZE must be in 2..8 and BM_NUM_GROUPS >= 1\n");
        return 1;
    }
    printf("Starting graph alignment optimization\n");
    printf("Using GPU device %lld\n\n", (long long)env_i64("BM_DEVICE", 0));
    printf("Loading input files:\n");
    bm_edges gm, gf;
    bm_pairs pairs;
    int ok = load_graph(argv[1], &gm) == 0;
    ok = ok && load_graph(argv[2], &gf) == 0;
    ok = ok && bm_pairs_read_permset(argv[3], &pairs) == 0;
    if (!ok) {
        fprintf(stderr, "Error loading input files\n");
        return 1;
    }
    printf("Loaded %lld mappings\n", (long long)pairs.m);
    int32_t n = k;
    if (gm.max_id > n) n = gm.max_id;
    if (gf.max_id > n) n = gf.max_id;
    if (pairs.max_id > n) n = pairs.max_id;
    int32_t *mapping = calloc((size_t)n + 1, 4);
    bm_pairs_to_mapping(&pairs, n, mapping);
    bm_pairs_free(&pairs);
    char output_path[4096];
    snprintf(output_path, sizeof output_path, "improved_%s", argv[3]);

    bm_ctx *ctx = cli_engine(n, &gm, &gf);
    bm_edges_free(&gm);
    bm_edges_free(&gf);
    if (!ctx || bm_set_mapping(ctx, mapping)) {
        fprintf(stderr, "Error loading input files\n");
        return 1;
    }
    printf("\nStarting optimization...\n");
    int64_t current_score = 0;
    bm_score(ctx, &current_score);
    int total = 1;
    for (int i = 2; i <= k; i++) total *= i;
    printf("Using complete enumeration: testing all %d permutations\n", total);
    printf("\nInitializing optimization with:\n");
    printf("  - Number of groups: %lld\n", (long long)num_groups);
    printf("  - Group size: %d\n", k);
    printf("  - Random permutation samples: %d\n", total);
    printf("  - Initial score: %lld\n", (long long)current_score);
    printf("  - Output path: %s\n\n", output_path);
    fflush(stdout);

    int32_t *sets = malloc((size_t)num_groups * k * 4);
    int32_t *bd = malloc((size_t)num_groups * 4), *br = malloc((size_t)num_groups * 4);
    int32_t perm[8];
    int64_t iters = 0, improvements = 0;
    double t0 = wall_s();
    while (max_iters <= 0 || iters < max_iters) {
        bm_gen_sets(seed + (uint64_t)iters * 0x9E3779B97F4A7C15ull, n, k, num_groups, sets);
        int rc = bm_eval_permset(ctx, k, sets, num_groups, bd, br);
        if (rc) {
            fprintf(stderr, "engine: %s (%s)\n", bm_strerror(rc), bm_last_error());
            return 1;
        }
        int64_t best_group = -1;
        int32_t best_delta = 0;
        for (int64_t g = 0; g < num_groups; g++)
            if (bd[g] > best_delta) {
                best_delta = bd[g];
                best_group = g;
            }
        if (best_delta > 0) {
            const int32_t *vs = sets + (size_t)best_group * k;
            bm_perm_from_rank(k, br[best_group], perm);
            bm_apply_perm(ctx, k, vs, perm);
            int64_t new_score = 0;
            bm_score(ctx, &new_score);
            if (new_score > current_score) {
                improvements++;
                printf("\nApplied permutation from group %lld:\n", (long long)best_group);
                printf("  - Delta: %d\n", best_delta);
                printf("  - New score: %lld\n", (long long)new_score);
                char fname[512];
                time_t now = time(NULL);
                struct tm *t = localtime(&now);
                snprintf(fname, sizeof fname, "improved_mapping_score_%lld_%04d%02d%02d_%02d%02d%02d.csv",
                         (long long)new_score, t->tm_year + 1900, t->tm_mon + 1, t->tm_mday, t->tm_hour, t->tm_min,
                         t->tm_sec);
                bm_get_mapping(ctx, mapping);
                save_mapping(fname, mapping, n, new_score);
                current_score = new_score;
            } else {
                bm_apply_perm(ctx, k, vs, perm); /* :722-728 applies the permutation again */
            }
        }
        iters++;
        if (iters % 100 == 0) {
            double el = wall_s() - t0;
            printf("\n=== Progress: iteration %lld, score %lld, improvements %lld, %.1f iterations/s ===\n",
                   (long long)iters, (long long)current_score, (long long)improvements, iters / el);
            fflush(stdout);
        }
    }
    printf("\nCleaning up...\n");
    free(sets);
    free(bd);
    free(br);
    free(mapping);
    bm_destroy(ctx);
    return 0;
}
