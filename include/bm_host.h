/*
 * bm_host.h -- host-side C layer of the MI355X move-delta engine: the
 * reference CLIs' file formats, seeded synthetic workloads and the local
 * search drivers that call the device engine (bmatch.h).
 *
 * Formats follow the reference programs byte for byte (zhihuidu/brain-match,
 * src/...); each function names the routine it mirrors.  Unlike the
 * reference, n is derived from the data instead of the fixed NUM_NODES =
 * 18524 (graph-alignment-optimized.c:15).
 */
#ifndef BM_HOST_H
#define BM_HOST_H

#include <stdint.h>
#include <stdio.h>

#include "bmatch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- edges */

typedef struct {
    int32_t *src, *dst, *w;
    int64_t m, cap;
    int32_t max_id; /* largest vertex id seen (-1 if none) */
    int64_t malformed;
} bm_edges;

/* How a loader reports a line that is not "%d,%d,%d":
 *   BM_CSV_LOG_ERROR : "[ERROR] load_graph_from_csv:666: Malformed line in
 *                      CSV: <line>" on stderr (graph-alignment-optimized.c:667)
 *   BM_CSV_WARN_LINE : "Warning: Invalid format at line <k>" on stderr
 *                      (refine-matching-hillclimb.c:423)
 *   BM_CSV_SILENT    : skipped silently (graph-alignment-permset-parallel.c:382)
 * The first line is always skipped as the header. */
#define BM_CSV_SILENT 0
#define BM_CSV_LOG_ERROR 1
#define BM_CSV_WARN_LINE 2
int bm_edges_read_csv(const char *path, int report, bm_edges *out);
int bm_edges_write_csv(const char *path, const char *header, const int32_t *src, const int32_t *dst,
                       const int32_t *w, int64_t m);
void bm_edges_free(bm_edges *e);

/* ---------------------------------------------------------------- mappings */

/* A parsed mapping file: pairs in file order. */
typedef struct {
    int32_t *male, *female;
    int64_t m;
    int32_t max_id; /* largest id on either side */
} bm_pairs;

/* load_benchmark_mapping, graph-alignment-optimized.c:428-475: "%d,%d",
 * else "%[^,],%s" with an optional leading 'm' / 'f' stripped by atoi. */
int bm_pairs_read_optimized(const char *path, bm_pairs *out);
/* load_mapping, graph-alignment-permset-parallel.c:441-468: "%[^,],%s",
 * the first character of each field always dropped (atoi(str + 1)). */
int bm_pairs_read_permset(const char *path, bm_pairs *out);
/* load_initial_matching, refine-matching-hillclimb.c:455-492: "%d,%d". */
int bm_pairs_read_plain(const char *path, bm_pairs *out);
void bm_pairs_free(bm_pairs *p);
/* mapping[male] = female (later lines win), size n+1, zero-filled;
 * pairs with an id outside [0,n] are ignored. */
void bm_pairs_to_mapping(const bm_pairs *p, int32_t n, int32_t *mapping);

/* write_mapping, graph-alignment-optimized.c:410-425 (and save_mapping,
 * graph-alignment-permset-parallel.c:391-407): header
 * "Male Node ID,Female Node ID", then "m%d,f%d" for i = 1..n with
 * mapping[i] != 0. */
int bm_mapping_write_mf(const char *path, const int32_t *mapping, int32_t n);

/* ---------------------------------------------------------------- workloads */

/* A seeded synthetic instance: G1, G2 (= hidden permutation pi of G1 plus
 * noise) and an initial mapping p0, all with ids 1..n. */
typedef struct {
    int32_t n;
    int64_t e1, e2;
    int32_t *s1, *d1, *w1, *s2, *d2, *w2;
    int32_t *pi; /* hidden alignment, n+1 entries (pi[0] = 0) */
    int32_t *p0; /* initial mapping, n+1 entries */
} bm_instance;

/* Power-law out-degrees (exponent 2.1, min 1, cap n-1) rescaled to sum to
 * m, targets uniform over V\{u} without repeats, weights geometric(0.3)+1
 * capped at 255; G2 = pi(G1) with del_frac of its lines deleted and the same
 * number of fresh uniform lines added; p0 = pi with shuffle_frac of the
 * vertices permuted among themselves (1.0 = a uniform random bijection). */
int bm_gen_powerlaw(uint64_t seed, int32_t n, int64_t m, double del_frac, double shuffle_frac, bm_instance *out);
/* Chung-Lu graph with power-law (2.1) out- and in-weights scaled to m lines,
 * no self-loops or repeated pairs, weights round(LogNormal(1.0, 1.2))
 * clipped to [1, 2000]; G2 = pi(G1) with rewire_frac of the lines rewired and
 * reweight_frac of the rest re-drawn; p0 = pi with shuffle_frac shuffled. */
int bm_gen_chunglu(uint64_t seed, int32_t n, int64_t m, double rewire_frac, double reweight_frac,
                   double shuffle_frac, bm_instance *out);
void bm_instance_free(bm_instance *inst);

/* count triples (u, v, w): u uniform in 1..n, v and w distinct members of
 * u's 2-hop neighbourhood over G1 out- and in-edges, each the end of a
 * random walk of one or two steps (equal odds) from u; u is redrawn when 64
 * walks do not yield two distinct vertices other than u.  out holds
 * 3*count ids. */
int bm_gen_triples_2hop(uint64_t seed, const bm_instance *inst, int64_t count, int32_t *out);
/* count sets of k distinct vertices in 1..n; out holds k*count ids. */
int bm_gen_sets(uint64_t seed, int32_t n, int32_t k, int64_t count, int32_t *out);

/* ---------------------------------------------------------------- drivers */

/* Per-row hook of the sweep drivers, called at the top of every row before
 * its moves are searched (like the progress/dump blocks of
 * graph-alignment-optimized.c:519-540).  rows = rows processed including
 * this one, score = current score, improvements = accepted moves so far.
 * A non-zero return stops the driver.  The hook may read the context (the
 * mapping, the score) but must not change it: the drivers evaluate up to
 * BM_ROW_GROUP_MAX rows ahead and use those results for as long as they apply
 * no swap themselves. */
typedef int (*bm_row_hook)(void *user, int64_t rows, int32_t row, int64_t score, int64_t improvements);

/* Per-swap mismatch hook of the drivers: called when a verified swap's exact
 * score change disagrees with its accepted delta, with the score the delta
 * gave (calculated) and the verified score that replaces it, at the point
 * where graph-alignment-optimized.c:570-576 logs "Score mismatch after swap -
 * calculated: %d, verified: %d" (also alignment-omp.c:652-656).  Registered
 * per calling thread for the drivers it calls afterwards; NULL turns it off.
 * No counterpart function in the reference. */
typedef void (*bm_mismatch_hook)(void *user, int64_t calculated, int64_t verified);
void bm_drive_on_mismatch(bm_mismatch_hook hook, void *user);

typedef struct {
    int64_t sweeps, rows, accepted, evals_rows;
    int64_t initial_score, final_score;
    int64_t mismatches; /* accepted deltas that disagreed with the rescore */
    /* rows not searched one by one because an all-pairs sweep of the current
     * mapping found no improving swap at all (counted in rows, not in evals_rows) */
    int64_t rows_cleared;
} bm_drive_stats;

/* First-improvement sweep of graph-alignment-optimized.c:513-581 on the
 * device: rows a = 1..rows_per_sweep (>= n; rows above n only count), for
 * each the smallest b (from the cursor) with Delta_B > 0 is applied at once
 * and the search resumes at b+1.  After each accept the score is verified
 * by a device rescore when verify != 0 (:570-576).  Stops after a sweep
 * without an accept, or after max_sweeps (> 0).  Accepted moves are written
 * as (a, b, delta) to acc (3 per move) up to acc_cap. */
int bm_drive_first_improvement(bm_ctx *ctx, int32_t rows_per_sweep, int64_t max_sweeps, int verify,
                               bm_row_hook hook, void *user, int32_t *acc, int64_t acc_cap,
                               bm_drive_stats *stats);

/* Row-best pass(es) of alignment-omp.c:598-658 at one thread: per row the
 * arg-max over b > a (lowest b on ties), applied when > 0. */
int bm_drive_row_best(bm_ctx *ctx, int64_t max_sweeps, int verify, bm_row_hook hook, void *user, int32_t *acc,
                      int64_t acc_cap, bm_drive_stats *stats);

/* Global best-swap steps of graph-re-alignment-gpu.cu:830-891 (all-pairs
 * arg-max, apply, repeat) until no positive pair or max_steps.  kind
 * BM_DELTA_DENSE ranks by that program's own kernel delta
 * (bm_sweep_all_pairs_dense) and, as it does (:874), adds the delta to the
 * score; BM_DELTA_LISTS ranks by Delta_B.  With verify the score follows the
 * exact change of every swap instead, and a delta that differs from it is
 * counted in stats->mismatches. */
int bm_drive_global_best(bm_ctx *ctx, int32_t kind, int64_t max_steps, int verify, int32_t *acc, int64_t acc_cap,
                         bm_drive_stats *stats);

/* Row-best passes with the stagnation restart of alignment-omp.c:543-678:
 * the start mapping is perturbed by initial_swaps swaps (:564); after each
 * pass, when more than idle_passes passes ran since the last restart and
 * the pass accepted nothing (:660), the best mapping so far is restored and
 * perturbed by restart_swaps swaps (:669-670).  Perturbations are
 * bm_perturb(BM_PERTURB_OMP) from a splitmix64 state seeded with seed.  The
 * score is re-verified after each accept when verify != 0 (:652-656).
 * Stops after max_passes passes (> 0) and leaves the context on the best
 * mapping seen (the one the reference writes out); best_mapping (n+1
 * entries, may be NULL) receives it. */
typedef struct {
    uint64_t seed;
    int32_t initial_swaps; /* 100 in the reference */
    int32_t restart_swaps; /* 50 */
    int32_t idle_passes;   /* 2: restart when pass > 2 */
    int32_t verify;
    int64_t max_passes;
} bm_restart_params;

typedef struct {
    int64_t passes, restarts, accepted, evals_rows;
    int64_t initial_score; /* of the start mapping, before the first perturbation */
    int64_t final_score;   /* of the mapping the run ends on, before the best is restored */
    int64_t best_score;
    int64_t mismatches;
} bm_restart_stats;

int bm_drive_row_best_restart(bm_ctx *ctx, const bm_restart_params *params, int32_t *best_mapping,
                              bm_restart_stats *stats);

/* The changed-vertex search of quick.cu:1137-1324 (one process, no
 * best-file synchronisation): perturb by 3 or 4 swaps
 * (bm_perturb(BM_PERTURB_WEIGHTED)), starting from the best mapping when the
 * last improvement is fewer than best_window evaluations old or the current
 * score trails the best by more than gap x gap_norm (:1141); evaluate every
 * pair touching a changed vertex (bm_eval_rows); then, while a pair has
 * Delta > 0 or = 0 and fewer than zero_limit zero moves ran in a row: with
 * Delta > 0 available take a random positive pair or the best pair (even
 * odds), else a random zero pair; swap, rescore, keep the best, evaluate the
 * two changed rows again.  Pairs are enumerated as the reference's kernel
 * numbers them (pairs inside the changed set, then changed x unchanged
 * ascending); the best is the first strict maximum.  Random choices come
 * from a splitmix64 state seeded with seed, in the reference's order of
 * rand() calls.  delta_kind ranks the pairs: BM_DELTA_DENSE with the
 * reference kernel's own dense-matrix delta (bm_eval_swaps_dense), or
 * BM_DELTA_LISTS with Delta_B; the two agree unless G1 has self-loops,
 * repeated lines, lines at vertex 0 or weights outside [1, 32767].  The
 * score is S_opt either way, kept by the exact change of every swap.  Stops
 * after max_evals evaluations and leaves the context on the best mapping. */
typedef struct {
    uint64_t seed;
    int64_t max_evals;
    int32_t zero_limit;  /* 1000 */
    int32_t best_window; /* 500 */
    double gap;          /* 0.002 */
    double gap_norm;     /* 5154247, the reference's benchmark score */
    int32_t delta_kind;  /* BM_DELTA_DENSE as quick.cu, or BM_DELTA_LISTS */
} bm_quick_params;

typedef struct {
    int64_t evals, perturbations, moves, positive_moves, zero_moves;
    int64_t initial_score, final_score, best_score;
    int64_t pairs; /* pairs evaluated over all evaluations */
} bm_quick_stats;

int bm_drive_quick(bm_ctx *ctx, const bm_quick_params *params, int32_t *best_mapping, bm_quick_stats *stats);

/* The triples q3.cu's calculateThreeNodeSwapKernel enumerates for the changed
 * vertices chg[0..k) (ascending, distinct, in [1, n]) of an n-vertex graph,
 * in its thread-index order (q3.cu:59-150): the triples inside the changed
 * set (lexicographic), then for each unchanged vertex (ascending) every
 * changed pair, then for each changed vertex every pair of consecutive
 * unchanged vertices; the unchanged vertices are 1..n without chg, ascending
 * (fillArray, :982-996).  Writes 3 ids per triple to out when non-NULL and
 * returns the count (negative on bad arguments). */
int64_t bm_q3_triples(int32_t n, const int32_t *chg, int32_t k, int32_t *out);

/* The changed-vertex triple search of q3.cu:1176-1440 (one process, no
 * best-file synchronisation): perturb by 3 + rand()%2 swaps
 * (bm_perturb(BM_PERTURB_WEIGHTED)), from the best mapping while fewer than
 * best_window evaluations ran (the reference never advances its `last`
 * marker) or while the score trails the best by more than gap x gap_norm;
 * score every triple of bm_q3_triples for the sorted changed vertices with
 * Delta_G3 over p = 1..5 (bm_eval_rot3, BM_ROT3_ALL); then, while a triple
 * has Delta > 0 or = 0 and fewer than zero_limit zero moves ran in a row:
 * with Delta > 0 a random positive triple or the first best one (even
 * odds), else a random zero triple; apply it (bm_apply_rot3); the score is
 * rescored after positive moves only, as the reference does (:1293);
 * evaluate the triples of the three moved vertices again.  Random choices
 * come from a splitmix64 state seeded with seed, in the reference's order
 * of rand() calls.  Stops after max_evals evaluations and leaves the
 * context on the best mapping. */
typedef struct {
    uint64_t seed;
    int64_t max_evals;
    int32_t zero_limit;  /* 1000 */
    int32_t best_window; /* 1500 */
    double gap;          /* 0.002 */
    double gap_norm;     /* 5154247 */
} bm_q3_params;

typedef struct {
    int64_t evals, perturbations, moves, positive_moves, zero_moves;
    int64_t initial_score, final_score, best_score;
    int64_t triples; /* triples evaluated over all evaluations */
} bm_q3_stats;

int bm_drive_q3(bm_ctx *ctx, const bm_q3_params *params, int32_t *best_mapping, bm_q3_stats *stats);

/* The triples m1 < m2 < m3 of 1..n in lexicographic order, from index start
 * (0-based): writes count triples (3 ids each) to out and returns how many
 * there were (fewer at the end of the range; negative on bad arguments).
 * The order triple.cu's calculateThreeNodeSwapKernel intends for its batch
 * index (triple.cu:54-75). */
int64_t bm_lex_triples(int32_t n, int64_t start, int64_t count, int32_t *out);

/* The batched triple pass of triple.cu:941-1196 (one process, no best-file
 * synchronisation): the lexicographic triples of 1..n in batches of `batch`
 * from the first; each batch scored with Delta_G3 over p = 1..5
 * (bm_eval_rot3, BM_ROT3_ALL); the first triple of the batch with the largest
 * positive delta is applied (bm_apply_rot3), the running score raised by its
 * delta, and the pass restarts from the first batch (:1148-1168); a batch
 * without one moves on to the next.  After a pass in which no batch had a
 * positive triple, the best mapping so far is perturbed by 3 + rand()%3
 * swaps of random_swap_k_vertices (:822-858, bm_perturb with BM_PERTURB_OMP)
 * and rescored (:1174-1181), and the next pass starts.  Random draws come
 * from a splitmix64 state seeded with seed.  Stops after max_evals batch
 * evaluations and leaves the context on the best mapping.  Deviation: the
 * reference advances by batch + total/3 after a batch without a move, which
 * takes its index map (whose blocks divide by 6 where the count of pairs
 * needs 2) beyond m1 = n from the second batch on; here the batches follow
 * one another through the whole range. */
typedef struct {
    uint64_t seed;
    int64_t max_evals; /* batch evaluations */
    int32_t batch;     /* 270000 in triple.cu */
} bm_triples_params;

typedef struct {
    int64_t evals, passes, perturbations, moves;
    int64_t initial_score, final_score, best_score;
    int64_t triples; /* triples evaluated over all batches */
} bm_triples_stats;

int bm_drive_triples(bm_ctx *ctx, const bm_triples_params *params, int32_t *best_mapping, bm_triples_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* BM_HOST_H */
