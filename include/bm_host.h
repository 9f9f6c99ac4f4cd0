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
 * A non-zero return stops the driver. */
typedef int (*bm_row_hook)(void *user, int64_t rows, int32_t row, int64_t score, int64_t improvements);

typedef struct {
    int64_t sweeps, rows, accepted, evals_rows;
    int64_t initial_score, final_score;
    int64_t mismatches; /* accepted deltas that disagreed with the rescore */
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
 * arg-max, apply, repeat) until no positive pair or max_steps. */
int bm_drive_global_best(bm_ctx *ctx, int64_t max_steps, int verify, int32_t *acc, int64_t acc_cap,
                         bm_drive_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* BM_HOST_H */
