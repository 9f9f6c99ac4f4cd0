/*
 * bm_oracle.h -- CPU restatement of brain-match's move-delta and score
 * arithmetic.  TEST INFRASTRUCTURE ONLY: this code is the checker that the
 * HIP path is compared against.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load it; the product library never
 * links or calls it.
 *
 * Every function restates one reference routine, cited by file:line into
 * zhihuidu/brain-match (src/...).  The data layout deliberately follows the
 * reference: per-vertex out/in lists in file order (the reference's EdgeMap,
 * graph-alignment-optimized.c:41-53, 183-204) and dense `short` adjacency
 * matrices of (n+1)^2 cells where the last CSV line for a pair wins
 * (graph-alignment-optimized.c:185).
 *
 * Vertex ids are 0..n; index 0 is the "unmapped" sentinel of the reference
 * (mapping[] is calloc'ed, W2 row/column 0 stay zero unless an input names
 * vertex 0).
 */
#ifndef BM_ORACLE_H
#define BM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_graphs orc_graphs;

/* Build from edge lists in file order.  n = largest vertex id. */
orc_graphs *orc_build(int n,
                      const int32_t *s1, const int32_t *d1, const int32_t *w1, int64_t e1,
                      const int32_t *s2, const int32_t *d2, const int32_t *w2, int64_t e2);
void orc_free(orc_graphs *g);
int orc_n(const orc_graphs *g);

/* S_opt: graph-alignment-optimized.c:292-307 (duplicates in G1 counted,
 * W2 through the dense short matrix). */
int64_t orc_score_opt(const orc_graphs *g, const int32_t *map);

/* S_ref: score-matching.c:558-598 with the loaders :339-358 (G1 and G2
 * de-duplicated, last weight wins, full int weights, absent G2 edge -> 0).
 * map[x] < 0 means "not in the matching file". */
int64_t orc_score_ref(const orc_graphs *g, const int32_t *map);

/* Delta_B: graph-alignment-optimized.c:321-408. */
int32_t orc_delta_b(const orc_graphs *g, const int32_t *map, int a, int b);

/* Delta_G2: graph-re-alignment-gpu.cu:60-128 / quick.cu:81-130 (dense scan
 * over i not in {a,b}; self-loops omitted). */
int32_t orc_delta_g2(const orc_graphs *g, const int32_t *map, int a, int b);

/* Delta_G3 for permutation p in 1..5 of triple (m1,m2,m3):
 * triple.cu:85-208.  p=3 -> (f2,f3,f1), p=4 -> (f3,f1,f2) are the rotations. */
int32_t orc_delta_g3(const orc_graphs *g, const int32_t *map, int m1, int m2, int m3, int p);
/* Best over p (strict >, from 0; triple.cu:94-214).  mode 0: p in {3,4};
 * mode 1: p in 1..5.  Returns best delta, *best_p = 0 when none positive. */
int32_t orc_best_g3(const orc_graphs *g, const int32_t *map, int m1, int m2, int m3, int mode, int *best_p);
/* apply_permutation, triple.cu:903-934 */
void orc_apply_g3(int32_t *map, int m1, int m2, int m3, int p);

/* Delta_P: graph-alignment-permset-parallel.c:304-339 (intra-set edges are
 * counted twice, exactly as written). */
int32_t orc_delta_p(const orc_graphs *g, const int32_t *map, const int32_t *set, const int32_t *perm, int k);
/* next_permutation, graph-alignment-permset-parallel.c:342-363 */
int orc_next_permutation(int32_t *arr, int n);
/* rank -> permutation in next_permutation order from the identity
 * (generate_permutation_samples :563-571). */
void orc_unrank_perm(int k, int64_t rank, int32_t *perm);
/* Best permutation of one set: strict > from 0 in lexicographic order
 * (:665-674).  Returns delta; *rank = -1 when none positive. */
int32_t orc_best_perm(const orc_graphs *g, const int32_t *map, const int32_t *set, int k, int64_t *rank);
/* apply_best_permutation :508-518 */
void orc_apply_perm(int32_t *map, const int32_t *set, const int32_t *perm, int k);

/* Delta_A: refine-matching-hillclimb.c:244-296 over the hash-table
 * adjacency (:101-137: de-duplicated, last weight wins; absent edge -> 0).
 * map[x] <= 0 is unmapped (contributes 0, as calculate_total_score skips it). */
int32_t orc_delta_a(const orc_graphs *g, const int32_t *map, int a, int b);

/* Exact change of S_opt for swapping a,b (full rescore, brute force). */
int64_t orc_delta_swap_exact(const orc_graphs *g, const int32_t *map, int a, int b);

/* --- drivers ------------------------------------------------------------ */

/* First-improvement sweep of graph-alignment-optimized.c:513-581.  Rows
 * m1 = 1..rows_per_sweep, m2 = 1..rows_per_sweep (m2 != m1); a swap is
 * applied as soon as Delta_B > 0 and the current score is replaced by a full
 * rescore when they disagree (:570-576).  rows_per_sweep >= n emulates the
 * reference's fixed NUM_NODES padding (rows above n never accept when all
 * weights are >= 0).  Stops after a sweep without an accepted swap or after
 * max_sweeps.  Records accepted moves (a,b,delta) into acc[3*i] (up to
 * acc_cap) and a progress sample (rows processed, score, improvements) every
 * log_every rows into logv[3*j] (up to log_cap), like :519-534.
 * Returns the number of sweeps run; *n_acc / *n_log get the counts. */
int orc_first_improvement(const orc_graphs *g, int32_t *map, int rows_per_sweep, int max_sweeps,
                          int32_t *acc, int64_t acc_cap, int64_t *n_acc,
                          int64_t *logv, int64_t log_cap, int64_t *n_log, int log_every,
                          int64_t *final_score);

/* Row-best pass of alignment-omp.c:598-658 at one thread: for m1 = 1..n,
 * arg-max over m2 > m1 (strict >, init -1, lowest m2 wins), apply if > 0.
 * Returns number of applied swaps; acc as above. */
int64_t orc_row_best_pass(const orc_graphs *g, int32_t *map, int32_t *acc, int64_t acc_cap);

/* All-pairs arg-max of graph-re-alignment-gpu.cu:859-868 over pairs
 * a < b in lexicographic order (strict >, from 0).  Uses Delta_B.  Fills
 * row_best_b/row_best_d (index a, for b > a; strict > from INT_MIN so the
 * lowest b wins ties) when non-NULL.  Returns best delta (0 if none > 0). */
int32_t orc_all_pairs(const orc_graphs *g, const int32_t *map, int32_t *best_a, int32_t *best_b,
                      int32_t *row_best_b, int32_t *row_best_d, int32_t *delta_matrix);

/* --- the reference OpenMP row-best loop, reference data layout ----------- */
/* Times rows [row_lo, row_hi) of alignment-omp.c:598-622 (Delta_B of
 * :321-408 over EdgeMap lists and `short` rows, `#pragma omp parallel for
 * schedule(dynamic)` over m2 > m1, critical arg-max) with the mapping held
 * fixed (no apply), and returns the number of Delta evaluations done.
 * threads <= 0 uses the OpenMP default. */
int64_t orc_omp_rowbest_rows(const orc_graphs *g, const int32_t *map, int row_lo, int row_hi, int threads,
                             int64_t *checksum);

/* --- the same arithmetic over many rows / moves (OpenMP, for the checks at
 * the full-size configurations; threads <= 0 uses the OpenMP default) ----- */

/* For each rows[i]: the arg-max over b > rows[i] of Delta_B (strict > from
 * INT32_MIN, lowest b on ties), as orc_all_pairs' row_best_b/row_best_d. */
void orc_row_best_rows(const orc_graphs *g, const int32_t *map, const int32_t *rows, int64_t count,
                       int32_t *best_b, int32_t *best_d, int threads);
/* Whole rows: out[i*(n+1) + b] = Delta_B(rows[i], b) for b in [1, n], b != rows[i];
 * entries b = 0 and b = rows[i] are 0 (bm_eval_rows' layout). */
void orc_delta_rows(const orc_graphs *g, const int32_t *map, const int32_t *rows, int64_t count, int32_t *out,
                    int threads);
/* The first max_rows rows of the first-improvement sweep of
 * graph-alignment-optimized.c:513-581 (m2 = 1..n from the row's cursor,
 * apply at the first Delta_B > 0, resume at m2+1), without the per-accept
 * rescore: on graphs without self-loops, repeated lines or lines at vertex 0
 * the delta is the exact change, so the rescore changes nothing.  Each
 * cursor's candidates are evaluated in parallel blocks and scanned in order.
 * Returns the number of accepted swaps, recorded as (a, b, delta) in acc. */
int64_t orc_first_improvement_rows(const orc_graphs *g, int32_t *map, int max_rows, int32_t *acc, int64_t acc_cap,
                                   int threads);
/* orc_best_g3 / orc_best_perm over lists of triples / k-sets. */
void orc_best_g3_many(const orc_graphs *g, const int32_t *map, const int32_t *triples, int64_t count, int mode,
                      int32_t *best_d, int32_t *best_p, int threads);
void orc_best_perm_many(const orc_graphs *g, const int32_t *map, const int32_t *sets, int k, int64_t count,
                        int32_t *best_d, int64_t *best_rank, int threads);

/* --- perturbation and the searches built on it --------------------------- */
/* splitmix64 step (the generator the engine's perturbation draws from). */
uint64_t orc_rng_next(uint64_t *state);

/* quick.cu:766-793 vertex weights into scores[0..n] (scores[0] = 0). */
void orc_vertex_scores(const orc_graphs *g, float *scores);

/* random_swap_k_vertices: mode 0 = alignment-omp.c:495-541 (2k distinct
 * indices rand() % n), mode 1 = quick.cu:816-887 (k weighted draws, then k
 * uniform rand() % n + 1); swaps map[sel[i]] <-> map[sel[i+k]].  rand() is
 * replaced by orc_rng_next(rng).  sel gets the 2k draws.  Returns 0, or -1
 * when 2k > n or k < 1. */
int orc_perturb(const orc_graphs *g, int32_t *map, int k, int mode, uint64_t *rng, int32_t *sel);

/* alignment-omp.c:543-678 row-best passes with restarts, max_passes passes;
 * map ends on the best mapping; out[0..5] = passes, restarts, accepted,
 * initial score, final (current) score, best score. */
int orc_row_best_restart(const orc_graphs *g, int32_t *map, uint64_t seed, int initial_swaps, int restart_swaps,
                         int idle_passes, int64_t max_passes, int64_t *out);

/* quick.cu:1137-1324 changed-vertex search for max_evals evaluations; map
 * ends on the best mapping; pairs are ranked by Delta_G2 (kind 1, the
 * reference kernel's delta) or Delta_B (kind 0); out[0..8] = evals,
 * perturbations, moves, positive moves, zero moves, initial, final (current),
 * best score, pairs evaluated. */
int orc_quick(const orc_graphs *g, int32_t *map, uint64_t seed, int64_t max_evals, int zero_limit,
              int best_window, double gap, double gap_norm, int kind, int64_t *out);

/* q3.cu:59-150: the triples calculateThreeNodeSwapKernel enumerates for the
 * changed vertices chg[0..k) (sorted ascending) of an n-vertex graph, in
 * its index order: (1) triples inside the changed set, lexicographic;
 * (2) for each unchanged vertex r (ascending), the changed pairs i < j
 * lexicographic with r; (3) for each changed vertex c, the consecutive
 * unchanged pairs (r_t, r_t+1).  The unchanged vertices are 1..n without
 * chg, ascending (fillArray, :982-996).  Writes 3 ids per triple to out
 * when non-NULL; returns the count. */
int64_t orc_q3_triples(int n, const int32_t *chg, int k, int32_t *out);
/* q3.cu:1176-1440 changed-vertex triple search for max_evals evaluations
 * (one process, no best-file sync): perturbation by 3 + rand()%2 weighted
 * swaps from the best mapping while iter < best_window (the reference never
 * moves `last` from 0) or while the score trails the best by more than
 * gap x gap_norm, else from the current one; every enumerated triple scored
 * with Delta_G3 over p = 1..5; then, while a triple has Delta > 0 or = 0 and
 * fewer than zero_limit zero moves ran in a row: with Delta > 0 a random
 * positive triple or the first best one (even odds), else a random zero
 * triple; apply_permutation; the score is rescored after positive moves only
 * (:1293), as there; the moved triple is evaluated again.  rand() is
 * orc_rng_next.  map ends on the best mapping; out[0..8] = evals,
 * perturbations, moves, positive moves, zero moves, initial, final
 * (current), best score, triples evaluated. */
int orc_q3(const orc_graphs *g, int32_t *map, uint64_t seed, int64_t max_evals, int zero_limit, int best_window,
           double gap, double gap_norm, int64_t *out);
/* The lexicographic triple of rank t (0-based) among the triples
 * m1 < m2 < m3 of 1..n, restated from its definition: the ranks count the
 * triples of smaller first element, then of smaller second element (the
 * order triple.cu:54-75 intends).  out gets count triples from rank start;
 * returns how many (fewer at the end of the range). */
int64_t orc_lex_triples(int n, int64_t start, int64_t count, int32_t *out);
/* triple.cu:941-1196 batched triple pass for max_evals batch evaluations:
 * batches of `batch` lexicographic triples from the first; the first triple
 * of the largest positive orc_best_g3 (p = 1..5) in a batch is applied, the
 * score raised by its delta, and the pass restarts from the first batch; a
 * pass without one perturbs the best mapping by 3 + rand()%3 uniform swaps
 * (orc_perturb mode 0, random_swap_k_vertices :822-858) and rescores.
 * Batches follow one another (the reference's stride of batch + total/3
 * leaves its index range).  rand() is orc_rng_next.  map ends on the best
 * mapping; out[0..7] = evals, passes, perturbations, moves, initial, final
 * (current), best score, triples evaluated. */
int orc_triples_pass(const orc_graphs *g, int32_t *map, uint64_t seed, int64_t max_evals, int batch, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif
