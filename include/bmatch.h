/*
 * bmatch.h -- C ABI of the MI355X move-delta engine for brain-match's
 * weighted directed graph alignment.
 *
 * Score of a mapping p (G1 vertex -> G2 vertex, 0 = unmapped):
 *     S(p) = sum over G1 edge-list entries (u,v,w) of min(w, W2[p(u)][p(v)])
 * where W2 is G2's dense `short` adjacency (last CSV line for a pair wins).
 * A move is accepted by the callers when its integer delta is > 0.
 *
 * The reference has no library or FFI: its deltas are file-local C functions
 * called from each program's inner loop.  Each entry point below names the
 * reference routine (zhihuidu/brain-match, src/file:line) whose results it
 * reproduces bit for bit, and the call site it replaces.
 *
 * Conventions
 *   - Vertex ids are 0..n (n = largest id); 0 is the reference's "unmapped"
 *     sentinel, drivers sweep 1..n.  mapping[] has n+1 entries.
 *   - All pointers are host pointers; the library copies inputs and owns all
 *     device memory.  Outputs go to caller-provided buffers.
 *   - Every entry point returns BM_OK (0) or a negative BM_E* code; none
 *     calls exit().  bm_last_error() gives a message for the calling thread.
 *   - A context is driven by one host thread at a time.  All device work of
 *     a context runs on its own HIP stream; calls return after that work is
 *     complete.
 */
#ifndef BMATCH_H
#define BMATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BM_ABI_VERSION 1

enum {
    BM_OK = 0,
    BM_EINVAL = -1,  /* bad argument (vertex id, k, mode, size) */
    BM_ENOMEM = -2,  /* device or host allocation failed / over budget */
    BM_EHIP = -3,    /* a HIP runtime call failed */
    BM_ESTATE = -4,  /* graphs or mapping not uploaded yet */
    BM_ENODEV = -5,  /* no such HIP device */
    BM_ERANGE = -6,  /* value outside what the engine represents */
    BM_EIO = -7      /* file I/O or parse failure (host helpers) */
};

typedef struct bm_ctx bm_ctx;

const char *bm_strerror(int code);
const char *bm_last_error(void);
int bm_abi_version(void);

/* Context on HIP device `device`. */
int bm_create(int device, bm_ctx **out);
int bm_destroy(bm_ctx *ctx);

/*
 * Upload G1 and G2 as edge lists in CSV line order (duplicates allowed).
 * Replaces load_graph_from_csv + add_edge of graph-alignment-optimized.c:
 * 183-204, 635-679 (and the identical loaders of the permset and
 * parallel programs): G1 keeps every line in per-vertex out/in lists, the
 * dense G1 cell and all of G2 take the last line for a pair, truncated to
 * `short` as the reference's adj_matrix does.  Vertex ids must lie in
 * [0, n]; G1 weights must be >= -32768.  Device footprint is about
 * 8*(n+1)^2 bytes + 72*|E1| + 28*|E2| bytes (+4*(n+1)^2 for the all-pairs
 * sweep); bm_job_footprint gives the bound the serving loop uses.
 */
int bm_upload_graphs(bm_ctx *ctx, int32_t n,
                     const int32_t *g1_src, const int32_t *g1_dst, const int32_t *g1_w, int64_t e1,
                     const int32_t *g2_src, const int32_t *g2_dst, const int32_t *g2_w, int64_t e2);

int32_t bm_num_vertices(const bm_ctx *ctx);
/* bytes of device memory held by the context */
int64_t bm_device_bytes(const bm_ctx *ctx);

/* Set / read the mapping (n+1 entries, values in [0, n]).  set_mapping
 * rebuilds the device-resident permuted W2 (P[x][y] = W2[p(x)][p(y)]) and
 * its transpose. */
int bm_set_mapping(bm_ctx *ctx, const int32_t *mapping);
int bm_get_mapping(bm_ctx *ctx, int32_t *mapping);

/* S_opt of calculate_alignment_score, graph-alignment-optimized.c:292-307
 * (called after every accepted swap at :571 and in every dump at :485). */
int bm_score(bm_ctx *ctx, int64_t *score);

/* Delta_B of calculate_swap_delta, graph-alignment-optimized.c:321-408
 * (call sites :554, alignment-omp.c:613,
 * graph-alignment-optimized-parallel.c:557), for count pairs (a[i], b[i]),
 * a != b, both in [0, n]. */
int bm_eval_swaps(bm_ctx *ctx, const int32_t *a, const int32_t *b, int64_t count, int32_t *delta);

/* One row of the sweep: Delta_B(a, b) for b in [b_lo, b_hi], b != a.
 *   BM_ROW_FIRST_POSITIVE: smallest b with Delta_B > 0 -- the m2 loop of
 *     graph-alignment-optimized.c:542-579; *best_b = 0 if none.
 *   BM_ROW_ARGMAX: largest Delta_B, lowest b on ties -- the m2 loop of
 *     alignment-omp.c:601-622 at one thread; *best_b = 0 if range empty. */
#define BM_ROW_FIRST_POSITIVE 0
#define BM_ROW_ARGMAX 1
int bm_eval_swap_row(bm_ctx *ctx, int32_t a, int32_t b_lo, int32_t b_hi, int mode,
                     int32_t *best_b, int32_t *best_delta);

/* bm_eval_swap_row for the consecutive rows a, a + 1, ..., a + count - 1
 * (count <= BM_ROW_GROUP_MAX, a + count - 1 <= n), all under the current
 * mapping and in one launch: row a over [b_lo, n], each later row r over
 * [b_lo_next, n], or [r + 1, n] when b_lo_next < 0.  best_b[i], best_delta[i]
 * are row a + i's result as bm_eval_swap_row gives it.  Replaces the next
 * count evaluations of the m2 loops of graph-alignment-optimized.c:542-579
 * (b_lo_next = 1) and alignment-omp.c:601-622 (b_lo_next = -1) for as long
 * as no swap is accepted: a row's result holds only until a swap is applied,
 * so the drivers use the later rows' results while the earlier rows accept
 * nothing, which keeps the reference's move sequence. */
#define BM_ROW_GROUP_MAX 16
int bm_eval_swap_rows(bm_ctx *ctx, int32_t a, int32_t count, int32_t b_lo, int32_t b_lo_next, int mode,
                      int32_t *best_b, int32_t *best_delta);

/* bm_eval_swaps that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation).  No counterpart in the reference.  The pairs and
 * deltas are staged in a fixed 12 MiB device array that every GPU holds from
 * the first such call until the process ends, instead of the context's
 * staging buffer; count <= 1,048,576 (BM_EINVAL above).  The context must
 * have its graphs and mapping set.  Calls on one GPU share the array and are
 * serialised (each returns after its deltas are copied out). */
int bm_eval_swaps_noalloc(bm_ctx *ctx, const int32_t *a, const int32_t *b, int64_t count, int32_t *delta);

/* The swap delta of the dense-matrix GPU kernels, calculateSwapDeltasKernel
 * and calculateSwapDeltasKernelQuick (quick.cu:48-250; the same arithmetic
 * in graph-re-alignment-gpu.cu:60-128): for every i in [1, n] other than a
 * and b, the terms of the G1 cells W1[a][i], W1[i][a], W1[b][i], W1[i][b]
 * that are positive as short (last CSV line of a pair wins), then the two
 * cells between a and b.  It equals Delta_B when G1 has no self-loop, no
 * repeated line, no line at vertex 0 and every weight in [1, 32767];
 * otherwise it is that kernel's value, not the score change.  Pairs in
 * [1, n], a != b. */
int bm_eval_swaps_dense(bm_ctx *ctx, const int32_t *a, const int32_t *b, int64_t count, int32_t *delta);

/* Whole rows: deltas[i * (n + 1) + b] = Delta(rows[i], b) for b in [1, n],
 * b != rows[i]; the entries b = 0 and b = rows[i] are 0.  These are the
 * pairs calculateSwapDeltasKernelQuick (quick.cu:141-250) evaluates after a
 * perturbation or an accepted swap: every pair that touches one of the
 * changed vertices (rows distinct, in [1, n], count <= 64).  kind selects
 * the delta: BM_DELTA_LISTS = Delta_B of bm_eval_swaps, BM_DELTA_DENSE = the
 * quick.cu kernel's own delta of bm_eval_swaps_dense. */
#define BM_DELTA_LISTS 0
#define BM_DELTA_DENSE 1
int bm_eval_rows(bm_ctx *ctx, const int32_t *rows, int32_t count, int32_t kind, int32_t *deltas);

/* Perturbation by k disjoint swaps, random_swap_k_vertices:
 *   BM_PERTURB_OMP      alignment-omp.c:495-541: 2k distinct indices drawn
 *                       uniformly from [0, n-1] (index 0 can be drawn and n
 *                       cannot, as there; 0 is the unused vertex slot)
 *   BM_PERTURB_WEIGHTED quick.cu:816-887: k vertices drawn one at a time
 *                       without replacement with probability proportional
 *                       to 0.7 x (sum of positive dense G1 cells of the
 *                       out-lines) + 0.3 x (the same over in-lines),
 *                       float arithmetic as quick.cu:766-814, then k more
 *                       uniformly from [1, n]
 * then mapping[sel[i]] <-> mapping[sel[i + k]] for i < k.  The draws come
 * from *rng, a splitmix64 state the call advances, in place of the
 * reference's time-seeded rand() (bm_rng_next gives the same stream to a
 * caller).  selected (2k entries, may be NULL) receives the drawn vertices
 * in draw order.  Requires 1 <= k and 2k <= n. */
#define BM_PERTURB_OMP 0
#define BM_PERTURB_WEIGHTED 1
int bm_perturb(bm_ctx *ctx, int32_t k, int mode, uint64_t *rng, int32_t *selected);

/* splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014): advances *state and returns the next value. */
uint64_t bm_rng_next(uint64_t *state);

/* All-pairs sweep over 1 <= a < b <= n with the current mapping: the
 * launch + host arg-max of graph-re-alignment-gpu.cu:837-868 (lexicographic
 * pair order, strict >, so the lowest (a,b) wins ties; *best_delta = 0 and
 * *best_a = *best_b = 0 when no pair is positive), computed with Delta_B.
 * Optional outputs (NULL to skip): per-row best b > a and its delta
 * (row_best_b[a] = 0, row_best_delta[a] = INT32_MIN when a has no b > a),
 * and the full symmetric delta matrix ((n+1)^2 int32, zero diagonal and
 * zero row/column 0) for small n. */
int bm_sweep_all_pairs(bm_ctx *ctx, int32_t *best_a, int32_t *best_b, int32_t *best_delta,
                       int32_t *row_best_b, int32_t *row_best_delta, int32_t *delta_matrix);

/* bm_sweep_all_pairs with the delta the reference's GPU kernel computes,
 * calculateSwapDeltasKernel (graph-re-alignment-gpu.cu:60-128,
 * quick.cu:48-138): the dense-matrix form of bm_eval_swaps_dense for every
 * pair, same outputs and tie-break.  Runs the dense all-pairs kernel over the
 * dense-form lists (n+1 x ld int32 of extra device memory). */
int bm_sweep_all_pairs_dense(bm_ctx *ctx, int32_t *best_a, int32_t *best_b, int32_t *best_delta,
                             int32_t *row_best_b, int32_t *row_best_delta, int32_t *delta_matrix);

/* Delta_G3 of calculateThreeNodeSwapKernel, triple.cu:85-214, per triple
 * (m1,m2,m3) of distinct vertices in [1, n]: best over the permutations
 * (strict >, from 0) and its index p (0 = none positive).
 *   BM_ROT3_ROTATIONS: p in {3, 4}, the two 3-cycles (f2,f3,f1), (f3,f1,f2)
 *   BM_ROT3_ALL: p in 1..5, as triple.cu */
#define BM_ROT3_ROTATIONS 0
#define BM_ROT3_ALL 1
int bm_eval_rot3(bm_ctx *ctx, const int32_t *triples, int64_t count, int mode,
                 int32_t *best_delta, int32_t *best_perm);

/* Delta_P of calculate_permutation_delta,
 * graph-alignment-permset-parallel.c:304-339, evaluated for every one of
 * the k! permutations of each k-set (2 <= k <= 8, distinct vertices in
 * [1, n]) in next_permutation order (:342-363, :563-571); per set the best
 * delta (strict >, from 0; :665-674) and the lexicographic rank of its
 * permutation (-1 when none is positive).  bm_perm_from_rank decodes it. */
int bm_eval_permset(bm_ctx *ctx, int32_t k, const int32_t *sets, int64_t count,
                    int32_t *best_delta, int32_t *best_rank);
int bm_perm_from_rank(int32_t k, int32_t rank, int32_t *perm);

/* Delta_A of refine-matching-hillclimb.c:244-296 (out-edges of a and b
 * only, de-duplicated G1, edge a<->b skipped) for count pairs; call site
 * :320.  Unmapped vertices (mapping 0) contribute 0. */
int bm_eval_swaps_hillclimb(bm_ctx *ctx, const int32_t *a, const int32_t *b, int64_t count, int32_t *delta);

/* Moves.  Update the mapping and patch the permuted W2 rows/columns on the
 * device.  apply_swap: graph-alignment-optimized.c:561-562;
 * apply_rot3: apply_permutation, triple.cu:903-934 (p in 1..5);
 * apply_perm: apply_best_permutation,
 * graph-alignment-permset-parallel.c:508-518 (new p(set[i]) =
 * old p(set[perm[i]])).  They return once the move is enqueued on the
 * context's stream; every later call on the context sees it. */
int bm_apply_swap(bm_ctx *ctx, int32_t a, int32_t b);
/* bm_apply_swap that also returns the exact score change S(after) - S(before),
 * summed over the G1 lines with an end at a or b (each line once): what the
 * full rescore after an accepted swap (graph-alignment-optimized.c:571,
 * alignment-omp.c:652, quick.cu:1248) establishes, at the cost of those
 * lines only. */
int bm_apply_swap_scored(bm_ctx *ctx, int32_t a, int32_t b, int64_t *change);
int bm_apply_rot3(bm_ctx *ctx, int32_t m1, int32_t m2, int32_t m3, int32_t p);
int bm_apply_perm(bm_ctx *ctx, int32_t k, const int32_t *set, const int32_t *perm);

/* Device-resident candidate list.  bm_moves_set copies `count` moves of
 * size k to the device once: k = 2 pairs (a, b) scored with Delta_B, k = 3
 * triples scored with Delta_G3, 4 <= k <= 8 sets scored with Delta_P over
 * all k! permutations (the move kinds of bm_eval_swaps, bm_eval_rot3 and
 * bm_eval_permset).  bm_moves_eval scores the whole list with the current
 * mapping on the device and reduces it there: the best move (largest delta,
 * lowest index on ties, only if > 0; *best_index = -1 otherwise), its delta
 * and aux (the permutation index p of triple.cu:85-92 for k = 3, selected by
 * `mode` as in bm_eval_rot3; the lexicographic rank for k >= 4; 0 for pairs).
 * bm_moves_results copies the per-move deltas / aux back.  This is the
 * launch + host arg-max of triple.cu:1082-1126 and quick.cu:1175-1217 with
 * the list and the reduction kept on the GPU. */
int bm_moves_set(bm_ctx *ctx, int32_t k, const int32_t *verts, int64_t count);
int bm_moves_eval(bm_ctx *ctx, int mode, int64_t *best_index, int32_t *best_delta, int32_t *best_aux);
int bm_moves_results(bm_ctx *ctx, int32_t *delta, int32_t *aux);

/*
 * Batched refinement.  Each step runs on the device without a host round
 * trip and is captured once in a hipGraph: candidates of all three move
 * kinds are scored with the current mapping (the row-best swap of every row
 * of the all-pairs Delta_B sweep, `triples` 3-vertex moves drawn from 2-hop
 * G1 neighbourhoods and scored with Delta_G3 over p = 1..5, `sets` k-sets
 * grown along G1 lines and scored over all k! permutations with intra-set
 * lines counted once); moves are selected in `rounds` rounds (largest
 * delta first, lowest candidate index on ties): a round's winners share no
 * vertex and no G1 line with each other and no vertex with earlier winners,
 * and each is scored exactly given the earlier rounds' winners (its lines to
 * the vertices they moved re-scored) and kept only if that change is
 * positive, so the changes add up exactly; they are applied, the step's
 * state is brought up to date, and S_opt is checked to equal the previous
 * score plus the applied changes.  Only vertices without self-loops or repeated lines move, where
 * every delta kind is the exact score change.  Results are deterministic for
 * a given instance, mapping and parameters.
 *
 * The reference applies one move per evaluation pass (host arg-max of
 * triple.cu:1082-1126, quick.cu:1175-1217, permset :651-730); this mode has
 * no counterpart there.
 */
typedef struct bm_refine_params {
    int32_t pairs;           /* 1: row-best swaps of the all-pairs sweep are candidates */
    int32_t triples;         /* 3-vertex candidates generated per step, 0..2^24 */
    int32_t sets;            /* k-set candidates generated per step, 0..2^20 */
    int32_t set_k;           /* k of the sets, 2..8 */
    int32_t set_gen;         /* 0: sets grown along G1 lines from a random vertex; 1: uniform random sets */
    int32_t rounds;          /* selection rounds per step, 1..64; 0: 64 when n >= 8,000, else 32 */
    int32_t steps_per_check; /* steps enqueued between host checks of the stop rules */
    uint64_t seed;           /* candidate generator seed */
} bm_refine_params;

typedef struct bm_refine_stats {
    int64_t steps, moves;                         /* in this run */
    int64_t pair_moves, triple_moves, set_moves;
    int64_t initial_score, final_score;           /* exact S_opt */
    int64_t idle_steps;                           /* trailing steps without a move */
    int64_t bad_step;                             /* first step failing the exact check, -1 if none */
    int32_t converged;                            /* stopped by `patience` */
    int32_t verified;                             /* every step passed the exact check */
} bm_refine_stats;

typedef struct bm_refine_step {
    int64_t score; /* exact S_opt after the step */
    int64_t gain;  /* sum of the applied deltas */
    int32_t moves; /* moves applied */
    int32_t ok;    /* score == previous score + gain, and rose if moves > 0 */
    int32_t moved; /* vertices whose image changed */
    int32_t full;  /* 1: the step began with a whole all-pairs sweep; 0: the gain matrix was kept current */
} bm_refine_step;

/* Allocates the step's buffers, autotunes kernel variants (results do not
 * depend on the choice) and captures the step graph.  Needs a mapping. */
int bm_refine_configure(bm_ctx *ctx, const bm_refine_params *params);
/* Runs steps from the current mapping until max_steps, until `moves` reaches
 * move_budget (> 0), or until `patience` consecutive steps applied nothing;
 * the rules are checked every steps_per_check steps.  The mapping (and
 * bm_get_mapping) then holds the result.  BM_ERANGE if an exact check failed. */
int bm_refine_run(bm_ctx *ctx, int64_t max_steps, int64_t move_budget, int32_t patience, bm_refine_stats *stats);
/* Runs `count` configured contexts (graph pairs of any sizes, one device) as
 * bm_refine_run would run each alone, all at once from the calling thread:
 * each context's step graph is replayed on its own stream, steps_per_check
 * steps per chunk, and a context whose chunk has finished gets its stop rules
 * applied and its next chunk enqueued while the others run.  stats[i] is
 * context i's result (identical to its solo run); BM_ERANGE (the first
 * failing context) when an exact check failed, after every context ended. */
int bm_refine_run_many(bm_ctx *const *ctxs, int32_t count, int64_t max_steps, int64_t move_budget, int32_t patience,
                       bm_refine_stats *stats);
/* Per-step records of the last run, steps [first, first+count) (the last 4096 are kept). */
int bm_refine_history(bm_ctx *ctx, int64_t first, int64_t count, bm_refine_step *out);
/* Keep a log of up to `capacity` applied moves per run (0: off).  Each record
 * is 22 int32: step, kind (0 pair, 1 triple, 2 set), k, delta (the exact
 * change given the earlier rounds' winners of the step), round, base (the
 * candidate's delta at the step's starting mapping), vertices[8], their new
 * images[8].  bm_refine_moves copies min(count, capacity) records. */
int bm_refine_log(bm_ctx *ctx, int32_t capacity);
int bm_refine_moves(bm_ctx *ctx, int64_t *count, int32_t *records, int64_t capacity);
/* Autotuned variants in use: lanes per triple (8/16/32/64), threads per set (64/256). */
int bm_refine_config(bm_ctx *ctx, int32_t *rot3_lanes, int32_t *permset_threads);

/*
 * Serving loop.  A server owns one device and takes refinement jobs (graph
 * pair, initial mapping, move budget, refinement parameters) as they arrive:
 * bm_server_submit copies the job and returns at once.  Preparation threads
 * admit queued jobs in arrival order while the job's device footprint fits
 * the HBM budget and an active slot is free (upload, mapping, step-graph
 * capture); the scheduler thread keeps every admitted job's step chunks
 * enqueued on the job's own stream, so jobs of any size overlap on the
 * device, and retires a job when its stop rule holds (move budget, patience,
 * step limit), freeing its slot and memory for the next.  Each job's result
 * is what bm_refine_run gives for it alone: the stop rules are checked at the
 * same step counts whatever else runs.  No counterpart in the reference.
 */
typedef struct bm_server bm_server;

typedef struct bm_server_config {
    int32_t max_active;     /* jobs on the device at once (>= 1) */
    int32_t prep_threads;   /* admission threads (>= 1) */
    int64_t hbm_budget;     /* bytes; 0 = the device's free memory at creation */
} bm_server_config;

typedef struct bm_job_spec {
    int32_t n;
    const int32_t *g1_src, *g1_dst, *g1_w;
    int64_t e1;
    const int32_t *g2_src, *g2_dst, *g2_w;
    int64_t e2;
    const int32_t *mapping;  /* n+1 entries */
    int64_t move_budget;     /* stop once this many moves were applied (0: no limit) */
    int64_t max_steps;       /* 0: no limit */
    int32_t patience;        /* stop after this many steps without a move (>= 1) */
    bm_refine_params params;
} bm_job_spec;

typedef struct bm_job_result {
    int64_t job_id;
    int32_t status;          /* BM_OK or the error code of the failing call */
    int32_t n;
    int64_t initial_score, final_score, steps, moves;
    int64_t device_bytes;    /* footprint while active */
    double t_submit, t_admit, t_start, t_done;  /* seconds since bm_server_create */
    char error[160];         /* message of the failing call ("" when status is BM_OK) */
} bm_job_result;

typedef struct bm_server_stats {
    int64_t submitted, done, failed, queued, active;
    int64_t peak_active;
    int64_t reserved_bytes, peak_reserved_bytes, hbm_budget;
} bm_server_stats;

/* Upper bound of the device bytes a job of this shape holds while active. */
int64_t bm_job_footprint(int32_t n, int64_t e1, int64_t e2, const bm_refine_params *params);
int bm_server_create(int device, const bm_server_config *config, bm_server **out);
/* BM_ENOMEM when the job alone exceeds the budget. */
int bm_server_submit(bm_server *srv, const bm_job_spec *job, int64_t *job_id);
/* Blocks until the job is done; mapping (n+1 entries) may be NULL.  A job's
 * result can be collected once. */
int bm_server_wait(bm_server *srv, int64_t job_id, bm_job_result *result, int32_t *mapping);
int bm_server_stats_get(bm_server *srv, bm_server_stats *stats);
/* Waits for every submitted job, then frees the server. */
int bm_server_destroy(bm_server *srv);

/* Which all-pairs gain kernel bm_sweep_all_pairs runs for this graph pair,
 * as chosen on its first call (0 before that): *kernel = 1 dense slab kernel
 * (with *dense_cols columns per lane) or 2 sparse kernel.  No counterpart in
 * the reference. */
int bm_sweep_config(bm_ctx *ctx, int *kernel, int *dense_cols);

/* Kernel timing.  When enabled, the engine brackets its kernels with HIP
 * events on the context stream; bm_kernel_time reports the launch count
 * and summed device milliseconds of the named kernel since the last reset.
 * No counterpart in the reference (it logs wall-clock seconds only). */
int bm_set_timing(bm_ctx *ctx, int enable);
int bm_kernel_time(bm_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms);
int bm_reset_timing(bm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* BMATCH_H */
