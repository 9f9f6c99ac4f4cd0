/*
 * bm_drivers.c -- local-search drivers over the device engine.  The move
 * order and tie-breaks are those of the reference loops they name; only the
 * delta evaluation moves to the GPU.
 */
#include <limits.h>
#include <string.h>

#include "bm_host.h"

static void record(int32_t *acc, int64_t cap, int64_t i, int32_t a, int32_t b, int32_t d) {
    if (acc && i < cap) {
        acc[3 * i] = a;
        acc[3 * i + 1] = b;
        acc[3 * i + 2] = d;
    }
}

/* apply, advance the score, and verify it like
 * graph-alignment-optimized.c:570-576 (the rescore is trusted) */
static int apply_and_verify(bm_ctx *ctx, int32_t a, int32_t b, int32_t d, int verify, bm_drive_stats *st) {
    int rc = bm_apply_swap(ctx, a, b);
    if (rc) return rc;
    st->final_score += d;
    st->accepted++;
    if (verify) {
        int64_t s;
        if ((rc = bm_score(ctx, &s))) return rc;
        if (s != st->final_score) {
            st->mismatches++;
            st->final_score = s;
        }
    }
    return BM_OK;
}

/* graph-alignment-optimized.c:513-581 */
int bm_drive_first_improvement(bm_ctx *ctx, int32_t rows_per_sweep, int64_t max_sweeps, int verify,
                               bm_row_hook hook, void *user, int32_t *acc, int64_t acc_cap,
                               bm_drive_stats *stats) {
    bm_drive_stats st;
    memset(&st, 0, sizeof st);
    int32_t n = bm_num_vertices(ctx);
    if (rows_per_sweep < n) rows_per_sweep = n;
    int rc = bm_score(ctx, &st.initial_score);
    if (rc) return rc;
    st.final_score = st.initial_score;
    int stop = 0;
    while (!stop && (max_sweeps <= 0 || st.sweeps < max_sweeps)) {
        int64_t before = st.accepted;
        for (int32_t a = 1; a <= rows_per_sweep && !stop; a++) {
            st.rows++;
            if (hook && hook(user, st.rows, a, st.final_score, st.accepted)) {
                stop = 1;
                break;
            }
            if (a > n) continue; /* padding rows: no edges, never accept */
            int32_t cursor = 1;
            while (cursor <= n) {
                int32_t b, d;
                if ((rc = bm_eval_swap_row(ctx, a, cursor, n, BM_ROW_FIRST_POSITIVE, &b, &d))) return rc;
                st.evals_rows++;
                if (b == 0) break;
                record(acc, acc_cap, st.accepted, a, b, d);
                if ((rc = apply_and_verify(ctx, a, b, d, verify, &st))) return rc;
                cursor = b + 1;
            }
        }
        st.sweeps++;
        if (st.accepted == before) break;
    }
    if (stats) *stats = st;
    return BM_OK;
}

/* alignment-omp.c:598-658 at one thread (the stagnation restart of
 * :660-678 is time-seeded and not reproduced: a pass without an accept ends
 * the run) */
int bm_drive_row_best(bm_ctx *ctx, int64_t max_sweeps, int verify, bm_row_hook hook, void *user, int32_t *acc,
                      int64_t acc_cap, bm_drive_stats *stats) {
    bm_drive_stats st;
    memset(&st, 0, sizeof st);
    int32_t n = bm_num_vertices(ctx);
    int rc = bm_score(ctx, &st.initial_score);
    if (rc) return rc;
    st.final_score = st.initial_score;
    int stop = 0;
    while (!stop && (max_sweeps <= 0 || st.sweeps < max_sweeps)) {
        int64_t before = st.accepted;
        for (int32_t a = 1; a <= n; a++) {
            st.rows++;
            if (hook && hook(user, st.rows, a, st.final_score, st.accepted)) {
                stop = 1;
                break;
            }
            if (a == n) continue;
            int32_t b, d;
            if ((rc = bm_eval_swap_row(ctx, a, a + 1, n, BM_ROW_ARGMAX, &b, &d))) return rc;
            st.evals_rows++;
            if (d > 0) {
                record(acc, acc_cap, st.accepted, a, b, d);
                if ((rc = apply_and_verify(ctx, a, b, d, verify, &st))) return rc;
            }
        }
        st.sweeps++;
        if (st.accepted == before) break;
    }
    if (stats) *stats = st;
    return BM_OK;
}

/* graph-re-alignment-gpu.cu:830-891 without the time-seeded restart */
int bm_drive_global_best(bm_ctx *ctx, int64_t max_steps, int verify, int32_t *acc, int64_t acc_cap,
                         bm_drive_stats *stats) {
    bm_drive_stats st;
    memset(&st, 0, sizeof st);
    int rc = bm_score(ctx, &st.initial_score);
    if (rc) return rc;
    st.final_score = st.initial_score;
    while (max_steps <= 0 || st.sweeps < max_steps) {
        int32_t a, b, d;
        if ((rc = bm_sweep_all_pairs(ctx, &a, &b, &d, NULL, NULL, NULL))) return rc;
        st.sweeps++;
        if (d <= 0) break;
        record(acc, acc_cap, st.accepted, a, b, d);
        if ((rc = apply_and_verify(ctx, a, b, d, verify, &st))) return rc;
    }
    if (stats) *stats = st;
    return BM_OK;
}
