/*
 * bm_drivers.c -- local-search drivers over the device engine.  The move
 * order and tie-breaks are those of the reference loops they name; only the
 * delta evaluation moves to the GPU.
 */
#include <limits.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bm_host.h"
#include "bm_private.h"

static void record(int32_t *acc, int64_t cap, int64_t i, int32_t a, int32_t b, int32_t d) {
    if (acc && i < cap) {
        acc[3 * i] = a;
        acc[3 * i + 1] = b;
        acc[3 * i + 2] = d;
    }
}

static __thread bm_mismatch_hook mismatch_hook;
static __thread void *mismatch_user;

void bm_drive_on_mismatch(bm_mismatch_hook hook, void *user) {
    mismatch_hook = hook;
    mismatch_user = user;
}

static void report_mismatch(int64_t calculated, int64_t verified) {
    if (mismatch_hook) mismatch_hook(mismatch_user, calculated, verified);
}

/* A verified swap whose exact score change has not been read yet. */
typedef struct {
    int pending;
    int64_t before; /* the score before the swap */
} pending_check;

/* the check of graph-alignment-optimized.c:570-576 for the last verified
 * swap: the verified score (the exact score change of the swap added to the
 * previous score, which equals the full rescore there) is trusted.  Waits for
 * the engine's stream if the change is not in yet; the drivers call it after
 * their next evaluation, which has waited already. */
static int settle(bm_ctx *ctx, bm_drive_stats *st, pending_check *pc) {
    if (!pc->pending) return BM_OK;
    int64_t change;
    int rc = bm_deferred_change(ctx, &change);
    if (rc) return rc;
    pc->pending = 0;
    if (pc->before + change != st->final_score) {
        st->mismatches++;
        report_mismatch(st->final_score, pc->before + change);
        st->final_score = pc->before + change;
    }
    return BM_OK;
}

/* apply and advance the score by the delta; with verify, the exact change is
 * enqueued and checked by the next settle() */
static int apply_and_verify(bm_ctx *ctx, int32_t a, int32_t b, int32_t d, int verify, bm_drive_stats *st,
                            pending_check *pc) {
    int rc = verify ? bm_apply_swap_deferred(ctx, a, b) : bm_apply_swap(ctx, a, b);
    if (rc) return rc;
    if (verify) {
        pc->pending = 1;
        pc->before = st->final_score;
    }
    st->final_score += d;
    st->accepted++;
    return BM_OK;
}

/* Rows evaluated ahead of the current one (bm_eval_swap_rows): their results
 * hold while the mapping stays as it was, so any applied swap drops them.  A
 * row that ends without an accept doubles the rows taken per launch (up to
 * BM_ROW_GROUP_MAX), an accept sets it back to one: while most rows accept,
 * nothing is evaluated in vain; once they stop accepting, one launch covers
 * up to BM_ROW_GROUP_MAX rows. */
typedef struct {
    int32_t first, count, width;
    int32_t b[BM_ROW_GROUP_MAX], d[BM_ROW_GROUP_MAX];
} rows_ahead;

static void ahead_drop(rows_ahead *ra) {
    ra->count = 0;
    ra->width = 1;
}

static void ahead_grow(rows_ahead *ra) {
    ra->width = ra->width * 2 > BM_ROW_GROUP_MAX ? BM_ROW_GROUP_MAX : ra->width * 2;
}

/* row a over [lo, n]; fresh: lo is the row's first lower end, the one the rows
 * ahead were evaluated with (lo_next, -1 for a + 1) */
static int eval_row_ahead(bm_ctx *ctx, rows_ahead *ra, int32_t n, int32_t a, int32_t lo, int32_t lo_next, int mode,
                          int fresh, int32_t *b, int32_t *d) {
    if (fresh && ra->count > 0 && a >= ra->first && a < ra->first + ra->count) {
        *b = ra->b[a - ra->first];
        *d = ra->d[a - ra->first];
        return BM_OK;
    }
    int32_t cnt = ra->width < n - a + 1 ? ra->width : n - a + 1;
    int32_t bb[BM_ROW_GROUP_MAX], dd[BM_ROW_GROUP_MAX];
    int rc = bm_eval_swap_rows(ctx, a, cnt, lo, lo_next, mode, bb, dd);
    if (rc) return rc;
    ra->first = a + 1;
    ra->count = cnt - 1;
    memcpy(ra->b, bb + 1, sizeof(int32_t) * (size_t)(cnt - 1));
    memcpy(ra->d, dd + 1, sizeof(int32_t) * (size_t)(cnt - 1));
    *b = bb[0];
    *d = dd[0];
    return BM_OK;
}

/* A stretch of rows without an accept asks one all-pairs sweep whether any
 * swap improves the current mapping.  If none does, no row can accept before
 * the mapping changes, and the mapping changes only by an accept, so the rest
 * of the pass is searched in vain: the reference finds nothing in those rows,
 * and the driver only counts them.  The sweep is asked once per stretch,
 * after at least BM_QUIET_ROWS rows and once the stretch has taken as long as
 * the last sweep did (about 1.8 ms at n = 18,524 on the sparse kernel, 15-26
 * ms on the dense one, against about 25 us per row searched): a sweep that
 * finds an improving swap then costs at most as much as the rows before it. */
#define BM_QUIET_ROWS 48

typedef struct {
    int64_t quiet;  /* rows searched since the last accept */
    double t0;      /* when the stretch began */
    double sweep_s; /* the last sweep's wall time */
    int asked;      /* the sweep was asked in this stretch */
    int cleared;    /* it found no improving swap: the pass's remaining rows are skipped */
} quiet_state;

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void quiet_accept(quiet_state *q) {
    q->quiet = 0;
    q->t0 = now_s();
    q->asked = 0;
    q->cleared = 0;
}

static void quiet_init(quiet_state *q) {
    q->sweep_s = 0.0;
    quiet_accept(q);
}

/* called before searching a row: 1 if the row cannot accept */
static int quiet_cleared(bm_ctx *ctx, quiet_state *q, int *rc) {
    *rc = BM_OK;
    if (q->cleared) return 1;
    if (q->asked || q->quiet < BM_QUIET_ROWS) return 0;
    const double t = now_s();
    if (t - q->t0 < q->sweep_s) return 0;
    q->asked = 1;
    int32_t a, b, d;
    if ((*rc = bm_sweep_all_pairs(ctx, &a, &b, &d, NULL, NULL, NULL))) return 0;
    q->sweep_s = now_s() - t;
    q->cleared = d <= 0;
    return q->cleared;
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
    pending_check pc = {0, 0};
    rows_ahead ra;
    ahead_drop(&ra);
    quiet_state qs;
    quiet_init(&qs);
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
            if (quiet_cleared(ctx, &qs, &rc)) {
                st.rows_cleared++;
                continue;
            }
            if (rc || (rc = settle(ctx, &st, &pc))) return rc;
            qs.quiet++;
            int32_t cursor = 1;
            while (cursor <= n) {
                int32_t b, d;
                /* the rows after a from 1, as their own first evaluation will be */
                if ((rc = eval_row_ahead(ctx, &ra, n, a, cursor, 1, BM_ROW_FIRST_POSITIVE, cursor == 1, &b, &d)))
                    return rc;
                if ((rc = settle(ctx, &st, &pc))) return rc;
                st.evals_rows++;
                if (b == 0) {
                    ahead_grow(&ra);
                    break;
                }
                record(acc, acc_cap, st.accepted, a, b, d);
                ahead_drop(&ra);
                quiet_accept(&qs);
                if ((rc = apply_and_verify(ctx, a, b, d, verify, &st, &pc))) return rc;
                cursor = b + 1;
            }
            if ((rc = settle(ctx, &st, &pc))) return rc;
        }
        st.sweeps++; /* a clearance holds into the next pass: the mapping is unchanged */
        if (st.accepted == before) break;
    }
    if ((rc = settle(ctx, &st, &pc))) return rc;
    if (stats) *stats = st;
    return BM_OK;
}

/* alignment-omp.c:598-658 at one thread; a pass without an accept ends the
 * run (bm_drive_row_best_restart below adds the restarts of :660-678) */
int bm_drive_row_best(bm_ctx *ctx, int64_t max_sweeps, int verify, bm_row_hook hook, void *user, int32_t *acc,
                      int64_t acc_cap, bm_drive_stats *stats) {
    bm_drive_stats st;
    memset(&st, 0, sizeof st);
    int32_t n = bm_num_vertices(ctx);
    int rc = bm_score(ctx, &st.initial_score);
    if (rc) return rc;
    st.final_score = st.initial_score;
    pending_check pc = {0, 0};
    rows_ahead ra;
    ahead_drop(&ra);
    quiet_state qs;
    quiet_init(&qs);
    int stop = 0;
    while (!stop && (max_sweeps <= 0 || st.sweeps < max_sweeps)) {
        int64_t before = st.accepted;
        for (int32_t a = 1; a <= n; a++) {
            st.rows++;
            if (hook && (rc = settle(ctx, &st, &pc))) return rc;  /* the hook sees the verified score */
            if (hook && hook(user, st.rows, a, st.final_score, st.accepted)) {
                stop = 1;
                break;
            }
            if (a == n) continue;
            if (quiet_cleared(ctx, &qs, &rc)) {
                st.rows_cleared++;
                continue;
            }
            if (rc) return rc;
            qs.quiet++;
            int32_t b, d;
            if ((rc = eval_row_ahead(ctx, &ra, n, a, a + 1, -1, BM_ROW_ARGMAX, 1, &b, &d))) return rc;
            if ((rc = settle(ctx, &st, &pc))) return rc;
            st.evals_rows++;
            if (d > 0) {
                record(acc, acc_cap, st.accepted, a, b, d);
                ahead_drop(&ra);
                quiet_accept(&qs);
                if ((rc = apply_and_verify(ctx, a, b, d, verify, &st, &pc))) return rc;
            } else {
                ahead_grow(&ra);
            }
        }
        if ((rc = settle(ctx, &st, &pc))) return rc;
        st.sweeps++;
        if (st.accepted == before) break;
    }
    if (stats) *stats = st;
    return BM_OK;
}

/* graph-re-alignment-gpu.cu:830-891 without the time-seeded restart */
int bm_drive_global_best(bm_ctx *ctx, int32_t kind, int64_t max_steps, int verify, int32_t *acc, int64_t acc_cap,
                         bm_drive_stats *stats) {
    if (kind != BM_DELTA_LISTS && kind != BM_DELTA_DENSE) return BM_EINVAL;
    bm_drive_stats st;
    memset(&st, 0, sizeof st);
    int rc = bm_score(ctx, &st.initial_score);
    if (rc) return rc;
    st.final_score = st.initial_score;
    pending_check pc = {0, 0};
    while (max_steps <= 0 || st.sweeps < max_steps) {
        int32_t a, b, d;
        rc = kind == BM_DELTA_DENSE ? bm_sweep_all_pairs_dense(ctx, &a, &b, &d, NULL, NULL, NULL)
                                    : bm_sweep_all_pairs(ctx, &a, &b, &d, NULL, NULL, NULL);
        if (rc) return rc;
        if ((rc = settle(ctx, &st, &pc))) return rc;
        st.sweeps++;
        if (d <= 0) break;
        record(acc, acc_cap, st.accepted, a, b, d);
        if ((rc = apply_and_verify(ctx, a, b, d, verify, &st, &pc))) return rc;
    }
    if ((rc = settle(ctx, &st, &pc))) return rc;
    if (stats) *stats = st;
    return BM_OK;
}

/* alignment-omp.c:543-678: row-best passes, perturbation at the start and
 * after an idle pass, the best mapping kept */
int bm_drive_row_best_restart(bm_ctx *ctx, const bm_restart_params *p, int32_t *best_mapping,
                              bm_restart_stats *stats) {
    if (!ctx || !p || p->max_passes < 1 || p->initial_swaps < 0 || p->restart_swaps < 1 || p->idle_passes < 0)
        return BM_EINVAL;
    bm_restart_stats st;
    memset(&st, 0, sizeof st);
    const int32_t n = bm_num_vertices(ctx);
    int32_t *best = malloc(sizeof(int32_t) * ((size_t)n + 1));
    if (!best) return BM_ENOMEM;
    uint64_t rng = p->seed;
    int64_t cur = 0;
    int rc;
    if ((rc = bm_get_mapping(ctx, best)) || (rc = bm_score(ctx, &st.initial_score))) goto out;
    st.best_score = st.initial_score;
    if (p->initial_swaps > 0 && (rc = bm_perturb(ctx, p->initial_swaps, BM_PERTURB_OMP, &rng, NULL))) goto out;
    if ((rc = bm_score(ctx, &cur))) goto out;
    int64_t pass = 0, improvements = 0, last = 0;
    rows_ahead ra;
    quiet_state qs;
    quiet_init(&qs);
    while (st.passes < p->max_passes) {
        ahead_drop(&ra); /* the mapping may have been reset and perturbed */
        for (int32_t a = 1; a < n; a++) { /* :571-655; row n has no b > a */
            int32_t b, d;
            if (quiet_cleared(ctx, &qs, &rc)) continue;
            if (rc) goto out;
            qs.quiet++;
            if ((rc = eval_row_ahead(ctx, &ra, n, a, a + 1, -1, BM_ROW_ARGMAX, 1, &b, &d))) goto out;
            st.evals_rows++;
            if (d <= 0) {
                ahead_grow(&ra);
                continue;
            }
            ahead_drop(&ra);
            quiet_accept(&qs);
            int64_t change = 0;
            if ((rc = p->verify ? bm_apply_swap_scored(ctx, a, b, &change) : bm_apply_swap(ctx, a, b))) goto out;
            const int64_t verified = cur + change;
            cur += d;
            improvements++;
            st.accepted++;
            if (p->verify && verified != cur) { /* :652-656, the verified score is trusted */
                st.mismatches++;
                report_mismatch(cur, verified);
                cur = verified;
            }
            if (cur > st.best_score) { /* :627-636 */
                st.best_score = cur;
                if ((rc = bm_get_mapping(ctx, best))) goto out;
            }
        }
        pass++;
        st.passes++;
        if (pass > p->idle_passes && improvements - last == 0) { /* :660-673 */
            if ((rc = bm_set_mapping(ctx, best))) goto out;
            if ((rc = bm_perturb(ctx, p->restart_swaps, BM_PERTURB_OMP, &rng, NULL))) goto out;
            if ((rc = bm_score(ctx, &cur))) goto out;
            improvements = last = pass = 0;
            st.restarts++;
            quiet_accept(&qs); /* a new mapping */
        }
        if (last < improvements) last = improvements; /* :674-676 */
    }
    st.final_score = cur;
    rc = bm_set_mapping(ctx, best);
out:
    if (!rc && best_mapping) memcpy(best_mapping, best, sizeof(int32_t) * ((size_t)n + 1));
    free(best);
    if (!rc && stats) *stats = st;
    return rc;
}

/* Pairs quick.cu's kernel evaluates after a change (:151-196), in its
 * numbering: inside the sorted changed set (i < j), then each changed vertex
 * with every unchanged one ascending.  Scans them like :1172-1191 / :1280-1299:
 * the first strict maximum above 0, and the number of positive and of zero
 * pairs; the pick-th positive or zero pair in that order is found by a second
 * scan when the search draws one (quick_pick), so the first scan keeps no
 * lists and takes no data-dependent branch per pair. */
typedef struct {
    int32_t *rows;       /* nc * (n + 1) deltas from bm_eval_rows; changed columns cleared */
    int32_t n, nc;
    int32_t changed[8];
    int32_t inside[28];  /* the deltas inside the changed set, in scan order */
    int64_t npos, nzero, pairs;
    int32_t best_a, best_b, max_delta;
} quick_scan;

static int quick_eval(bm_ctx *ctx, int32_t n, const int32_t *changed, int nc, int32_t kind, quick_scan *q) {
    int rc = bm_eval_rows(ctx, changed, nc, kind, q->rows);
    if (rc) return rc;
    const size_t N1 = (size_t)n + 1;
    q->n = n;
    q->nc = nc;
    memcpy(q->changed, changed, sizeof(int32_t) * (size_t)nc);
    q->pairs += (int64_t)nc * (nc - 1) / 2 + (int64_t)nc * (n - nc);
    int64_t npos = 0, nzero = 0;
    int32_t maxd = 0, ba = 0, bb = 0;
    int m = 0;
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++) {
            const int32_t d = q->rows[i * N1 + changed[j]];
            q->inside[m++] = d;
            if (d > maxd) {
                maxd = d;
                ba = changed[i];
                bb = changed[j];
            }
            npos += d > 0;
            nzero += d == 0;
        }
    /* the changed columns are not unchanged vertices: neither positive nor zero */
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) q->rows[i * N1 + changed[j]] = INT32_MIN;
    for (int i = 0; i < nc; i++) {
        const int32_t *r = q->rows + i * N1;
        int64_t cp = 0, cz = 0;
        int32_t rmax = INT32_MIN;
        for (int32_t v = 1; v <= n; v++) {
            const int32_t d = r[v];
            cp += d > 0;
            cz += d == 0;
            rmax = d > rmax ? d : rmax;
        }
        npos += cp;
        nzero += cz;
        if (rmax > maxd) { /* the first v holding the row's maximum */
            int32_t v = 1;
            while (r[v] != rmax) v++;
            maxd = rmax;
            ba = changed[i];
            bb = v;
        }
    }
    q->npos = npos;
    q->nzero = nzero;
    q->max_delta = maxd;
    q->best_a = ba;
    q->best_b = bb;
    return BM_OK;
}

/* the pick-th positive (positive = 1) or zero pair of the last scan, in scan order */
static void quick_pick(const quick_scan *q, int positive, int64_t pick, int32_t *a, int32_t *b) {
    const size_t N1 = (size_t)q->n + 1;
    int64_t k = 0;
    int m = 0;
    for (int i = 0; i < q->nc; i++)
        for (int j = i + 1; j < q->nc; j++) {
            const int32_t d = q->inside[m++];
            if (positive ? d > 0 : d == 0) {
                if (k++ == pick) {
                    *a = q->changed[i];
                    *b = q->changed[j];
                    return;
                }
            }
        }
    for (int i = 0; i < q->nc; i++) {
        const int32_t *r = q->rows + i * N1;
        for (int32_t v = 1; v <= q->n; v++) {
            const int32_t d = r[v];
            k += positive ? d > 0 : d == 0;
            if (k > pick) {
                *a = q->changed[i];
                *b = v;
                return;
            }
        }
    }
    *a = *b = 0; /* not reached: pick < the count of the scan */
}

static void sort_small(int32_t *a, int m) {
    for (int i = 1; i < m; i++)
        for (int j = i; j > 0 && a[j - 1] > a[j]; j--) {
            int32_t t = a[j];
            a[j] = a[j - 1];
            a[j - 1] = t;
        }
}

/* quick.cu:1137-1324 */
int bm_drive_quick(bm_ctx *ctx, const bm_quick_params *p, int32_t *best_mapping, bm_quick_stats *stats) {
    if (!ctx || !p || p->max_evals < 1 || p->zero_limit < 1 || p->best_window < 0 || !(p->gap_norm > 0) ||
        (p->delta_kind != BM_DELTA_LISTS && p->delta_kind != BM_DELTA_DENSE))
        return BM_EINVAL;
    const int32_t n = bm_num_vertices(ctx);
    if (n < 8) return BM_EINVAL; /* 4 swaps need 8 distinct vertices */
    bm_quick_stats st;
    memset(&st, 0, sizeof st);
    const size_t N1 = (size_t)n + 1;
    quick_scan q;
    memset(&q, 0, sizeof q);
    int32_t *best = malloc(sizeof(int32_t) * N1);
    q.rows = malloc(sizeof(int32_t) * 8 * N1);
    int rc = BM_OK;
    if (!best || !q.rows) {
        rc = BM_ENOMEM;
        goto out;
    }
    uint64_t rng = p->seed;
    int64_t cur, last = 0, iter = 0;
    if ((rc = bm_get_mapping(ctx, best)) || (rc = bm_score(ctx, &cur))) goto out;
    st.initial_score = st.best_score = cur;
    while (iter < p->max_evals) {
        /* :1139-1171 perturbation (max_delta == 0 here) */
        const int k = 3 + (int)(bm_rng_next(&rng) % 2);
        if (iter - last < p->best_window || (double)(st.best_score - cur) / p->gap_norm > p->gap) {
            if ((rc = bm_set_mapping(ctx, best))) goto out;
        }
        int32_t changed[8];
        if ((rc = bm_perturb(ctx, k, BM_PERTURB_WEIGHTED, &rng, changed))) goto out;
        sort_small(changed, 2 * k);
        if ((rc = bm_score(ctx, &cur))) goto out;
        st.perturbations++;
        if ((rc = quick_eval(ctx, n, changed, 2 * k, p->delta_kind, &q))) goto out;
        iter++;
        int zerocheck = 0;
        /* :1203-1318 */
        while ((q.max_delta > 0 || q.nzero > 0) && zerocheck < p->zero_limit && iter < p->max_evals) {
            int32_t m1, m2;
            if (q.max_delta > 0) {
                if (bm_rng_next(&rng) % 2 == 0) {
                    const int64_t pick = (int64_t)(bm_rng_next(&rng) % (uint64_t)q.npos);
                    quick_pick(&q, 1, pick, &m1, &m2);
                } else {
                    m1 = q.best_a;
                    m2 = q.best_b;
                }
                zerocheck = 0;
                st.positive_moves++;
            } else {
                const int64_t pick = (int64_t)(bm_rng_next(&rng) % (uint64_t)q.nzero);
                quick_pick(&q, 0, pick, &m1, &m2);
                zerocheck++;
                st.zero_moves++;
            }
            if ((rc = bm_apply_swap_deferred(ctx, m1, m2))) goto out;
            st.moves++;
            changed[0] = m1 < m2 ? m1 : m2;
            changed[1] = m1 < m2 ? m2 : m1;
            if ((rc = quick_eval(ctx, n, changed, 2, p->delta_kind, &q))) goto out;
            int64_t change; /* in by now: the rows' copy back waited for the swap */
            if ((rc = bm_deferred_change(ctx, &change))) goto out;
            cur += change; /* :1248, the rescore after the swap */
            if (cur > st.best_score) {
                st.best_score = cur;
                if ((rc = bm_get_mapping(ctx, best))) goto out;
                last = iter;
            }
            iter++;
        }
    }
    st.evals = iter;
    st.pairs = q.pairs;
    st.final_score = cur;
    rc = bm_set_mapping(ctx, best);
out:
    if (!rc && best_mapping) memcpy(best_mapping, best, sizeof(int32_t) * N1);
    if (!rc && stats) *stats = st;
    free(best);
    free(q.rows);
    return rc;
}

/* ------------------------------------------------------------------------
 * q3.cu changed-vertex triple search */

int64_t bm_q3_triples(int32_t n, const int32_t *chg, int32_t k, int32_t *out) {
    if (n < 3 || k < 1 || k > n || !chg) return BM_EINVAL;
    for (int32_t i = 0; i < k; i++)
        if (chg[i] < 1 || chg[i] > n || (i > 0 && chg[i] <= chg[i - 1])) return BM_EINVAL;
    const int64_t nr = (int64_t)n - k;
    const int64_t count = (int64_t)k * (k - 1) * (k - 2) / 6 + (int64_t)k * (k - 1) / 2 * nr + (nr > 0 ? k * (nr - 1) : 0);
    if (!out) return count;
    int32_t *rest = malloc(sizeof(int32_t) * (size_t)(nr > 0 ? nr : 1));
    if (!rest) return BM_ENOMEM;
    for (int32_t v = 1, i = 0, r = 0; v <= n; v++) { /* fillArray: the unchanged vertices ascending */
        if (i < k && chg[i] == v) i++;
        else rest[r++] = v;
    }
    int64_t c = 0;
    for (int32_t i = 0; i < k; i++) /* inside the changed set (:70-102) */
        for (int32_t j = i + 1; j < k; j++)
            for (int32_t l = j + 1; l < k; l++, c++) {
                out[3 * c] = chg[i];
                out[3 * c + 1] = chg[j];
                out[3 * c + 2] = chg[l];
            }
    for (int64_t r = 0; r < nr; r++) /* a changed pair and one unchanged vertex (:105-131) */
        for (int32_t i = 0; i < k; i++)
            for (int32_t j = i + 1; j < k; j++, c++) {
                out[3 * c] = chg[i];
                out[3 * c + 1] = chg[j];
                out[3 * c + 2] = rest[r];
            }
    for (int32_t i = 0; i < k; i++) /* a changed vertex and consecutive unchanged ones (:133-150) */
        for (int64_t t = 0; t + 1 < nr; t++, c++) {
            out[3 * c] = chg[i];
            out[3 * c + 1] = rest[t];
            out[3 * c + 2] = rest[t + 1];
        }
    free(rest);
    return c;
}

typedef struct {
    int64_t count, npos, nzero, best_idx, triples;
    int32_t best_delta;
    int32_t *tri, *d, *p;
    int64_t *pos, *zero;
} q3_scan;

/* score the triples of the changed set; the first maximum from -1 and the
 * positive and zero triples in index order (q3.cu:1264-1289) */
static int q3_eval(bm_ctx *ctx, int32_t n, const int32_t *chg, int32_t k, q3_scan *q) {
    q->count = bm_q3_triples(n, chg, k, q->tri);
    if (q->count < 0) return (int)q->count;
    int rc = bm_eval_rot3(ctx, q->tri, q->count, BM_ROT3_ALL, q->d, q->p);
    if (rc) return rc;
    q->best_delta = -1;
    q->best_idx = -1;
    q->npos = q->nzero = 0;
    for (int64_t i = 0; i < q->count; i++) {
        const int32_t d = q->d[i];
        if (d > q->best_delta) {
            q->best_delta = d;
            q->best_idx = i;
        }
        if (d > 0) q->pos[q->npos++] = i;
        else if (d == 0) q->zero[q->nzero++] = i;
    }
    q->triples += q->count;
    return BM_OK;
}

/* q3.cu:1176-1440 */
int bm_drive_q3(bm_ctx *ctx, const bm_q3_params *p, int32_t *best_mapping, bm_q3_stats *stats) {
    if (!ctx || !p || p->max_evals < 1 || p->zero_limit < 1 || p->best_window < 0 || !(p->gap_norm > 0))
        return BM_EINVAL;
    const int32_t n = bm_num_vertices(ctx);
    if (n < 8) return BM_EINVAL; /* 4 swaps need 8 distinct vertices */
    const int32_t eight[8] = {1, 2, 3, 4, 5, 6, 7, 8};
    const int64_t cap = bm_q3_triples(n, eight, 8, NULL); /* the most a perturbation can ask for */
    const size_t N1 = (size_t)n + 1;
    bm_q3_stats st;
    memset(&st, 0, sizeof st);
    q3_scan q;
    memset(&q, 0, sizeof q);
    int32_t *best = malloc(sizeof(int32_t) * N1);
    q.tri = malloc(sizeof(int32_t) * 3 * (size_t)cap);
    q.d = malloc(sizeof(int32_t) * (size_t)cap);
    q.p = malloc(sizeof(int32_t) * (size_t)cap);
    q.pos = malloc(sizeof(int64_t) * (size_t)cap);
    q.zero = malloc(sizeof(int64_t) * (size_t)cap);
    int rc = BM_OK;
    if (!best || !q.tri || !q.d || !q.p || !q.pos || !q.zero) {
        rc = BM_ENOMEM;
        goto out;
    }
    uint64_t rng = p->seed;
    int64_t cur, iter = 0;
    const int64_t last = 0; /* q3.cu never advances it */
    if ((rc = bm_get_mapping(ctx, best)) || (rc = bm_score(ctx, &cur))) goto out;
    st.initial_score = st.best_score = cur;
    while (iter < p->max_evals) {
        /* :1186-1211 perturbation (best_delta <= 0 whenever the move loop ended) */
        const int k = 3 + (int)(bm_rng_next(&rng) % 2);
        if (iter - last < p->best_window || (double)(st.best_score - cur) / p->gap_norm > p->gap) {
            if ((rc = bm_set_mapping(ctx, best))) goto out;
        }
        int32_t chg[8];
        if ((rc = bm_perturb(ctx, k, BM_PERTURB_WEIGHTED, &rng, chg))) goto out;
        sort_small(chg, 2 * k);
        if ((rc = bm_score(ctx, &cur))) goto out;
        st.perturbations++;
        if ((rc = q3_eval(ctx, n, chg, 2 * k, &q))) goto out;
        iter++;
        int zerocheck = 0;
        /* :1292-1434 */
        while ((q.best_delta > 0 || q.nzero > 0) && zerocheck < p->zero_limit && iter < p->max_evals) {
            int64_t t;
            if (q.best_delta > 0) {
                t = bm_rng_next(&rng) % 2 == 0 ? q.pos[bm_rng_next(&rng) % (uint64_t)q.npos] : q.best_idx;
                if ((rc = bm_apply_rot3(ctx, q.tri[3 * t], q.tri[3 * t + 1], q.tri[3 * t + 2], q.p[t]))) goto out;
                if ((rc = bm_score(ctx, &cur))) goto out; /* :1293 */
                zerocheck = 0;
                st.positive_moves++;
            } else {
                t = q.zero[bm_rng_next(&rng) % (uint64_t)q.nzero];
                if ((rc = bm_apply_rot3(ctx, q.tri[3 * t], q.tri[3 * t + 1], q.tri[3 * t + 2], q.p[t]))) goto out;
                zerocheck++;
                st.zero_moves++;
            }
            st.moves++;
            if (cur > st.best_score) {
                st.best_score = cur;
                if ((rc = bm_get_mapping(ctx, best))) goto out;
            }
            int32_t m3[3] = {q.tri[3 * t], q.tri[3 * t + 1], q.tri[3 * t + 2]};
            sort_small(m3, 3);
            if ((rc = q3_eval(ctx, n, m3, 3, &q))) goto out;
            iter++;
        }
    }
    st.evals = iter;
    st.triples = q.triples;
    st.final_score = cur;
    rc = bm_set_mapping(ctx, best);
out:
    if (!rc && best_mapping) memcpy(best_mapping, best, sizeof(int32_t) * N1);
    if (!rc && stats) *stats = st;
    free(best);
    free(q.tri);
    free(q.d);
    free(q.p);
    free(q.pos);
    free(q.zero);
    return rc;
}

/* lexicographic rank -> triple: with first element m1 there are C(n - m1, 2)
 * triples, then with second element m2 there are n - m2 */
int64_t bm_lex_triples(int32_t n, int64_t start, int64_t count, int32_t *out) {
    if (n < 3 || start < 0 || count < 0) return -1;
    const int64_t total = (int64_t)n * (n - 1) * (n - 2) / 6;
    if (start >= total || count == 0) return 0;
    if (count > total - start) count = total - start;
    if (!out) return count;
    int64_t t = start;
    int32_t m1 = 1, m2, m3;
    for (;;) {
        const int64_t r = (int64_t)(n - m1) * (n - m1 - 1) / 2;
        if (t < r) break;
        t -= r;
        m1++;
    }
    m2 = m1 + 1;
    while (t >= n - m2) {
        t -= n - m2;
        m2++;
    }
    m3 = m2 + 1 + (int32_t)t;
    for (int64_t i = 0; i < count; i++) {
        out[3 * i] = m1;
        out[3 * i + 1] = m2;
        out[3 * i + 2] = m3;
        if (++m3 > n) {
            if (++m2 > n - 1) {
                m1++;
                m2 = m1 + 1;
            }
            m3 = m2 + 1;
        }
    }
    return count;
}

/* triple.cu:941-1196 */
int bm_drive_triples(bm_ctx *ctx, const bm_triples_params *p, int32_t *best_mapping, bm_triples_stats *stats) {
    if (!ctx || !p || p->max_evals < 1 || p->batch < 1) return BM_EINVAL;
    const int32_t n = bm_num_vertices(ctx);
    if (n < 10) return BM_EINVAL; /* a perturbation of 5 swaps needs 10 distinct indices */
    const int64_t total = (int64_t)n * (n - 1) * (n - 2) / 6;
    const int64_t B = p->batch < total ? p->batch : total;
    const size_t N1 = (size_t)n + 1;
    bm_triples_stats st;
    memset(&st, 0, sizeof st);
    int32_t *best = malloc(sizeof(int32_t) * N1);
    int32_t *tri = malloc(sizeof(int32_t) * 3 * (size_t)B);
    int32_t *d = malloc(sizeof(int32_t) * (size_t)B);
    int32_t *pp = malloc(sizeof(int32_t) * (size_t)B);
    int rc = BM_OK;
    if (!best || !tri || !d || !pp) {
        rc = BM_ENOMEM;
        goto out;
    }
    uint64_t rng = p->seed;
    int64_t cur;
    if ((rc = bm_get_mapping(ctx, best)) || (rc = bm_score(ctx, &cur))) goto out;
    st.initial_score = st.best_score = cur;
    while (st.evals < p->max_evals) {
        int found = 0;
        int64_t start = 0;
        st.passes++;
        for (; start < total && st.evals < p->max_evals; start += B) {
            const int64_t m = bm_lex_triples(n, start, B, tri);
            if ((rc = bm_eval_rot3(ctx, tri, m, BM_ROT3_ALL, d, pp))) goto out;
            st.evals++;
            st.triples += m;
            int32_t bd = 0;
            int64_t bi = -1;
            for (int64_t i = 0; i < m; i++)
                if (d[i] > bd) {
                    bd = d[i];
                    bi = i;
                }
            if (bd > 0) { /* :1148-1168 */
                if ((rc = bm_apply_rot3(ctx, tri[3 * bi], tri[3 * bi + 1], tri[3 * bi + 2], pp[bi]))) goto out;
                cur += bd;
                st.moves++;
                if (cur > st.best_score) {
                    st.best_score = cur;
                    if ((rc = bm_get_mapping(ctx, best))) goto out;
                }
                found = 1;
                break;
            }
        }
        if (!found && start >= total) { /* :1174-1181 */
            if ((rc = bm_set_mapping(ctx, best))) goto out;
            const int k = 3 + (int)(bm_rng_next(&rng) % 3);
            if ((rc = bm_perturb(ctx, k, BM_PERTURB_OMP, &rng, NULL))) goto out;
            if ((rc = bm_score(ctx, &cur))) goto out;
            st.perturbations++;
        }
    }
    st.final_score = cur;
    rc = bm_set_mapping(ctx, best);
out:
    if (!rc && best_mapping) memcpy(best_mapping, best, sizeof(int32_t) * N1);
    if (!rc && stats) *stats = st;
    free(best);
    free(tri);
    free(d);
    free(pp);
    return rc;
}
