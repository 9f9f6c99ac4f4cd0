/*
 * bm_io.c -- the reference CLIs' edge-CSV and matching-CSV formats.
 * Each reader/writer names the routine of zhihuidu/brain-match it mirrors.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bm_host.h"

#define LINE_MAX_REF 1024 /* MAX_LINE_LENGTH, graph-alignment-optimized.c:10 */

static int grow_edges(bm_edges *e) {
    int64_t cap = e->cap ? e->cap * 2 : 1024;
    int32_t *s = realloc(e->src, (size_t)cap * sizeof(int32_t));
    if (!s) return BM_ENOMEM;
    e->src = s;
    int32_t *d = realloc(e->dst, (size_t)cap * sizeof(int32_t));
    if (!d) return BM_ENOMEM;
    e->dst = d;
    int32_t *w = realloc(e->w, (size_t)cap * sizeof(int32_t));
    if (!w) return BM_ENOMEM;
    e->w = w;
    e->cap = cap;
    return BM_OK;
}

/* load_graph_from_csv, graph-alignment-optimized.c:635-679 (header skipped,
 * "%d,%d,%d" per line); the hill-climb (:405-453) and permset (:365-389)
 * loaders parse the same way and differ only in how they report a bad line. */
int bm_edges_read_csv(const char *path, int report, bm_edges *out) {
    memset(out, 0, sizeof *out);
    out->max_id = -1;
    FILE *f = fopen(path, "r");
    if (!f) return BM_EIO;
    char line[LINE_MAX_REF];
    long line_number = 1;
    if (!fgets(line, sizeof line, f)) {
        fclose(f);
        return BM_OK; /* empty file: no edges */
    }
    while (fgets(line, sizeof line, f)) {
        line_number++;
        int a, b, w;
        if (sscanf(line, "%d,%d,%d", &a, &b, &w) == 3) {
            if (out->m == out->cap && grow_edges(out)) {
                fclose(f);
                bm_edges_free(out);
                return BM_ENOMEM;
            }
            out->src[out->m] = a;
            out->dst[out->m] = b;
            out->w[out->m] = w;
            out->m++;
            if (a > out->max_id) out->max_id = a;
            if (b > out->max_id) out->max_id = b;
        } else {
            out->malformed++;
            if (report == BM_CSV_LOG_ERROR)
                fprintf(stderr, "[ERROR] load_graph_from_csv:667: Malformed line in CSV: %s\n", line);
            else if (report == BM_CSV_WARN_LINE)
                fprintf(stderr, "Warning: Invalid format at line %ld\n", line_number);
        }
    }
    fclose(f);
    return BM_OK;
}

int bm_edges_write_csv(const char *path, const char *header, const int32_t *src, const int32_t *dst,
                       const int32_t *w, int64_t m) {
    FILE *f = fopen(path, "w");
    if (!f) return BM_EIO;
    fprintf(f, "%s\n", header ? header : "Source,Target,Weight");
    for (int64_t i = 0; i < m; i++) fprintf(f, "%d,%d,%d\n", src[i], dst[i], w[i]);
    return fclose(f) == 0 ? BM_OK : BM_EIO;
}

void bm_edges_free(bm_edges *e) {
    free(e->src);
    free(e->dst);
    free(e->w);
    memset(e, 0, sizeof *e);
    e->max_id = -1;
}

static int push_pair(bm_pairs *p, int64_t *cap, int32_t a, int32_t b) {
    if (p->m == *cap) {
        int64_t nc = *cap ? *cap * 2 : 1024;
        int32_t *x = realloc(p->male, (size_t)nc * sizeof(int32_t));
        if (!x) return BM_ENOMEM;
        p->male = x;
        int32_t *y = realloc(p->female, (size_t)nc * sizeof(int32_t));
        if (!y) return BM_ENOMEM;
        p->female = y;
        *cap = nc;
    }
    p->male[p->m] = a;
    p->female[p->m] = b;
    p->m++;
    if (a > p->max_id) p->max_id = a;
    if (b > p->max_id) p->max_id = b;
    return BM_OK;
}

/* mode 0: load_benchmark_mapping (graph-alignment-optimized.c:452-469)
 * mode 1: load_mapping (graph-alignment-permset-parallel.c:455-462)
 * mode 2: load_initial_matching (refine-matching-hillclimb.c:471-488) */
static int read_pairs(const char *path, int mode, bm_pairs *out) {
    memset(out, 0, sizeof *out);
    out->max_id = -1;
    FILE *f = fopen(path, "r");
    if (!f) return BM_EIO;
    char line[LINE_MAX_REF];
    int64_t cap = 0;
    if (!fgets(line, sizeof line, f)) {
        fclose(f);
        return BM_OK;
    }
    while (fgets(line, sizeof line, f)) {
        int a, b, got = 0;
        char ms[LINE_MAX_REF], fs[LINE_MAX_REF];
        if (mode == 0) {
            if (sscanf(line, "%d,%d", &a, &b) == 2) {
                got = 1;
            } else if (sscanf(line, "%[^,],%s", ms, fs) == 2) {
                a = atoi(ms + (ms[0] == 'm' ? 1 : 0));
                b = atoi(fs + (fs[0] == 'f' ? 1 : 0));
                got = 1;
            }
        } else if (mode == 1) {
            if (sscanf(line, "%[^,],%s", ms, fs) == 2) {
                a = atoi(ms + 1);
                b = atoi(fs + 1);
                got = 1;
            }
        } else {
            if (sscanf(line, "%d,%d", &a, &b) == 2) got = 1;
        }
        if (got && push_pair(out, &cap, a, b)) {
            fclose(f);
            bm_pairs_free(out);
            return BM_ENOMEM;
        }
    }
    fclose(f);
    return BM_OK;
}

int bm_pairs_read_optimized(const char *path, bm_pairs *out) { return read_pairs(path, 0, out); }
int bm_pairs_read_permset(const char *path, bm_pairs *out) { return read_pairs(path, 1, out); }
int bm_pairs_read_plain(const char *path, bm_pairs *out) { return read_pairs(path, 2, out); }

void bm_pairs_free(bm_pairs *p) {
    free(p->male);
    free(p->female);
    memset(p, 0, sizeof *p);
    p->max_id = -1;
}

void bm_pairs_to_mapping(const bm_pairs *p, int32_t n, int32_t *mapping) {
    memset(mapping, 0, ((size_t)n + 1) * sizeof(int32_t));
    for (int64_t i = 0; i < p->m; i++) {
        int32_t a = p->male[i], b = p->female[i];
        if (a >= 0 && a <= n && b >= 0 && b <= n) mapping[a] = b;
    }
}

/* write_mapping, graph-alignment-optimized.c:410-425 */
int bm_mapping_write_mf(const char *path, const int32_t *mapping, int32_t n) {
    FILE *f = fopen(path, "w");
    if (!f) return BM_EIO;
    fprintf(f, "Male Node ID,Female Node ID\n");
    for (int32_t i = 1; i <= n; i++)
        if (mapping[i] != 0) fprintf(f, "m%d,f%d\n", i, mapping[i]);
    return fclose(f) == 0 ? BM_OK : BM_EIO;
}
