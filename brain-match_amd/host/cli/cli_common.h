/*
 * cli_common.h -- shared helpers of the drop-in CLIs: the reference's log
 * macros and number formatting (graph-alignment-optimized.c:17-38, 82-106),
 * engine setup, and GPU knobs read from the environment.
 */
#ifndef BM_CLI_COMMON_H
#define BM_CLI_COMMON_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bm_host.h"
#include "bmatch.h"

#define LOG_ERROR(fmt, ...) fprintf(stderr, "[ERROR] %s:%d: " fmt "\n", __func__, __LINE__, ##__VA_ARGS__)
#define LOG_INFO(fmt, ...) printf("[INFO] " fmt "\n", ##__VA_ARGS__)

/* format_number with thousands separators; rotating buffers so several
 * calls can appear in one printf */
static inline const char *fmt_num(int64_t num) {
    static char bufs[8][40];
    static int slot = 0;
    char *out = bufs[slot++ & 7];
    char tmp[32];
    int neg = num < 0;
    uint64_t u = neg ? (uint64_t)(-num) : (uint64_t)num;
    int len = snprintf(tmp, sizeof tmp, "%llu", (unsigned long long)u);
    int j = 0;
    if (neg) out[j++] = '-';
    for (int i = 0; i < len; i++) {
        if (i > 0 && (len - i) % 3 == 0) out[j++] = ',';
        out[j++] = tmp[i];
    }
    out[j] = '\0';
    return out;
}

static inline int64_t env_i64(const char *name, int64_t dflt) {
    const char *s = getenv(name);
    return (s && *s) ? strtoll(s, NULL, 10) : dflt;
}

static inline double wall_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* engine on BM_DEVICE (default 0) with both graphs uploaded */
static inline bm_ctx *cli_engine(int32_t n, const bm_edges *g1, const bm_edges *g2) {
    bm_ctx *ctx = NULL;
    int rc = bm_create((int)env_i64("BM_DEVICE", 0), &ctx);
    if (rc) {
        fprintf(stderr, "Error: GPU engine unavailable: %s (%s)\n", bm_strerror(rc), bm_last_error());
        return NULL;
    }
    rc = bm_upload_graphs(ctx, n, g1->src, g1->dst, g1->w, g1->m, g2->src, g2->dst, g2->w, g2->m);
    if (rc) {
        fprintf(stderr, "Error: graph upload failed: %s (%s)\n", bm_strerror(rc), bm_last_error());
        bm_destroy(ctx);
        return NULL;
    }
    return ctx;
}

#endif
