/* Engine calls the search drivers use to save a stream synchronisation per
 * move.  Not part of the C ABI: hidden in libbmatch.so. */
#ifndef BM_PRIVATE_H
#define BM_PRIVATE_H

#include "bmatch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bm_apply_swap_scored without the wait: the swap and its exact score change
 * are enqueued on the engine's stream; bm_deferred_change returns the change
 * once the stream has drained (it waits if needed).  At most one change is
 * pending per context: the next deferred swap overwrites it. */
__attribute__((visibility("hidden"))) int bm_apply_swap_deferred(bm_ctx *ctx, int32_t a, int32_t b);
__attribute__((visibility("hidden"))) int bm_deferred_change(bm_ctx *ctx, int64_t *change);

#ifdef __cplusplus
}
#endif

#endif
