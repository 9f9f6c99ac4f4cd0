// Device-resident candidate lists and their on-device arg-max.
//
// The reference GPU programs copy every per-move delta back and scan it on
// the host (triple.cu:1104-1126, quick.cu:1193-1217, 2 GB per step in
// graph-re-alignment-gpu.cu:856-868).  Here the list stays in HBM, the move
// kernels write their per-move results next to it and one reduction kernel
// folds them into a single 64-bit key: (delta biased) << 32 | ~index, so the
// largest delta wins and the lowest index breaks ties (strict > in the
// reference's scan order).
#include <string.h>

#include <algorithm>
#include <vector>

#include "bm_internal.h"
#include "bm_wave.h"

using namespace bm;

__global__ void k_fill_u64(unsigned long long *__restrict__ p, unsigned long long v) {
    if (threadIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(256) void k_argmax_moves(const int32_t *__restrict__ delta, int64_t count,
                                                      unsigned long long *__restrict__ key) {
    __shared__ unsigned long long red[4];
    unsigned long long best = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const int d = delta[i];
        if (d > 0) {
            const unsigned long long k =
                ((unsigned long long)((unsigned)d ^ 0x80000000u) << 32) | (0xFFFFFFFFu - (unsigned)i);
            best = k > best ? k : best;
        }
    }
    best = bm_wave::wave_max_u64(best);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long k = max(max(red[0], red[1]), max(red[2], red[3]));
        if (k) atomicMax(key, k);
    }
}

int bm::free_moves(bm_ctx *c) {
    const size_t cap = (size_t)c->mv_cap;
    if (c->mv_verts && c->stream) (void)hipStreamSynchronize(c->stream);  // queued work may still read them
    dev_free(c, (void **)&c->mv_verts, cap * 8 * 4);
    dev_free(c, (void **)&c->mv_delta, cap * 4);
    dev_free(c, (void **)&c->mv_aux, cap * 4);
    dev_free(c, (void **)&c->mv_flag, 64);
    c->mv_cap = c->mv_count = 0;
    c->mv_k = 0;
    return BM_OK;
}

static int ensure_moves(bm_ctx *c, int64_t count) {
    if (count <= c->mv_cap) return BM_OK;
    free_moves(c);
    const size_t cap = (size_t)std::max<int64_t>(count, 1024);
    int rc;
    if ((rc = dev_alloc(c, (void **)&c->mv_verts, cap * 8 * 4))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->mv_delta, cap * 4))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->mv_aux, cap * 4))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->mv_flag, 64))) return rc;
    c->mv_cap = (int64_t)cap;
    return BM_OK;
}

extern "C" int bm_moves_set(bm_ctx *c, int32_t k, const int32_t *verts, int64_t count) {
    if (!c || (count > 0 && !verts)) return set_error(BM_EINVAL, "NULL argument");
    if (!c->graphs) return set_error(BM_ESTATE, "graphs not uploaded");
    if (k < 2 || k > 8) return set_error(BM_EINVAL, "k=%d outside [2,8]", k);
    if (count < 0 || count >= (int64_t)1 << 31) return set_error(BM_ERANGE, "count %lld", (long long)count);
    for (int64_t i = 0; i < count; i++) {
        const int32_t *m = verts + (size_t)k * i;
        for (int j = 0; j < k; j++) {
            if (m[j] < (k == 2 ? 0 : 1) || m[j] > c->n)
                return set_error(BM_EINVAL, "move %lld: vertex %d outside the graph", (long long)i, m[j]);
            for (int q = 0; q < j; q++)
                if (m[q] == m[j]) return set_error(BM_EINVAL, "move %lld repeats vertex %d", (long long)i, m[j]);
        }
    }
    BM_HIP(hipSetDevice(c->device));
    int rc = ensure_moves(c, count);
    if (rc) return rc;
    if (count > 0) {
        if (k == 2) {  // pairs as two planes (a..., b...) for the pair kernel
            std::vector<int32_t> planes(2 * (size_t)count);
            for (int64_t i = 0; i < count; i++) {
                planes[i] = verts[2 * i];
                planes[count + i] = verts[2 * i + 1];
            }
            if (int rc_ = copy_sync(c->mv_verts, planes.data(), planes.size() * 4, hipMemcpyHostToDevice)) return rc_;
        } else {
            if (int rc_ = copy_sync(c->mv_verts, verts, (size_t)count * k * 4, hipMemcpyHostToDevice)) return rc_;
        }
    }
    c->mv_k = k;
    c->mv_count = count;
    return BM_OK;
}

// Scores the resident list; the best key lands in c->keys[N1 + 3].
static int eval_resident(bm_ctx *c, int mode) {
    const int64_t m = c->mv_count;
    int rc = BM_OK;
    int table = 0;
    if (c->mv_k == 3) rc = choose_rot3_table(c, c->mv_verts, m, c->mv_delta, c->mv_aux, &table);
    if (!rc && c->mv_k >= 4 && c->permset_threads_k != c->mv_k && m >= 256)
        rc = tune_permset(c, c->mv_k, c->mv_verts, m, c->mv_delta, c->mv_aux, c->mv_flag, 2);
    if (rc) return rc;
    if (c->mv_k == 2) rc = launch_eval_pairs(c, c->mv_verts, c->mv_verts + m, m, c->mv_delta);
    else if (c->mv_k == 3) rc = launch_rot3_moves(c, c->mv_verts, m, mode, c->mv_delta, c->mv_aux, table, true);
    else rc = launch_permset(c, c->mv_k, c->mv_verts, m, c->mv_delta, c->mv_aux, c->mv_flag);
    if (rc) return rc;
    unsigned long long *key = c->keys + c->N1 + 3;
    BM_LAUNCH(c, "fill", hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, key, 0ull));
    const unsigned blocks = (unsigned)std::min<int64_t>((m + 255) / 256, 2048);
    if (m > 0)
        BM_LAUNCH(c, "argmax_moves", hipLaunchKernelGGL(k_argmax_moves, dim3(blocks), dim3(256), 0, c->stream,
                                                        c->mv_delta, m, key));
    return BM_OK;
}

extern "C" int bm_moves_eval(bm_ctx *c, int mode, int64_t *best_index, int32_t *best_delta, int32_t *best_aux) {
    if (!c || !best_index || !best_delta || !best_aux) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (!c->mv_k) return set_error(BM_ESTATE, "no resident move list (bm_moves_set)");
    if (c->mv_k == 3 && mode != BM_ROT3_ROTATIONS && mode != BM_ROT3_ALL)
        return set_error(BM_EINVAL, "bad mode %d", mode);
    BM_HIP(hipSetDevice(c->device));
    if (c->mv_k >= 4) BM_HIP(hipMemsetAsync(c->mv_flag, 0, 4, c->stream));
    int rc = eval_resident(c, mode);
    if (rc) return rc;
    c->hres[1] = 0;
    BM_HIP(hipMemcpyAsync(c->hres, c->keys + c->N1 + 3, 8, hipMemcpyDeviceToHost, c->stream));
    if (c->mv_k >= 4) BM_HIP(hipMemcpyAsync(c->hres + 1, c->mv_flag, 4, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    const unsigned long long key = c->hres[0];
    const int flag = (int)(uint32_t)c->hres[1];
    if (flag) return set_error(BM_ERANGE, "a set has too many intra-set G1 lines");
    *best_index = -1;
    *best_delta = 0;
    *best_aux = 0;
    if (key) {
        const int64_t i = (int64_t)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu));
        *best_index = i;
        *best_delta = (int32_t)((uint32_t)(key >> 32) ^ 0x80000000u);
        if (c->mv_k >= 3) {
            if (int rc_ = copy_sync(best_aux, c->mv_aux + i, 4, hipMemcpyDeviceToHost)) return rc_;
        }
    }
    return BM_OK;
}

extern "C" int bm_moves_results(bm_ctx *c, int32_t *delta, int32_t *aux) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!c->mv_k) return set_error(BM_ESTATE, "no resident move list");
    BM_HIP(hipSetDevice(c->device));
    const size_t m = (size_t)c->mv_count;
    if (delta && m) {
        if (int rc_ = copy_sync(delta, c->mv_delta, m * 4, hipMemcpyDeviceToHost)) return rc_;
    }
    if (aux && m) {
        if (c->mv_k >= 3) {
            if (int rc_ = copy_sync(aux, c->mv_aux, m * 4, hipMemcpyDeviceToHost)) return rc_;
        } else {
            memset(aux, 0, m * 4);
        }
    }
    return BM_OK;
}
