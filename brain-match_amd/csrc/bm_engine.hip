// Context, upload, mapping, score and move application of the MI355X
// move-delta engine.  See bm_internal.h for the device data layout.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <thread>
#include <functional>
#include <vector>

#include "bm_internal.h"
#include "../host/bm_private.h"
#include "bm_wave.h"

namespace bm {

static thread_local char g_err[512];

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? BM_ENOMEM : BM_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return BM_OK;
}

std::mutex &sync_mutex() {
    static std::mutex mu;
    return mu;
}

// Device memory is taken through a process-wide cache: a freed block goes
// back to its device's free list instead of to hipFree, and an allocation
// takes the smallest cached block of at least its size and at most a quarter
// more.  hipFree waits for the device, and under the serving loop every
// retired job freed some forty buffers while other jobs ran; the cache also
// spares the next job most of its hipMalloc calls.  Sizes are rounded to
// 256 B below 1 MiB and to 2 MiB above.  The caller frees a block only when
// its context's stream has no work left that uses it.  The free lists hold at
// most kCacheMax bytes per device (the largest blocks go back first), and a
// failed hipMalloc returns every cached block of the device and tries again.
namespace {
constexpr int kCacheDevices = 64;
constexpr size_t kCacheMax = (size_t)48 << 30;
struct BlockCache {
    std::mutex mu;
    std::map<void *, std::pair<int, size_t>> live;  // block -> (device, bytes)
    std::multimap<size_t, void *> idle[kCacheDevices];
    size_t idle_bytes[kCacheDevices] = {};
};
BlockCache &block_cache() {
    static BlockCache *bc = new BlockCache;  // never destroyed: frees may come during process exit
    return *bc;
}
size_t block_size(size_t bytes) {
    if (bytes < ((size_t)1 << 20)) return (bytes + 255) & ~(size_t)255;
    return (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
}
void hip_free_all(const std::vector<void *> &blocks) {
    if (blocks.empty()) return;
    std::lock_guard<std::mutex> g(sync_mutex());
    for (void *p : blocks) (void)hipFree(p);
}
}  // namespace

hipError_t block_alloc(int device, void **p, size_t bytes) {
    BlockCache &bc = block_cache();
    const size_t want = block_size(bytes ? bytes : 16);
    const int d = device >= 0 && device < kCacheDevices ? device : 0;
    {
        std::lock_guard<std::mutex> g(bc.mu);
        auto it = bc.idle[d].lower_bound(want);
        if (it != bc.idle[d].end() && it->first <= want + want / 4) {
            *p = it->second;
            bc.live[*p] = {d, it->first};
            bc.idle_bytes[d] -= it->first;
            bc.idle[d].erase(it);
            return hipSuccess;
        }
    }
    hipError_t e;
    {
        std::lock_guard<std::mutex> g(sync_mutex());
        e = hipMalloc(p, want);
    }
    if (e != hipSuccess) {  // give this device's cached blocks back and try once more
        (void)hipGetLastError();
        std::vector<void *> drop;
        {
            std::lock_guard<std::mutex> g(bc.mu);
            for (auto &kv : bc.idle[d]) drop.push_back(kv.second);
            bc.idle[d].clear();
            bc.idle_bytes[d] = 0;
        }
        if (drop.empty()) {
            *p = nullptr;
            return e;
        }
        hip_free_all(drop);
        {
            std::lock_guard<std::mutex> g(sync_mutex());
            e = hipMalloc(p, want);
        }
        if (e != hipSuccess) {
            *p = nullptr;
            return e;
        }
    }
    std::lock_guard<std::mutex> g(bc.mu);
    bc.live[*p] = {d, want};
    return hipSuccess;
}

void block_free(void *p) {
    if (!p) return;
    BlockCache &bc = block_cache();
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> g(bc.mu);
        auto it = bc.live.find(p);
        if (it == bc.live.end()) {
            drop.push_back(p);  // not from block_alloc
        } else {
            const int d = it->second.first;
            const size_t sz = it->second.second;
            bc.live.erase(it);
            bc.idle[d].emplace(sz, p);
            bc.idle_bytes[d] += sz;
            while (bc.idle_bytes[d] > kCacheMax && !bc.idle[d].empty()) {
                auto big = std::prev(bc.idle[d].end());
                bc.idle_bytes[d] -= big->first;
                drop.push_back(big->second);
                bc.idle[d].erase(big);
            }
        }
    }
    hip_free_all(drop);
}

int dev_alloc(bm_ctx *c, void **p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    const hipError_t e = block_alloc(c->device, p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    c->dev_bytes += (int64_t)bytes;
    return BM_OK;
}

int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    std::lock_guard<std::mutex> g(sync_mutex());
    BM_HIP(hipMemcpy(dst, src, bytes, kind));
    return BM_OK;
}

void dev_free(bm_ctx *c, void **p, size_t bytes) {
    if (*p) {
        block_free(*p);
        c->dev_bytes -= (int64_t)(bytes ? bytes : 16);
    }
    *p = nullptr;
}

int ensure_stage(bm_ctx *c, size_t bytes) {
    if (c->stage_bytes >= bytes) return BM_OK;
    if (c->stage) {
        BM_HIP(hipStreamSynchronize(c->stream));  // the stream's last use of the buffer is over
        block_free(c->stage);
        c->dev_bytes -= (int64_t)c->stage_bytes;
        c->stage = nullptr;
        c->stage_bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    int rc = dev_alloc(c, &c->stage, want);
    if (rc) return rc;
    c->stage_bytes = want;
    return BM_OK;
}

static int kernel_index(bm_ctx *c, const char *name) {
    for (size_t i = 0; i < c->kstats.size(); i++)
        if (c->kstats[i].name == name) return (int)i;
    c->kstats.push_back(KernelStat{name, 0, 0.0});
    return (int)c->kstats.size() - 1;
}

static hipEvent_t take_event(bm_ctx *c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int timer_begin(bm_ctx *c, const char *name) {
    if (!c->timing) return -1;
    TimerRec r;
    r.kernel = kernel_index(c, name);
    r.start = take_event(c);
    r.stop = take_event(c);
    if (!r.start || !r.stop) return -1;
    (void)hipEventRecord(r.start, c->stream);
    c->pending.push_back(r);
    return (int)c->pending.size() - 1;
}

void timer_end(bm_ctx *c, int token) {
    if (token < 0 || !c->timing) return;
    (void)hipEventRecord(c->pending[token].stop, c->stream);
}

static int resolve_timers(bm_ctx *c) {
    if (c->pending.empty()) return BM_OK;
    BM_HIP(hipStreamSynchronize(c->stream));
    for (auto &r : c->pending) {
        float ms = 0.f;
        BM_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
        c->kstats[r.kernel].launches++;
        c->kstats[r.kernel].ms += ms;
        c->ev_pool.push_back(r.start);
        c->ev_pool.push_back(r.stop);
    }
    c->pending.clear();
    return BM_OK;
}

}  // namespace bm

using namespace bm;

// ---------------------------------------------------------------- kernels

// W2[u][v] = W2T[v][u] = w for the de-duplicated G2 lines (last line per pair wins).
__global__ void k_scatter_w2(int16_t *__restrict__ W2, int N1, int ld, const int32_t *__restrict__ tri, int64_t m) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int u = tri[3 * i], v = tri[3 * i + 1], w = tri[3 * i + 2];
    W2[(size_t)u * ld + v] = (int16_t)w;
    W2[((size_t)N1 + v) * ld + u] = (int16_t)w;
}

// (neighbour, weight) per comb entry, from the two arrays
__global__ void k_interleave(const int32_t *__restrict__ a, const int32_t *__restrict__ b, int64_t m,
                             int2 *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = make_int2(a[i], b[i]);
}

// The flat out-line list in CSR order from the comb lists: out entry e of
// vertex v (cptr[v] <= e < cmid[v]) is line outptr[v] + e - cptr[v].
__global__ void k_flat_out(const int32_t *__restrict__ cptr, const int32_t *__restrict__ cmid,
                           const int32_t *__restrict__ outptr, const int32_t *__restrict__ cr,
                           const int32_t *__restrict__ cw, const int32_t *__restrict__ cdw, int N1, int64_t nc,
                           int32_t *__restrict__ es, int32_t *__restrict__ ed, int32_t *__restrict__ ew,
                           int32_t *__restrict__ edw) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc) return;
    int lo = 0, hi = N1 - 1;  // the vertex whose entries hold e: last v with cptr[v] <= e
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cptr[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    if (e >= cmid[lo]) return;  // an in entry
    const int64_t k = outptr[lo] + (e - cptr[lo]);
    es[k] = lo;
    ed[k] = cr[e] - N1;
    ew[k] = cw[e];
    edw[k] = cdw[e];
}

typedef int16_t bm_h8 __attribute__((ext_vector_type(8)));
// LDS limit of the row-gather kernel (160 KiB per CU on gfx950); one row of
// ld <= 65536 int16 always fits
constexpr size_t kGatherLdsMax = 160 * 1024;

// Q = [P; P^T] from [W2; W2^T]: P[x][y] = W2[p(x)][p(y)] and
// P^T[x][y] = W2^T[p(x)][p(y)], zero in the padding columns.  One block per
// x: the rows p(x) of W2 and W2^T are staged in LDS with 16-byte loads, then
// both output rows are gathered from LDS through one read of the mapping,
// 8 columns per thread, 16-byte stores.  (Gathering straight from global
// memory puts every lane on its own cache line.)
// Rows wider than fit twice in LDS take one half per block (grid y = 2,
// `both` = 0): the same code with r1 / d1 aliased to the single half.
// With `rows` (a device list of nrows[0] vertices) only those rows are
// rebuilt, grid-stride; `flag` (optional) skips the launch when *flag != want.
__global__ __launch_bounds__(512) void k_gather_rows(int16_t *__restrict__ Q, const int16_t *__restrict__ W2,
                                                     const int32_t *__restrict__ map, int N1, int ld, int both,
                                                     const int32_t *__restrict__ rows, const int *__restrict__ nrows,
                                                     const int *__restrict__ flag, int want) {
    extern __shared__ __align__(16) int16_t srow[];  // 2 * ld (both) or ld
    if (flag && *flag != want) return;
    const int count = rows ? *nrows : N1;
    for (int bi = blockIdx.x; bi < count; bi += gridDim.x) {
        const int x = rows ? rows[bi] : bi, tid = threadIdx.x, z = both ? 0 : blockIdx.y;
        const int g = map[x];
        const bm_h8 *s0 = (const bm_h8 *)(W2 + ((size_t)z * N1 + g) * ld);
        const bm_h8 *s1 = (const bm_h8 *)(W2 + ((size_t)N1 + g) * ld);
        bm_h8 *l0 = (bm_h8 *)srow, *l1 = (bm_h8 *)(srow + ld);
        for (int i = tid; i < ld / 8; i += 512) {
            l0[i] = s0[i];
            if (both) l1[i] = s1[i];
        }
        __syncthreads();
        const int16_t *r0 = srow, *r1 = both ? srow + ld : srow;
        int16_t *d0 = Q + ((size_t)z * N1 + x) * ld, *d1 = both ? Q + ((size_t)N1 + x) * ld : d0;
        for (int y0 = tid * 8; y0 < ld; y0 += 4096) {
            bm_h8 v0, v1;
            if (y0 + 8 <= N1) {
                const int4 m0 = *(const int4 *)(map + y0), m1 = *(const int4 *)(map + y0 + 4);
                const int m[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    v0[j] = r0[m[j]];
                    v1[j] = r1[m[j]];
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const bool in = y0 + j < N1;
                    const int m = in ? map[y0 + j] : 0;
                    v0[j] = in ? r0[m] : (int16_t)0;
                    v1[j] = in ? r1[m] : (int16_t)0;
                }
            }
            __builtin_nontemporal_store(v0, (bm_h8 *)(d0 + y0));
            if (both) __builtin_nontemporal_store(v1, (bm_h8 *)(d1 + y0));
        }
        __syncthreads();  // srow is reused by the next row
    }
}

// The columns of a vertex list in every row of P and P^T:
// P[x][v] = W2[p(x)][p(v)], P^T[x][v] = W2^T[p(x)][p(v)]; one wave per
// output row, lanes over the list, which each workgroup first stages in LDS
// with the images p(v) (lists of up to kPatchStage vertices; longer ones are
// read from memory).  Skipped when *flag != want.
constexpr int kPatchStage = 1024;

__global__ __launch_bounds__(256) void k_patch_cols(int16_t *__restrict__ Q, const int16_t *__restrict__ W2,
                                                    const int32_t *__restrict__ map, int N1, int ld,
                                                    const int32_t *__restrict__ cols, const int *__restrict__ ncols,
                                                    const int *__restrict__ flag, int want) {
    __shared__ int lv[kPatchStage], lg[kPatchStage];
    if (*flag != want) return;
    const int n = *ncols, lane = threadIdx.x & 63;
    const bool staged = n <= kPatchStage;
    if (staged)
        for (int i = threadIdx.x; i < n; i += 256) {
            const int v = cols[i];
            lv[i] = v;
            lg[i] = map[v];
        }
    __syncthreads();
    const int nwaves = gridDim.x * 4;
    for (int r = blockIdx.x * 4 + (threadIdx.x >> 6); r < 2 * N1; r += nwaves) {
        const int z = r >= N1, x = r - z * N1;
        const int16_t *src = W2 + ((size_t)z * N1 + map[x]) * ld;
        int16_t *dst = Q + (size_t)r * ld;
        for (int i0 = lane; i0 < n; i0 += 256) {  // four cells in flight per lane
            int v[4];
            int16_t w[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int i = i0 + 64 * k;
                v[k] = i < n ? (staged ? lv[i] : cols[i]) : -1;
                w[k] = v[k] >= 0 ? src[staged ? lg[i] : map[v[k]]] : (int16_t)0;
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (v[k] >= 0) dst[v[k]] = w[k];
        }
    }
}

// diag[v] = sum_{e in comb(v)} min(w_e, Q[r_e][v]); one wave per vertex.
__global__ __launch_bounds__(256) void k_diag(int32_t *__restrict__ diag, const int32_t *__restrict__ cptr,
                                              const int32_t *__restrict__ cr, const int32_t *__restrict__ cw,
                                              const int16_t *__restrict__ Q, int N1, int ld) {
    int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (v >= N1) return;
    int s = 0;
    for (int e = cptr[v] + lane; e < cptr[v + 1]; e += 64) s += min(cw[e], (int)Q[(size_t)cr[e] * ld + v]);
    s = bm_wave::wave_sum(s);
    if (lane == 0) diag[v] = s;
}

// Exact S_opt change of swapping the images of a and b, before the swap: the
// G1 lines with an end at a or b and a source other than vertex 0, old term min(w, P[x][y]) from the current
// P, new term min(w, W2[p'(x)][p'(y)]) at the swapped images.  Each line
// once: the out entries of a and b, then their in entries whose source is
// neither a nor b.  One workgroup; the total goes to *out.
__global__ __launch_bounds__(256) void k_swap_change(long long *__restrict__ out, int a, int b,
                                                     const int32_t *__restrict__ cptr, const int2 *__restrict__ crw,
                                                     const int32_t *__restrict__ map, const int16_t *__restrict__ Q,
                                                     const int16_t *__restrict__ W2, int N1, int ld) {
    __shared__ long long part[4];
    const int pa = map[a], pb = map[b];
    long long s = 0;
    for (int side = 0; side < 2; side++) {
        const int v = side ? b : a;
        for (int e = cptr[v] + threadIdx.x; e < cptr[v + 1]; e += 256) {
            const int2 rw = crw[e];
            const bool out_entry = rw.x >= N1;  // v -> rw.x - N1, else rw.x -> v
            const int x = out_entry ? v : rw.x, y = out_entry ? rw.x - N1 : v;
            if (!out_entry && (x == a || x == b)) continue;  // counted as an out entry of x
            if (x == 0) continue;                              // lines leaving vertex 0 are not scored
            const int px = x == a ? pb : x == b ? pa : map[x];
            const int py = y == a ? pb : y == b ? pa : map[y];
            s += min(rw.y, (int)W2[(size_t)px * ld + py]) - min(rw.y, (int)Q[(size_t)x * ld + y]);
        }
    }
    s = bm_wave::wave_sum64(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = part[0] + part[1] + part[2] + part[3];
}

// S_opt = sum over G1 lines (x->y, w), x >= 1, of min(w, P[x][y]): the flat
// list is in source order, so the lines of vertex 0 are its first `lo`
// entries, which calculate_alignment_score (graph-alignment-optimized.c:295)
// never visits.
// With pmap the cells are read as W2[p(x)][p(y)] (P = W2), else as P[x][y].
__global__ __launch_bounds__(256) void k_score(unsigned long long *__restrict__ out, const int32_t *__restrict__ es,
                                               const int32_t *__restrict__ ed, const int32_t *__restrict__ ew,
                                               int64_t lo, int64_t m, const int16_t *__restrict__ P, int ld,
                                               const int32_t *__restrict__ pmap) {
    __shared__ long long part[4];
    long long s = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // four lines per thread per pass: their list loads, then their four cell gathers, in flight together
    for (int64_t i0 = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < m; i0 += 4 * stride) {
        int x[4], y[4], w[4], q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t i = i0 + k * stride;
            const bool in = i < m;
            x[k] = in ? es[i] : 0;
            y[k] = in ? ed[i] : 0;
            w[k] = in ? ew[i] : 0;
        }
        if (pmap) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                x[k] = pmap[x[k]];
                y[k] = pmap[y[k]];
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) q[k] = i0 + k * stride < m ? (int)P[(size_t)x[k] * ld + y[k]] : 0;
#pragma unroll
        for (int k = 0; k < 4; k++) s += i0 + k * stride < m ? min(w[k], q[k]) : 0;
    }
    s = bm_wave::wave_sum64(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = part[0] + part[1] + part[2] + part[3];
        atomicAdd(out, (unsigned long long)t);
    }
}

// X(v,v) after a move on a vertex set S (new p(S[i]) = old p(S[src[i]])).
// For v outside S only the entries of comb(v) that are G1 lines to S change,
// so they are patched from comb(S), before the permutation, with old and
// new cells both read from the old P (the new row of S[i] is the old row of
// S[src[i]]):
//   line s -> v:  diag[v] += min(w, P[S[src]][v]) - min(w, P[s][v])
//   line v -> s:  diag[v] += min(w, P[v][S[src]]) - min(w, P[v][s])
// The vertices of S are recomputed over their whole lists afterwards.
struct TouchSet {
    int k;
    int v[8];
    int src[8];  // the old row / column that S[i] takes
    int inv[8];  // the position that S[i]'s old row / column moves to
    int off[9];  // prefix of the comb-list lengths over the set
};

__device__ __forceinline__ int diag_of(int v, int lane, const int32_t *__restrict__ cptr,
                                       const int32_t *__restrict__ cr, const int32_t *__restrict__ cw,
                                       const int16_t *__restrict__ Q, int ld) {
    int s = 0;
    for (int e = cptr[v] + lane; e < cptr[v + 1]; e += 64) s += min(cw[e], (int)Q[(size_t)cr[e] * ld + v]);
    s = bm_wave::wave_sum(s);
    return s;
}

// X(v,v) after a move, from the moved P and P^T, in one launch.
//   blocks [0, nb): a thread per line between a set vertex s and a vertex v
//     outside the set: X(v,v) changes by min(w, cell now) - min(w, cell before),
//     the cell before being the one that now sits at s's new position inv;
//   block nb: one wave per set vertex over its whole list.
// The two parts write X(v,v) of disjoint vertices and read only P and P^T.
__global__ __launch_bounds__(512) void k_diag_move(int32_t *__restrict__ diag, const int32_t *__restrict__ cptr,
                                                   const int32_t *__restrict__ cr, const int32_t *__restrict__ cw,
                                                   const int16_t *__restrict__ Q, int N1, int ld, TouchSet ts, int nb) {
    if ((int)blockIdx.x == nb) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (w >= ts.k) return;
        const int v = ts.v[w];
        const int d = diag_of(v, lane, cptr, cr, cw, Q, ld);
        if (lane == 0) diag[v] = d;
        return;
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ts.off[ts.k]) return;
    int i = 0;
    while (t >= ts.off[i + 1]) i++;
    if (ts.src[i] == i) return;  // s keeps its row and column
    const int s = ts.v[i], s0 = ts.v[ts.inv[i]];
    const int e = cptr[s] + (t - ts.off[i]);
    const int r = cr[e], w = cw[e];
    const bool out = r >= N1;
    const int v = out ? r - N1 : r;
    for (int q = 0; q < ts.k; q++)
        if (ts.v[q] == v) return;  // a line inside S: v is recomputed in full
    const int o = out ? (int)Q[(size_t)s0 * ld + v] : (int)Q[(size_t)v * ld + s0];
    const int nw = out ? (int)Q[(size_t)s * ld + v] : (int)Q[(size_t)v * ld + s];
    const int d = min(w, nw) - min(w, o);
    if (d) atomicAdd(&diag[v], d);
}

// The translated G2 lists (target as p^-1(target), 0xFFFF when unmapped)
// after a move: only entries whose target is one of the moved images change;
// they are the G2 lines at those images, found through the cross positions.
struct ImageSet {
    int k;
    int g[8];    // the moved images
    int y[8];    // their preimages after the move
    int off[9];  // prefix of (in-degree + out-degree) over the images
};

struct SetPerm {
    int k;
    int set[8];
    int src[8];  // new row/col set[i] <- old row/col set[src[i]]
};

// One move on P and P^T (rows and columns of the set permuted), the mapping
// and the translated G2 lists, in one launch.  Thread ranges:
//   [0, ld)              column j outside the set: the set's rows at j, both halves
//   [ld, ld + 2(n+1))    row x outside the set: the set's columns in row x
//   next npatch          the translated entries of the G2 lines at the moved images
//                        (target -> its preimage after the move, is.y)
//   the last one         the set x set block of both halves, the mapping and p^-1
// The four parts write disjoint elements and read only elements no other part
// writes, so they need no ordering among themselves.
__global__ __launch_bounds__(256) void k_apply_move(int16_t *__restrict__ Q, int N1, int ld, SetPerm sp,
                                                    int32_t *__restrict__ map, int32_t *__restrict__ pinv,
                                                    int with_pinv, ImageSet is, int npatch,
                                                    uint32_t *__restrict__ g2oy, uint32_t *__restrict__ g2iy,
                                                    const uint32_t *__restrict__ g2o, const uint32_t *__restrict__ g2i,
                                                    const int32_t *__restrict__ g2optr,
                                                    const int32_t *__restrict__ g2iptr,
                                                    const int32_t *__restrict__ g2o_at_in,
                                                    const int32_t *__restrict__ g2i_at_out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = sp.k;
    auto in_set = [&](int v) {
        bool hit = false;
        for (int i = 0; i < k; i++) hit |= sp.set[i] == v;
        return hit;
    };
    if (t < ld) {  // rows part
        const int j = (int)t;
        if (in_set(j)) return;
        for (int half = 0; half < 2; half++) {
            int16_t *M = Q + (size_t)half * N1 * ld;
            int16_t old[8];
            for (int i = 0; i < k; i++) old[i] = M[(size_t)sp.set[sp.src[i]] * ld + j];
            for (int i = 0; i < k; i++) M[(size_t)sp.set[i] * ld + j] = old[i];
        }
        return;
    }
    int64_t u = t - ld;
    if (u < 2 * (int64_t)N1) {  // columns part
        const int x = (int)u;
        if (in_set(x >= N1 ? x - N1 : x)) return;
        int16_t *row = Q + (size_t)x * ld;
        int16_t old[8];
        for (int i = 0; i < k; i++) old[i] = row[sp.set[sp.src[i]]];
        for (int i = 0; i < k; i++) row[sp.set[i]] = old[i];
        return;
    }
    u -= 2 * (int64_t)N1;
    if (u < npatch) {  // translated G2 entries at the moved images
        const int l0 = (int)u;
        int i = 0;
        while (l0 >= is.off[i + 1]) i++;
        const int g = is.g[i], l = l0 - is.off[i];
        const uint32_t hi = (uint32_t)is.y[i] << 16;
        const int nin = g2iptr[g + 1] - g2iptr[g];
        if (l < nin) {  // line s -> g: entry of s's out-list
            const int j = g2o_at_in[g2iptr[g] + l];
            g2oy[j] = hi | (g2o[j] & 0xFFFFu);
        } else {        // line g -> v: entry of v's in-list
            const int j = g2i_at_out[g2optr[g] + l - nin];
            g2iy[j] = hi | (g2i[j] & 0xFFFFu);
        }
        return;
    }
    if (u == npatch) {  // the set x set block, the mapping
        for (int half = 0; half < 2; half++) {
            int16_t *M = Q + (size_t)half * N1 * ld;
            int16_t old[8][8];
            for (int i = 0; i < k; i++)
                for (int j = 0; j < k; j++) old[i][j] = M[(size_t)sp.set[sp.src[i]] * ld + sp.set[sp.src[j]]];
            for (int i = 0; i < k; i++)
                for (int j = 0; j < k; j++) M[(size_t)sp.set[i] * ld + sp.set[j]] = old[i][j];
        }
        int img[8];
        for (int i = 0; i < k; i++) img[i] = map[sp.set[sp.src[i]]];
        for (int i = 0; i < k; i++) map[sp.set[i]] = img[i];
        // the images are permuted among the set, so p stays injective
        if (with_pinv)
            for (int i = 0; i < k; i++)
                if (img[i] != 0) pinv[img[i]] = sp.set[i];
    }
}

// ---------------------------------------------------------------- helpers

namespace bm {

int launch_rebuild_p(bm_ctx *c, const int *patch_flag, const int32_t *verts, const int *nverts) {
    // both source rows in LDS when they fit (BM_GATHER_SPLIT=1 forces one row per block)
    const int both = (size_t)c->ld * 4 <= kGatherLdsMax && !getenv("BM_GATHER_SPLIT") ? 1 : 0;
    const size_t lds = (size_t)c->ld * 2 * (both ? 2 : 1);
    if (!c->gather_attr) {  // once per context, outside any capture
        BM_HIP(hipFuncSetAttribute((const void *)k_gather_rows, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kGatherLdsMax));
        c->gather_attr = true;
    }
    // whole matrices (unless *patch_flag says patch); grid-stride over the rows, so that a
    // launch that the flag skips has few workgroups to retire
    BM_LAUNCH(c, "build_p", hipLaunchKernelGGL(k_gather_rows, dim3(patch_flag ? std::min(c->N1, 2048) : c->N1,
                                                                   both ? 1 : 2), dim3(512), lds, c->stream,
                                               c->Q, c->W2, c->map, c->N1, c->ld, both, nullptr, nullptr,
                                               patch_flag, 0));
    if (patch_flag) {  // or the rows and columns of the moved vertices only
        BM_LAUNCH(c, "patch_p", hipLaunchKernelGGL(k_gather_rows, dim3(1024, both ? 1 : 2), dim3(512), lds, c->stream,
                                                   c->Q, c->W2, c->map, c->N1, c->ld, both, verts, nverts,
                                                   patch_flag, 1));
        BM_LAUNCH(c, "patch_p", hipLaunchKernelGGL(k_patch_cols, dim3(2048), dim3(256), 0, c->stream, c->Q, c->W2,
                                                   c->map, c->N1, c->ld, verts, nverts, patch_flag, 1));
    }
    return BM_OK;
}

int launch_rebuild_p_if(bm_ctx *c, const int *flag) {
    const int both = (size_t)c->ld * 4 <= kGatherLdsMax && !getenv("BM_GATHER_SPLIT") ? 1 : 0;
    const size_t lds = (size_t)c->ld * 2 * (both ? 2 : 1);
    if (!c->gather_attr) return set_error(BM_ESTATE, "P rebuild before the first mapping");
    BM_LAUNCH(c, "build_p", hipLaunchKernelGGL(k_gather_rows, dim3(std::min(c->N1, 2048), both ? 1 : 2), dim3(512),
                                               lds, c->stream, c->Q, c->W2, c->map, c->N1, c->ld, both, nullptr,
                                               nullptr, flag, 1));
    return BM_OK;
}

int ensure_p(bm_ctx *c) {
    if (!c->p_stale) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    int rc = launch_rebuild_p(c);
    if (rc) return rc;
    c->p_stale = false;
    return BM_OK;
}

int launch_diag(bm_ctx *c) {
    BM_LAUNCH(c, "diag", hipLaunchKernelGGL(k_diag, dim3((c->N1 + 3) / 4), dim3(256), 0, c->stream, c->diag, c->cptr,
                                            c->cr, c->cw, c->Q, c->N1, c->ld));
    return BM_OK;
}

// S_opt into *out (which must hold 0 beforehand); no host synchronisation
int launch_score(bm_ctx *c, unsigned long long *out, bool via_map) {
    int64_t m = c->e1;
    int blocks = (int)std::min<int64_t>((m - c->e1_v0 + 255) / 256, 4096);
    if (blocks < 1) blocks = 1;
    BM_LAUNCH(c, "score", hipLaunchKernelGGL(k_score, dim3(blocks), dim3(256), 0, c->stream, out, c->esrc, c->edst,
                                             c->ew, c->e1_v0, m, via_map ? c->W2 : c->Q, c->ld, via_map ? c->map : nullptr));
    return BM_OK;
}

int launch_list_diag(bm_ctx *c, int32_t *out, const int32_t *ptr, const int32_t *r, const int32_t *w) {
    BM_LAUNCH(c, "list_diag", hipLaunchKernelGGL(k_diag, dim3((c->N1 + 3) / 4), dim3(256), 0, c->stream, out, ptr, r,
                                                 w, c->Q, c->N1, c->ld));
    return BM_OK;
}

int ensure_diag(bm_ctx *c) {
    if (c->diag_valid) return BM_OK;
    int blocks = (c->N1 + 3) / 4;
    BM_LAUNCH(c, "diag", hipLaunchKernelGGL(k_diag, dim3(blocks), dim3(256), 0, c->stream, c->diag, c->cptr, c->cr,
                                            c->cw, c->Q, c->N1, c->ld));
    c->diag_valid = true;
    return BM_OK;
}

static void touch_set(const std::vector<int32_t> &deg, int k, const int32_t *set, const int32_t *src_of,
                      TouchSet &ts) {
    memset(&ts, 0, sizeof ts);
    ts.k = k;
    for (int i = 0; i < k; i++) {
        ts.v[i] = set[i];
        ts.src[i] = src_of[i];
        ts.inv[src_of[i]] = i;
        ts.off[i + 1] = ts.off[i] + deg[set[i]];
    }
}

int permute_rows_cols(bm_ctx *c, int k, const int32_t *set, const int32_t *src_of) {
    c->map_epoch++;
    const bool diag = c->diag_valid;      // keep X(v,v) current where it changes
    const bool dfdiag = c->dfdiag_valid;  // the same for the dense-form lists
    SetPerm sp;
    memset(&sp, 0, sizeof sp);
    sp.k = k;
    for (int i = 0; i < k; i++) {
        sp.set[i] = set[i];
        sp.src[i] = src_of[i];
    }
    int32_t old[8];
    for (int i = 0; i < k; i++) old[i] = c->hmap[set[src_of[i]]];
    for (int i = 0; i < k; i++) c->hmap[set[i]] = old[i];
    if (c->sparse_ok)
        for (int i = 0; i < k; i++)
            if (old[i] != 0) c->hpinv[old[i]] = set[i];
    // keep the translated G2 lists current: the lines at the moved images
    ImageSet is;
    memset(&is, 0, sizeof is);
    const bool patch = c->sparse_ok && !c->g2y_dirty;
    if (patch) {
        for (int i = 0; i < k; i++) {
            const int g = old[i];
            if (g == 0) continue;
            is.g[is.k] = g;
            is.y[is.k] = set[i];
            is.off[is.k + 1] = is.off[is.k] + (c->hg2iptr[g + 1] - c->hg2iptr[g]) + (c->hg2optr[g + 1] - c->hg2optr[g]);
            is.k++;
        }
    } else {
        c->g2y_dirty = true;
    }
    const int npatch = is.off[is.k];
    const int64_t threads = (int64_t)c->ld + 2 * (int64_t)c->N1 + npatch + 1;
    BM_LAUNCH(c, "apply_move",
              hipLaunchKernelGGL(k_apply_move, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream, c->Q,
                                 c->N1, c->ld, sp, c->map, c->pinv, c->sparse_ok ? 1 : 0, is, npatch, c->g2oy, c->g2iy,
                                 c->g2o, c->g2i, c->g2optr, c->g2iptr, c->g2o_at_in, c->g2i_at_out));
    TouchSet ts;
    if (diag) {
        touch_set(c->hcdeg, k, set, src_of, ts);
        const int nb = (ts.off[k] + 511) / 512;
        BM_LAUNCH(c, "diag_move", hipLaunchKernelGGL(k_diag_move, dim3(nb + 1), dim3(512), 0, c->stream, c->diag,
                                                     c->cptr, c->cr, c->cw, c->Q, c->N1, c->ld, ts, nb));
    }
    if (dfdiag) {
        touch_set(c->hdfdeg, k, set, src_of, ts);
        const int nb = (ts.off[k] + 511) / 512;
        BM_LAUNCH(c, "diag_move", hipLaunchKernelGGL(k_diag_move, dim3(nb + 1), dim3(512), 0, c->stream, c->dfdiag,
                                                     c->dfptr, c->dfr, c->dfw, c->Q, c->N1, c->ld, ts, nb));
    }
    return BM_OK;
}

}  // namespace bm

template <typename T>
static int upload(bm_ctx *c, T **dst, const std::vector<T> &src) {
    int rc = dev_alloc(c, (void **)dst, src.size() * sizeof(T));
    if (rc) return rc;
    if (!src.empty()) return copy_sync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    return BM_OK;
}

// ---------------------------------------------------------------- C ABI

extern "C" {

const char *bm_strerror(int code) {
    switch (code) {
        case BM_OK: return "ok";
        case BM_EINVAL: return "invalid argument";
        case BM_ENOMEM: return "out of memory";
        case BM_EHIP: return "HIP runtime error";
        case BM_ESTATE: return "graphs or mapping not uploaded";
        case BM_ENODEV: return "no such HIP device";
        case BM_ERANGE: return "value out of range";
        case BM_EIO: return "I/O error";
        default: return "unknown error";
    }
}

const char *bm_last_error(void) { return bm::g_err; }

int bm_abi_version(void) { return BM_ABI_VERSION; }

int bm_create(int device, bm_ctx **out) {
    if (!out) return set_error(BM_EINVAL, "bm_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return set_error(BM_ENODEV, "no HIP device (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return set_error(BM_ENODEV, "device %d not in [0,%d)", device, ndev);
    BM_HIP(hipSetDevice(device));
    bm_ctx *c = new (std::nothrow) bm_ctx();
    if (!c) return set_error(BM_ENOMEM, "bm_create: host allocation");
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    {
        std::lock_guard<std::mutex> g(sync_mutex());
        e = hipHostMalloc((void **)&c->hres, 8 * sizeof(unsigned long long), hipHostMallocDefault);
    }
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "hipHostMalloc");
    }
    *out = c;
    return BM_OK;
}

static void free_graph_buffers(bm_ctx *c) {
    size_t N1 = (size_t)c->N1, ld = (size_t)c->ld;
    c->hvscore.clear();
    dev_free(c, (void **)&c->W2, 2 * N1 * ld * 2);
    dev_free(c, (void **)&c->Q, 2 * N1 * ld * 2);
    dev_free(c, (void **)&c->map, N1 * 4);
    dev_free(c, (void **)&c->cptr, (N1 + 1) * 4);
    dev_free(c, (void **)&c->cr, (size_t)c->ncomb * 4);
    dev_free(c, (void **)&c->cw, (size_t)c->ncomb * 4);
    dev_free(c, (void **)&c->cdw, (size_t)c->ncomb * 4);
    dev_free(c, (void **)&c->crw, (size_t)c->ncomb * 8);
    dev_free(c, (void **)&c->cmid, N1 * 4);
    dev_free(c, (void **)&c->cneg, N1 * 4);
    dev_free(c, (void **)&c->rowpart, c->rowpart_bytes);
    c->rowpart_bytes = 0;
    dev_free(c, (void **)&c->grpres, (size_t)BM_ROW_GROUP_MAX * 8);
    dev_free(c, (void **)&c->ddptr, (N1 + 1) * 4);
    dev_free(c, (void **)&c->ddmid, N1 * 4);
    dev_free(c, (void **)&c->ddr, (size_t)c->ndd * 4);
    dev_free(c, (void **)&c->ddw, (size_t)c->ndd * 4);
    dev_free(c, (void **)&c->dfptr, (N1 + 1) * 4);
    dev_free(c, (void **)&c->dfmid, N1 * 4);
    dev_free(c, (void **)&c->dfr, (size_t)c->ndf * 4);
    dev_free(c, (void **)&c->dfw, (size_t)c->ndf * 4);
    dev_free(c, (void **)&c->dfrw, (size_t)c->ndf * 8);
    dev_free(c, (void **)&c->dfneg, N1 * 4);
    dev_free(c, (void **)&c->dfdiag, N1 * 4);
    dev_free(c, (void **)&c->dfes, (size_t)c->ndfout * 4);
    dev_free(c, (void **)&c->dfed, (size_t)c->ndfout * 4);
    dev_free(c, (void **)&c->dfew, (size_t)c->ndfout * 4);
    c->ndfout = 0;
    dev_free(c, (void **)&c->dfX, (N1 + 63) / 64 * 64 * ld * 4);
    dev_free(c, (void **)&c->dfW1, N1 * ld * 2);
    dev_free(c, (void **)&c->dfXP, N1 * ld * 8);
    c->ndf = 0;
    c->dfmaxw = 0;
    c->dfdiag_valid = false;
    c->hdfdeg.clear();
    c->hcptr.clear(), c->hcmid.clear(), c->hcr.clear(), c->hcw.clear(), c->hddptr.clear(), c->hddmid.clear();
    c->hlast.clear();
    dev_free(c, (void **)&c->esrc, (size_t)c->e1 * 4);
    dev_free(c, (void **)&c->edst, (size_t)c->e1 * 4);
    dev_free(c, (void **)&c->ew, (size_t)c->e1 * 4);
    dev_free(c, (void **)&c->edw, (size_t)c->e1 * 4);
    dev_free(c, (void **)&c->diag, N1 * 4);
    dev_free(c, (void **)&c->g2optr, (N1 + 1) * 4);
    dev_free(c, (void **)&c->g2iptr, (N1 + 1) * 4);
    dev_free(c, (void **)&c->g2o, (size_t)c->ng2 * 4);
    dev_free(c, (void **)&c->g2i, (size_t)c->ng2 * 4);
    dev_free(c, (void **)&c->g2o_at_in, (size_t)c->ng2 * 4);
    dev_free(c, (void **)&c->g2i_at_out, (size_t)c->ng2 * 4);
    dev_free(c, (void **)&c->pinv, N1 * 4);
    dev_free(c, (void **)&c->g2oy, (size_t)c->ng2 * 8);  // g2iy lives in the same allocation
    c->g2iy = nullptr;
    dev_free(c, (void **)&c->g2q, (size_t)c->nquads * 16);
    dev_free(c, (void **)&c->qoptr, (N1 + 1) * 4);
    dev_free(c, (void **)&c->qiptr, (N1 + 1) * 4);
    dev_free(c, (void **)&c->g2qpos, (size_t)c->ng2 * 8);
    c->nquads = 0;
    dev_free(c, (void **)&c->crws, (size_t)c->ncomb * 8);
    dev_free(c, (void **)&c->cdws, (size_t)c->ncomb * 4);
    dev_free(c, (void **)&c->pdg, N1 * 4);
    dev_free(c, (void **)&c->X, (N1 + 63) / 64 * 64 * (size_t)c->ldx * 4);
    dev_free(c, (void **)&c->keys, (N1 + 8) * 8);
    dev_free(c, (void **)&c->vsimple, N1);
    dev_free(c, (void **)&c->plan, (size_t)c->nplan * 16);
    dev_free(c, (void **)&c->split_rows, (size_t)c->nsplit * 4);
    c->nplan = 0;
    c->nsplit = 0;
    free_moves(c);
    free_batch(c);
    c->graphs = c->mapped = c->diag_valid = c->dfdiag_valid = false;
    c->gain_cols = 0;
    c->sweep_kind = 0;
    c->sparse_ok = c->all_mapped = false;
    c->rot3_tuned = false;
    c->permset_threads_k = 0;
}

int bm_destroy(bm_ctx *c) {
    if (!c) return BM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_graph_buffers(c);
    block_free(c->stage);
    std::lock_guard<std::mutex> g(sync_mutex());
    for (auto &r : c->pending) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->hres) (void)hipHostFree(c->hres);
    if (c->hrows) (void)hipHostFree(c->hrows);
    if (c->hgrpres) (void)hipHostFree(c->hgrpres);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BM_OK;
}

int32_t bm_num_vertices(const bm_ctx *c) { return c ? c->n : 0; }
int64_t bm_device_bytes(const bm_ctx *c) { return c ? c->dev_bytes : 0; }

int bm_upload_graphs(bm_ctx *c, int32_t n, const int32_t *s1, const int32_t *d1, const int32_t *w1, int64_t e1,
                     const int32_t *s2, const int32_t *d2, const int32_t *w2, int64_t e2) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (n < 1 || n > 65534) return set_error(BM_ERANGE, "n=%d outside [1, 65534]", n);
    if (e1 < 0 || e2 < 0 || (e1 && (!s1 || !d1 || !w1)) || (e2 && (!s2 || !d2 || !w2)))
        return set_error(BM_EINVAL, "bad edge arrays");
    if (2 * e1 >= (int64_t)INT32_MAX) return set_error(BM_ERANGE, "too many G1 edges (%lld)", (long long)e1);
    for (int64_t i = 0; i < e1; i++) {
        if (s1[i] < 0 || s1[i] > n || d1[i] < 0 || d1[i] > n)
            return set_error(BM_EINVAL, "G1 line %lld: vertex outside [0,%d]", (long long)i, n);
        if (w1[i] < -32768) return set_error(BM_ERANGE, "G1 line %lld: weight %d < -32768", (long long)i, w1[i]);
    }
    for (int64_t i = 0; i < e2; i++)
        if (s2[i] < 0 || s2[i] > n || d2[i] < 0 || d2[i] > n)
            return set_error(BM_EINVAL, "G2 line %lld: vertex outside [0,%d]", (long long)i, n);
    BM_HIP(hipSetDevice(c->device));
    BM_HIP(hipStreamSynchronize(c->stream));
    free_graph_buffers(c);

    const int N1 = n + 1;
    c->n = n;
    c->N1 = N1;
    c->ld = (N1 + 63) / 64 * 64;
    c->e1 = e1;

    // The host list builds run on a few threads each (stable counting sorts
    // over contiguous chunks of the input: a chunk's entries keep their file
    // order, and the chunks are placed one after another, so the lists come
    // out exactly as a single pass would lay them down).
    auto threads_for = [](int64_t work) { return (int)std::max<int64_t>(1, std::min<int64_t>(4, work / 200000)); };
    auto parallel = [](int T, const std::function<void(int)> &fn) {
        std::vector<std::thread> th;
        for (int t = 1; t < T; t++) {
            try {
                th.emplace_back(fn, t);
            } catch (...) {
                fn(t);  // no thread: run the part here
            }
        }
        fn(0);
        for (auto &x : th) x.join();
    };
    // per-thread stable placement: counts[t][v] (of chunk t) become the
    // chunk's first position for key v, given each key's base position
    auto chunk_offsets = [](std::vector<int32_t> &counts, int T, int K, const std::function<int32_t(int)> &base) {
        for (int v = 0; v < K; v++) {
            int32_t o = base(v);
            for (int t = 0; t < T; t++) {
                const int32_t c = counts[(size_t)t * K + v];
                counts[(size_t)t * K + v] = o;
                o += c;
            }
        }
    };

    // ---- G2 (on a second host thread while this one builds the G1 lists):
    // last line per ordered pair, then its out- and in-lists packed (vertex << 16 | weight)
    std::vector<int32_t> tri;
    bool g2nonneg = true, g2_vertex0 = false, g2small = true;
    std::vector<int32_t> g2optr(N1 + 1, 0), g2iptr(N1 + 1, 0);
    std::vector<uint32_t> g2o, g2i;
    std::vector<int32_t> g2o_at_in, g2i_at_out;
    int64_t m2 = 0;
    auto build_g2 = [&]() {
        const int T = threads_for(e2);
        auto lo = [&](int64_t m, int t) { return m * t / T; };
        // lines by source, file order within a source
        std::vector<int32_t> p2(N1 + 1, 0), cnt((size_t)T * N1, 0);
        parallel(T, [&](int t) {
            int32_t *c = cnt.data() + (size_t)t * N1;
            for (int64_t i = lo(e2, t); i < lo(e2, t + 1); i++) c[s2[i]]++;
        });
        for (int v = 0; v < N1; v++) {
            int32_t sum = 0;
            for (int t = 0; t < T; t++) sum += cnt[(size_t)t * N1 + v];
            p2[v + 1] = p2[v] + sum;
        }
        chunk_offsets(cnt, T, N1, [&](int v) { return p2[v]; });
        std::vector<int32_t> od(e2), ow(e2);
        parallel(T, [&](int t) {
            int32_t *pos = cnt.data() + (size_t)t * N1;
            for (int64_t i = lo(e2, t); i < lo(e2, t + 1); i++) {
                const int p = pos[s2[i]]++;
                od[p] = d2[i];
                ow[p] = w2[i];
            }
        });
        // the last line of each (source, target) pair, sources in increasing order: each thread a range of
        // sources (split by line count) into its own part, the parts then laid end to end
        std::vector<int> vs(T + 1, N1);
        vs[0] = 0;
        for (int t = 1; t < T; t++)
            vs[t] = (int)(std::upper_bound(p2.begin(), p2.end(), (int32_t)lo(e2, t)) - p2.begin()) - 1;
        for (int t = 1; t <= T; t++) vs[t] = std::max(vs[t], vs[t - 1]);
        std::vector<std::vector<int32_t>> part(T);
        std::vector<char> neg(T, 0), big(T, 0), v0(T, 0);
        parallel(T, [&](int t) {
            std::vector<int32_t> seen(N1, -1);
            auto &out = part[t];
            out.reserve(3 * (size_t)(p2[vs[t + 1]] - p2[vs[t]]));
            for (int u = vs[t]; u < vs[t + 1]; u++)
                for (int e = p2[u + 1] - 1; e >= p2[u]; e--)
                    if (seen[od[e]] != u) {
                        seen[od[e]] = u;
                        out.push_back(u);
                        out.push_back(od[e]);
                        out.push_back(ow[e]);
                        if ((int16_t)ow[e] < 0) neg[t] = 1;
                        if ((int16_t)ow[e] < 0 || (int16_t)ow[e] > 16383) big[t] = 1;
                        if (u == 0 || od[e] == 0) v0[t] = 1;
                    }
        });
        size_t all = 0;
        for (int t = 0; t < T; t++) all += part[t].size();
        tri.reserve(all);
        for (int t = 0; t < T; t++) {
            tri.insert(tri.end(), part[t].begin(), part[t].end());
            std::vector<int32_t>().swap(part[t]);
            if (neg[t]) g2nonneg = false;
            if (big[t]) g2small = false;
            if (v0[t]) g2_vertex0 = true;
        }
        m2 = (int64_t)tri.size() / 3;
        // the out- and in-lists of the de-duplicated lines, in their order
        std::vector<int32_t> co((size_t)T * N1, 0), ci((size_t)T * N1, 0);
        parallel(T, [&](int t) {
            int32_t *a = co.data() + (size_t)t * N1, *b = ci.data() + (size_t)t * N1;
            for (int64_t i = lo(m2, t); i < lo(m2, t + 1); i++) {
                a[tri[3 * i]]++;
                b[tri[3 * i + 1]]++;
            }
        });
        for (int v = 0; v < N1; v++) {
            int32_t so = 0, si = 0;
            for (int t = 0; t < T; t++) so += co[(size_t)t * N1 + v], si += ci[(size_t)t * N1 + v];
            g2optr[v + 1] = g2optr[v] + so;
            g2iptr[v + 1] = g2iptr[v] + si;
        }
        chunk_offsets(co, T, N1, [&](int v) { return g2optr[v]; });
        chunk_offsets(ci, T, N1, [&](int v) { return g2iptr[v]; });
        g2o.resize(m2);
        g2i.resize(m2);
        g2o_at_in.resize(m2);
        g2i_at_out.resize(m2);
        parallel(T, [&](int t) {
            int32_t *po = co.data() + (size_t)t * N1, *pi = ci.data() + (size_t)t * N1;
            for (int64_t i = lo(m2, t); i < lo(m2, t + 1); i++) {
                const uint32_t u = (uint32_t)tri[3 * i], v = (uint32_t)tri[3 * i + 1];
                const uint32_t wv = (uint16_t)(int16_t)tri[3 * i + 2];
                const int jo = po[u]++, ji = pi[v]++;
                g2o[jo] = (v << 16) | wv;
                g2i[ji] = (u << 16) | wv;
                g2o_at_in[ji] = jo;  // the same line in the other list
                g2i_at_out[jo] = ji;
            }
        });
    };
    std::thread g2_thread;
    try {
        g2_thread = std::thread(build_g2);
    } catch (...) {
        build_g2();  // no second thread: build it here
    }

    // ---- G1 comb lists: out entries (file order) then in entries (file order)
    const int T1 = threads_for(e1);
    auto lo1 = [&](int64_t m, int t) { return m * t / T1; };
    std::vector<int32_t> cnt_out(N1, 0), cnt_in(N1, 0), po_t((size_t)T1 * N1, 0), pi_t((size_t)T1 * N1, 0);
    parallel(T1, [&](int t) {
        int32_t *a = po_t.data() + (size_t)t * N1, *b = pi_t.data() + (size_t)t * N1;
        for (int64_t i = lo1(e1, t); i < lo1(e1, t + 1); i++) {
            a[s1[i]]++;
            b[d1[i]]++;
        }
    });
    for (int v = 0; v < N1; v++)
        for (int t = 0; t < T1; t++) cnt_out[v] += po_t[(size_t)t * N1 + v], cnt_in[v] += pi_t[(size_t)t * N1 + v];
    std::vector<int32_t> cptr(N1 + 1, 0);
    for (int v = 0; v < N1; v++) cptr[v + 1] = cptr[v] + cnt_out[v] + cnt_in[v];
    chunk_offsets(po_t, T1, N1, [&](int v) { return cptr[v]; });
    chunk_offsets(pi_t, T1, N1, [&](int v) { return cptr[v] + cnt_out[v]; });
    const int64_t nc = 2 * e1;
    std::vector<int32_t> cr(nc), cw(nc), cdw(nc, 0), nb(nc);
    parallel(T1, [&](int t) {
        int32_t *pos_out = po_t.data() + (size_t)t * N1, *pos_in = pi_t.data() + (size_t)t * N1;
        for (int64_t i = lo1(e1, t); i < lo1(e1, t + 1); i++) {
            const int s = s1[i], d = d1[i];
            const int po = pos_out[s]++;
            cr[po] = N1 + d;
            cw[po] = w1[i];
            nb[po] = d;
            const int pi = pos_in[d]++;
            cr[pi] = s;
            cw[pi] = w1[i];
            nb[pi] = s;
        }
    });
    std::vector<int32_t>().swap(po_t);
    std::vector<int32_t>().swap(pi_t);
    // owner of the dense G1 cell: the last line of each ordered pair; the same
    // entries make the de-duplicated view (last weight wins), out part then
    // in part per vertex, whose lists are built on first use (host_dd)
    std::vector<int32_t> ddptr(N1 + 1, 0), ddmid(N1, 0), ddo(N1, 0), ddi(N1, 0);
    std::vector<uint8_t> last(nc, 0);
    {
        std::vector<int> vs(T1 + 1, N1);
        vs[0] = 0;
        for (int t = 1; t < T1; t++)
            vs[t] = (int)(std::upper_bound(cptr.begin(), cptr.end(), (int32_t)lo1(nc, t)) - cptr.begin()) - 1;
        for (int t = 1; t <= T1; t++) vs[t] = std::max(vs[t], vs[t - 1]);
        parallel(T1, [&](int t) {
            std::vector<int32_t> stamp(N1, -1);
            for (int v = vs[t]; v < vs[t + 1]; v++) {
                const int mid = cptr[v] + cnt_out[v];
                for (int e = mid - 1; e >= cptr[v]; e--)
                    if (stamp[nb[e]] != 2 * v) {
                        stamp[nb[e]] = 2 * v;
                        cdw[e] = (int16_t)cw[e];
                        last[e] = 1;
                        ddo[v]++;
                    }
                for (int e = cptr[v + 1] - 1; e >= mid; e--)
                    if (stamp[nb[e]] != 2 * v + 1) {
                        stamp[nb[e]] = 2 * v + 1;
                        cdw[e] = (int16_t)cw[e];
                        last[e] = 1;
                        ddi[v]++;
                    }
            }
        });
    }
    int64_t ndd = 0;
    for (int v = 0; v < N1; v++) {
        ddptr[v] = (int32_t)ndd;
        ddmid[v] = (int32_t)(ndd + ddo[v]);
        ndd += ddo[v] + ddi[v];
    }
    ddptr[N1] = (int32_t)ndd;
    // the flat out-line list in CSR order (score, all-pairs corrections) is
    // built on the device from the comb lists; the lines of vertex 0 come
    // first and the score skips them
    c->e1_v0 = cnt_out[0];
    std::vector<int32_t> outptr(N1 + 1, 0);
    for (int v = 0; v < N1; v++) outptr[v + 1] = outptr[v] + cnt_out[v];
    c->ncomb = nc;
    c->ndd = ndd;

    // weight facts used by the packed fast paths
    bool nonneg = true;
    int32_t maxw = 0;
    for (int64_t i = 0; i < e1; i++) {
        if (w1[i] < 0) nonneg = false;
        maxw = std::max(maxw, w1[i]);
    }

    if (g2_thread.joinable()) g2_thread.join();
    c->g2nonneg = g2nonneg;
    c->g2q_small = g2small;
    c->g1nonneg = nonneg;
    c->nonneg = nonneg && g2nonneg;
    // every line its own dense cell: no self-loop, no repeated pair, no vertex 0, weights in [1, 32767]
    c->dense_equal = c->ndd == c->ncomb;
    for (int64_t i = 0; i < e1 && c->dense_equal; i++)
        if (s1[i] == d1[i] || s1[i] == 0 || d1[i] == 0 || w1[i] < 1 || w1[i] > 32767) c->dense_equal = false;
    c->max_w1 = maxw;
    c->g2_vertex0 = g2_vertex0;
    c->ng2 = m2;

    // ---- device buffers
    size_t ld = (size_t)c->ld;
    int rc;
    if ((rc = dev_alloc(c, (void **)&c->W2, 2 * (size_t)N1 * ld * 2))) goto fail;
    if ((rc = dev_alloc(c, (void **)&c->Q, 2 * (size_t)N1 * ld * 2))) goto fail;
    if ((rc = dev_alloc(c, (void **)&c->map, (size_t)N1 * 4))) goto fail;
    if ((rc = dev_alloc(c, (void **)&c->diag, (size_t)N1 * 4))) goto fail;
    if ((rc = dev_alloc(c, (void **)&c->keys, ((size_t)N1 + 8) * 8))) goto fail;
    if ((rc = upload(c, &c->cptr, cptr))) goto fail;
    if ((rc = upload(c, &c->cr, cr))) goto fail;
    if ((rc = upload(c, &c->cw, cw))) goto fail;
    if ((rc = upload(c, &c->cdw, cdw))) goto fail;
    {
        std::vector<int32_t> mid(N1), neg(N1, 0);
        for (int v = 0; v < N1; v++) {
            mid[v] = cptr[v] + cnt_out[v];
            for (int e = cptr[v]; e < cptr[v + 1]; e++) neg[v] += std::min(cw[e], 0);
        }
        if ((rc = upload(c, &c->cmid, mid))) goto fail;
        if ((rc = upload(c, &c->cneg, neg))) goto fail;
        // (neighbour, weight) pairs and the flat out-line list, from what is on the device now
        if ((rc = dev_alloc(c, (void **)&c->crw, (size_t)nc * 8))) goto fail;
        if ((rc = dev_alloc(c, (void **)&c->esrc, (size_t)e1 * 4))) goto fail;
        if ((rc = dev_alloc(c, (void **)&c->edst, (size_t)e1 * 4))) goto fail;
        if ((rc = dev_alloc(c, (void **)&c->ew, (size_t)e1 * 4))) goto fail;
        if ((rc = dev_alloc(c, (void **)&c->edw, (size_t)e1 * 4))) goto fail;
        if ((rc = ensure_stage(c, (size_t)(N1 + 1) * 4))) goto fail;
        hipError_t e = hipMemcpyAsync(c->stage, outptr.data(), (size_t)(N1 + 1) * 4, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "copy outptr"); goto fail; }
        if (nc > 0) {
            hipLaunchKernelGGL(k_interleave, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream, c->cr, c->cw,
                               nc, (int2 *)c->crw);
            hipLaunchKernelGGL(k_flat_out, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream, c->cptr,
                               c->cmid, (const int32_t *)c->stage, c->cr, c->cw, c->cdw, N1, nc, c->esrc, c->edst,
                               c->ew, c->edw);
            if ((rc = launch_check("flat_out"))) goto fail;
        }
        // the stage buffer is reused by the W2 scatter below: wait for these kernels first
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "flat_out sync"); goto fail; }
    }
    if ((rc = upload(c, &c->g2optr, g2optr))) goto fail;
    if ((rc = upload(c, &c->g2iptr, g2iptr))) goto fail;
    if ((rc = upload(c, &c->g2o, g2o))) goto fail;
    if ((rc = upload(c, &c->g2i, g2i))) goto fail;
    if ((rc = upload(c, &c->g2o_at_in, g2o_at_in))) goto fail;
    if ((rc = upload(c, &c->g2i_at_out, g2i_at_out))) goto fail;
    {
        // vertices on which every delta kind equals the exact score change:
        // no self-loop line and no repeated (u,v) line touches them, and no
        // line leaves vertex 0 for them (the deltas count such a line as an
        // in-entry, S_opt does not score it)
        std::vector<uint8_t> simple(N1, 1);
        for (int64_t i = 0; i < e1; i++)
            if (s1[i] == d1[i] || s1[i] == 0) simple[d1[i]] = 0;
        simple[0] = 0;
        for (int v = 0; v < N1; v++)
            if (ddptr[v + 1] - ddptr[v] != cptr[v + 1] - cptr[v]) simple[v] = 0;
        if ((rc = upload(c, &c->vsimple, simple))) goto fail;
    }
    if ((rc = dev_alloc(c, (void **)&c->pinv, (size_t)N1 * 4))) goto fail;
    // the translated out- and in-lists back to back: g2iy = g2oy + ng2
    if ((rc = dev_alloc(c, (void **)&c->g2oy, (size_t)c->ng2 * 8))) goto fail;
    c->g2iy = c->g2oy + c->ng2;
    {
        hipError_t e = hipMemsetAsync(c->W2, 0, 2 * (size_t)N1 * ld * 2, c->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "memset W2"); goto fail; }
        e = hipMemsetAsync(c->Q, 0, 2 * (size_t)N1 * ld * 2, c->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "memset Q"); goto fail; }
        if (m2 > 0) {
            if ((rc = ensure_stage(c, tri.size() * 4))) goto fail;
            e = hipMemcpyAsync(c->stage, tri.data(), tri.size() * 4, hipMemcpyHostToDevice, c->stream);
            if (e != hipSuccess) { rc = hip_fail(e, "copy G2"); goto fail; }
            hipLaunchKernelGGL(k_scatter_w2, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, c->stream, c->W2,
                               N1, c->ld, (const int32_t *)c->stage, m2);
            if ((rc = launch_check("scatter_w2"))) goto fail;
        }
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "upload sync"); goto fail; }
    }
    c->g2_maxlen = 0;
    for (int v = 0; v < N1; v++)
        c->g2_maxlen = std::max(c->g2_maxlen, std::max(g2optr[v + 1] - g2optr[v], g2iptr[v + 1] - g2iptr[v]));
    c->hg2optr = g2optr;
    c->hg2iptr = g2iptr;
    c->hmap.assign(N1, 0);
    c->hcdeg.resize(N1);
    c->hg2deg.resize(N1);
    for (int v = 0; v < N1; v++) {
        c->hcdeg[v] = cptr[v + 1] - cptr[v];
        c->hg2deg[v] = (g2optr[v + 1] - g2optr[v]) + (g2iptr[v + 1] - g2iptr[v]);
    }
    {
        std::vector<int32_t> mid(N1);
        for (int v = 0; v < N1; v++) mid[v] = cptr[v] + cnt_out[v];
        c->hcmid.swap(mid);
    }
    c->hcptr.swap(cptr);
    c->hcr.swap(cr);
    c->hcw.swap(cw);
    c->hlast.swap(last);
    c->hddptr.swap(ddptr);
    c->hddmid.swap(ddmid);
    c->graphs = true;
    c->mapped = false;
    return BM_OK;
fail:
    free_graph_buffers(c);
    return rc;
}

int bm_set_mapping(bm_ctx *c, const int32_t *mapping) {
    if (!c || !mapping) return set_error(BM_EINVAL, "NULL argument");
    if (!c->graphs) return set_error(BM_ESTATE, "graphs not uploaded");
    for (int v = 0; v < c->N1; v++)
        if (mapping[v] < 0 || mapping[v] > c->n)
            return set_error(BM_EINVAL, "mapping[%d]=%d outside [0,%d]", v, mapping[v], c->n);
    BM_HIP(hipSetDevice(c->device));
    c->hmap.assign(mapping, mapping + c->N1);
    c->map_epoch++;
    BM_HIP(hipMemcpyAsync(c->map, c->hmap.data(), (size_t)c->N1 * 4, hipMemcpyHostToDevice, c->stream));
    // inverse of the mapping on mapped vertices; the sparse all-pairs kernel
    // needs it to be a function (p injective where p != 0) and W2 row/column 0 empty
    c->hpinv.assign(c->N1, -1);
    c->sparse_ok = !c->g2_vertex0;
    for (int v = 0; v < c->N1 && c->sparse_ok; v++) {
        const int g = mapping[v];
        if (g == 0) continue;
        if (c->hpinv[g] >= 0) c->sparse_ok = false;
        else c->hpinv[g] = v;
    }
    c->all_mapped = c->sparse_ok;
    for (int g = 1; g < c->N1 && c->all_mapped; g++)
        if (c->hpinv[g] < 0) c->all_mapped = false;
    if (c->sparse_ok)
        BM_HIP(hipMemcpyAsync(c->pinv, c->hpinv.data(), (size_t)c->N1 * 4, hipMemcpyHostToDevice, c->stream));
    c->g2y_dirty = true;
    int rc = launch_rebuild_p(c);
    if (rc) return rc;
    c->p_stale = false;
    BM_HIP(hipStreamSynchronize(c->stream));
    c->mapped = true;
    c->diag_valid = c->dfdiag_valid = false;
    return BM_OK;
}

int bm_get_mapping(bm_ctx *c, int32_t *mapping) {
    if (!c || !mapping) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    memcpy(mapping, c->hmap.data(), (size_t)c->N1 * 4);
    return BM_OK;
}

int bm_score(bm_ctx *c, int64_t *score) {
    if (!c || !score) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    BM_HIP(hipSetDevice(c->device));
    unsigned long long *out = c->keys + c->N1 + 2;
    BM_HIP(hipMemsetAsync(out, 0, 8, c->stream));
    int64_t m = c->e1;
    int blocks = (int)std::min<int64_t>((m - c->e1_v0 + 255) / 256, 4096);
    if (blocks < 1) blocks = 1;
    BM_LAUNCH(c, "score", hipLaunchKernelGGL(k_score, dim3(blocks), dim3(256), 0, c->stream, out, c->esrc, c->edst,
                                             c->ew, c->e1_v0, m, c->Q, c->ld, nullptr));
    BM_HIP(hipMemcpyAsync(c->hres, out, 8, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    *score = (int64_t)c->hres[0];
    return BM_OK;
}

static int check_vertex(const bm_ctx *c, int32_t v) { return v >= 0 && v <= c->n; }

int bm_apply_swap_deferred(bm_ctx *c, int32_t a, int32_t b) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (!check_vertex(c, a) || !check_vertex(c, b) || a == b)
        return set_error(BM_EINVAL, "apply_swap(%d,%d): bad vertices", a, b);
    BM_HIP(hipSetDevice(c->device));
    long long *slot = (long long *)(c->keys + c->N1 + 4);
    BM_LAUNCH(c, "swap_change", hipLaunchKernelGGL(k_swap_change, dim3(1), dim3(256), 0, c->stream, slot, a, b,
                                                   c->cptr, (const int2 *)c->crw, c->map, c->Q, c->W2, c->N1,
                                                   c->ld));
    int32_t set[2] = {a, b}, src[2] = {1, 0};
    int rc = permute_rows_cols(c, 2, set, src);
    if (rc) return rc;
    BM_HIP(hipMemcpyAsync(c->hres + 2, slot, 8, hipMemcpyDeviceToHost, c->stream));  // pinned
    return BM_OK;
}

int bm_deferred_change(bm_ctx *c, int64_t *change) {
    if (!c || !change) return set_error(BM_EINVAL, "NULL argument");
    BM_HIP(hipSetDevice(c->device));
    BM_HIP(hipStreamSynchronize(c->stream));
    *change = (int64_t)(long long)c->hres[2];
    return BM_OK;
}

int bm_apply_swap_scored(bm_ctx *c, int32_t a, int32_t b, int64_t *change) {
    if (!c || !change) return set_error(BM_EINVAL, "NULL argument");
    const int rc = bm_apply_swap_deferred(c, a, b);
    return rc ? rc : bm_deferred_change(c, change);
}

int bm_apply_swap(bm_ctx *c, int32_t a, int32_t b) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (!check_vertex(c, a) || !check_vertex(c, b) || a == b)
        return set_error(BM_EINVAL, "apply_swap(%d,%d): bad vertices", a, b);
    BM_HIP(hipSetDevice(c->device));
    int32_t set[2] = {a, b}, src[2] = {1, 0};
    return permute_rows_cols(c, 2, set, src);  // enqueued: every call that returns data waits for it
}

// triple.cu:903-934: new (p(m1),p(m2),p(m3)) = table[p] applied to (f1,f2,f3)
int bm_apply_rot3(bm_ctx *c, int32_t m1, int32_t m2, int32_t m3, int32_t p) {
    static const int tab[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (!check_vertex(c, m1) || !check_vertex(c, m2) || !check_vertex(c, m3) || m1 == m2 || m1 == m3 || m2 == m3 ||
        p < 0 || p > 5)
        return set_error(BM_EINVAL, "apply_rot3: bad arguments");
    if (p == 0) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    int32_t set[3] = {m1, m2, m3}, src[3] = {tab[p][0], tab[p][1], tab[p][2]};
    return permute_rows_cols(c, 3, set, src);  // enqueued: every call that returns data waits for it
}

int bm_apply_perm(bm_ctx *c, int32_t k, const int32_t *set, const int32_t *perm) {
    if (!c || !set || !perm) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (k < 1 || k > 8) return set_error(BM_EINVAL, "k=%d outside [1,8]", k);
    int seen = 0;
    for (int i = 0; i < k; i++) {
        if (!check_vertex(c, set[i])) return set_error(BM_EINVAL, "set[%d]=%d bad vertex", i, set[i]);
        for (int j = 0; j < i; j++)
            if (set[j] == set[i]) return set_error(BM_EINVAL, "set has duplicate vertex %d", set[i]);
        if (perm[i] < 0 || perm[i] >= k || (seen & (1 << perm[i])))
            return set_error(BM_EINVAL, "perm is not a permutation of 0..%d", k - 1);
        seen |= 1 << perm[i];
    }
    BM_HIP(hipSetDevice(c->device));
    return permute_rows_cols(c, k, set, perm);  // enqueued: every call that returns data waits for it
}

uint64_t bm_rng_next(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // extern "C"

// Perturbation weights of quick.cu:766-793: 0.7 x out + 0.3 x in, where out
// (in) sums the positive dense G1 cells W1[v][j] (W1[j][v]) over j = 1..n in
// ascending j, in float.  The cells are the de-duplicated lists' weights as
// int16 (the reference's short matrix, last line wins); they are copied back
// from the device once and sorted by j so that the float sums round in the
// reference's order.
static int vertex_scores(bm_ctx *c) {
    if (!c->hvscore.empty()) return BM_OK;
    const int N1 = c->N1;
    const std::vector<int32_t> &ptr = c->hddptr, &mid = c->hddmid;
    std::vector<int32_t> r, w;
    host_dd(c, r, w);
    std::vector<float> score(N1, 0.f);
    std::vector<std::pair<int, int>> cells;
    auto sum_sorted = [&](int e0, int e1, bool out) {
        cells.clear();
        for (int e = e0; e < e1; e++) {
            const int j = out ? r[e] - N1 : r[e];
            const int cell = (int16_t)w[e];
            if (j >= 1 && cell > 0) cells.emplace_back(j, cell);
        }
        std::sort(cells.begin(), cells.end());
        float s = 0.f;
        for (const auto &jc : cells) s += (float)jc.second;
        return s;
    };
    for (int v = 1; v < N1; v++) {
        const float out_w = sum_sorted(ptr[v], mid[v], true);
        const float in_w = sum_sorted(mid[v], ptr[v + 1], false);
        const float a = 0.7f * out_w;  // two statements: no fused multiply-add
        const float b = 0.3f * in_w;
        score[v] = a + b;
    }
    c->hvscore.swap(score);
    return BM_OK;
}

// Lists of the dense-matrix swap delta of quick.cu:48-250: for every i in
// 1..n other than a and b it adds the terms of the cells W1[a][i], W1[i][a]
// (and the same for b) that are positive as short, then the two cells between
// a and b.  Those are the de-duplicated entries whose int16 cell is > 0,
// minus self-loops and vertex 0; with each entry owning its cell (dense-cell
// weight = weight) the list-form kernels compute exactly that delta.
void bm::host_dd(const bm_ctx *c, std::vector<int32_t> &r, std::vector<int32_t> &w) {
    // the last entry of each (direction, neighbour) pair, taken from the back of each part as upload marked them
    r.resize((size_t)c->ndd);
    w.resize((size_t)c->ndd);
    int64_t k = 0;
    for (int v = 0; v < c->N1; v++) {
        for (int e = c->hcmid[v] - 1; e >= c->hcptr[v]; e--)
            if (c->hlast[e]) r[k] = c->hcr[e], w[k++] = c->hcw[e];
        for (int e = c->hcptr[v + 1] - 1; e >= c->hcmid[v]; e--)
            if (c->hlast[e]) r[k] = c->hcr[e], w[k++] = c->hcw[e];
    }
}

int bm::ensure_dd(bm_ctx *c) {
    if (c->ddptr) return BM_OK;
    std::vector<int32_t> r, w;
    host_dd(c, r, w);
    int rc;
    if ((rc = upload(c, &c->ddptr, c->hddptr)) || (rc = upload(c, &c->ddmid, c->hddmid)) ||
        (rc = upload(c, &c->ddr, r)) || (rc = upload(c, &c->ddw, w)))
        return rc;
    return BM_OK;
}

int bm::ensure_dense_lists(bm_ctx *c) {
    if (c->dfptr) return BM_OK;
    const int N1 = c->N1;
    const std::vector<int32_t> &ptr = c->hddptr;
    std::vector<int32_t> r, w;
    host_dd(c, r, w);
    // out entries (r >= N1) stay ahead of in entries, as in the lists they come from
    std::vector<int32_t> fptr(N1 + 1, 0), fmid(N1, 0), fr, fw, frw, deg(N1, 0), es, ed, ew;
    fr.reserve((size_t)c->ndd);
    fw.reserve((size_t)c->ndd);
    frw.reserve(2 * (size_t)c->ndd);
    for (int v = 0; v < N1; v++) {
        fptr[v] = fmid[v] = (int32_t)fr.size();
        if (v == 0) continue;
        for (int e = ptr[v]; e < ptr[v + 1]; e++) {
            const int nb = r[e] >= N1 ? r[e] - N1 : r[e];
            const int cell = (int16_t)w[e];
            if (cell <= 0 || nb == 0 || nb == v) continue;
            fr.push_back(r[e]);
            fw.push_back(cell);
            frw.push_back(r[e]);
            frw.push_back(cell);
            c->dfmaxw = std::max(c->dfmaxw, cell);
            if (r[e] >= N1) {
                fmid[v] = (int32_t)fr.size();
                es.push_back(v);
                ed.push_back(nb);
                ew.push_back(cell);
            }
        }
        deg[v] = (int32_t)fr.size() - fptr[v];
    }
    fptr[N1] = (int32_t)fr.size();
    int rc;
    c->ndf = (int64_t)fr.size();
    c->ndfout = (int64_t)es.size();
    const std::vector<int32_t> zeros(N1, 0);
    if ((rc = upload(c, &c->dfr, fr)) || (rc = upload(c, &c->dfw, fw)) || (rc = upload(c, &c->dfrw, frw)) ||
        (rc = upload(c, &c->dfmid, fmid)) || (rc = upload(c, &c->dfneg, zeros)) ||
        (rc = upload(c, &c->dfes, es)) || (rc = upload(c, &c->dfed, ed)) || (rc = upload(c, &c->dfew, ew)) ||
        (rc = dev_alloc(c, (void **)&c->dfdiag, (size_t)N1 * 4)) || (rc = upload(c, &c->dfptr, fptr))) {
        dev_free(c, (void **)&c->dfes, (size_t)c->ndfout * 4);
        dev_free(c, (void **)&c->dfed, (size_t)c->ndfout * 4);
        dev_free(c, (void **)&c->dfew, (size_t)c->ndfout * 4);
        c->ndfout = 0;
        dev_free(c, (void **)&c->dfr, (size_t)c->ndf * 4);
        dev_free(c, (void **)&c->dfw, (size_t)c->ndf * 4);
        dev_free(c, (void **)&c->dfrw, (size_t)c->ndf * 8);
        dev_free(c, (void **)&c->dfmid, (size_t)N1 * 4);
        dev_free(c, (void **)&c->dfneg, (size_t)N1 * 4);
        dev_free(c, (void **)&c->dfdiag, (size_t)N1 * 4);
        dev_free(c, (void **)&c->dfptr, (size_t)(N1 + 1) * 4);
        c->ndf = 0;
        return rc;
    }
    c->hdfdeg.swap(deg);
    c->dfdiag_valid = false;
    return BM_OK;
}

// quick.cu:796-814: r = U[0,1] x (remaining total); the first unused vertex
// whose running sum reaches r, else the first unused vertex.
static int draw_weighted(const std::vector<float> &score, const std::vector<uint8_t> &used, int n, float total,
                         uint64_t *rng) {
    const float u = (float)(bm_rng_next(rng) >> 40) * (1.0f / 16777216.0f);
    const float r = u * total;
    float cum = 0.f;
    for (int i = 1; i <= n; i++)
        if (!used[i]) {
            cum += score[i];
            if (cum >= r) return i;
        }
    for (int i = 1; i <= n; i++)
        if (!used[i]) return i;
    return 1;
}

extern "C" {

int bm_perturb(bm_ctx *c, int32_t k, int mode, uint64_t *rng, int32_t *selected) {
    if (!c || !rng) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (mode != BM_PERTURB_OMP && mode != BM_PERTURB_WEIGHTED) return set_error(BM_EINVAL, "bad mode %d", mode);
    const int n = c->n;
    if (k < 1 || 2 * (int64_t)k > n) return set_error(BM_EINVAL, "k=%d: need 1 <= k and 2k <= n=%d", k, n);
    BM_HIP(hipSetDevice(c->device));
    std::vector<int32_t> sel;
    sel.reserve(2 * (size_t)k);
    std::vector<uint8_t> used((size_t)n + 1, 0);
    if (mode == BM_PERTURB_OMP) {  // alignment-omp.c:506-513: rand() % n
        while ((int)sel.size() < 2 * k) {
            const int idx = (int)(bm_rng_next(rng) % (uint64_t)n);
            if (!used[idx]) {
                used[idx] = 1;
                sel.push_back(idx);
            }
        }
    } else {
        int rc = vertex_scores(c);
        if (rc) return rc;
        float total = 0.f;  // quick.cu:834-838
        for (int i = 1; i <= n; i++) total += c->hvscore[i];
        while ((int)sel.size() < k) {  // :842-850
            const int idx = draw_weighted(c->hvscore, used, n, total, rng);
            used[idx] = 1;
            sel.push_back(idx);
            total -= c->hvscore[idx];
        }
        while ((int)sel.size() < 2 * k) {  // :853-860: rand() % n + 1
            const int idx = (int)(bm_rng_next(rng) % (uint64_t)n) + 1;
            if (!used[idx]) {
                used[idx] = 1;
                sel.push_back(idx);
            }
        }
    }
    // disjoint swaps sel[i] <-> sel[i + k]: one row/column patch each, one sync
    for (int i = 0; i < k; i++) {
        int32_t set[2] = {sel[i], sel[i + k]}, src[2] = {1, 0};
        int rc = permute_rows_cols(c, 2, set, src);
        if (rc) return rc;
    }
    BM_HIP(hipStreamSynchronize(c->stream));
    if (selected) memcpy(selected, sel.data(), sizeof(int32_t) * 2 * (size_t)k);
    return BM_OK;
}

int bm_set_timing(bm_ctx *c, int enable) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!enable) {
        int rc = resolve_timers(c);
        if (rc) return rc;
    }
    c->timing = enable != 0;
    return BM_OK;
}

int bm_kernel_time(bm_ctx *c, const char *kernel, int64_t *launches, double *total_ms) {
    if (!c || !kernel) return set_error(BM_EINVAL, "NULL argument");
    int rc = resolve_timers(c);
    if (rc) return rc;
    for (auto &k : c->kstats)
        if (k.name == kernel) {
            if (launches) *launches = k.launches;
            if (total_ms) *total_ms = k.ms;
            return BM_OK;
        }
    if (launches) *launches = 0;
    if (total_ms) *total_ms = 0.0;
    return BM_OK;
}

int bm_reset_timing(bm_ctx *c) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    int rc = resolve_timers(c);
    if (rc) return rc;
    for (auto &k : c->kstats) {
        k.launches = 0;
        k.ms = 0.0;
    }
    return BM_OK;
}

}  // extern "C"
