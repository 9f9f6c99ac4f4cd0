// Delta_B pair-swap kernels (graph-alignment-optimized.c:321-408) for gfx950.
//
//   k_eval_pairs : a move list, 16 lanes per pair walking the four neighbour
//                  lists exactly as the reference loop does.
//   k_row_eval   : one row a against a range of b (first-positive or argmax):
//                  push part X(a,b) streamed coalesced over b, pull part
//                  X(b,a) from column a of P staged in LDS, list corrections
//                  for b adjacent to a, block reduction, one atomic per block.
//   k_gain + k_corr + k_combine : the all-pairs sweep.  X(a,b) for every a, b
//                  (sparse x dense min-plus product over the comb lists), then
//                  Delta_B(a,b) = X(a,b) + X(b,a) - X(a,a) - X(b,b) + Corr(a,b)
//                  on 64x64 tiles with row-best and global arg-max reductions.
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include <hipcub/hipcub.hpp>

#include "bm_device.h"
#include "bm_wave.h"
#include "bm_internal.h"

using namespace bm;

__device__ __forceinline__ int q16(const int16_t *__restrict__ Q, size_t row, int ld, int col) {
    return (int)Q[row * (size_t)ld + col];
}


// ------------------------------------------------------------ move list

// Delta_B for pairs (A[i], B[i]).  16-lane groups, 4 pairs per wave.
__global__ __launch_bounds__(256) void k_eval_pairs(const int32_t *__restrict__ A, const int32_t *__restrict__ B,
                                                    int64_t count, int32_t *__restrict__ out,
                                                    const int32_t *__restrict__ cptr, const int32_t *__restrict__ cr,
                                                    const int32_t *__restrict__ cw, const int32_t *__restrict__ cdw,
                                                    const int16_t *__restrict__ Q, int N1, int ld) {
    const int64_t pair = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int gl = threadIdx.x & 15;
    int sum = 0;
    int a = 0, b = 0;
    if (pair < count) {
        a = A[pair];
        b = B[pair];
        const int pab = q16(Q, a, ld, b), pba = q16(Q, b, ld, a);
        // lists of a: entries toward b are skipped, the owner line of a->b
        // carries the dense-cell term (:392-397)
        for (int e = cptr[a] + gl; e < cptr[a + 1]; e += 16) {
            const int r = cr[e], w = cw[e];
            const int nb = r >= N1 ? r - N1 : r;
            if (nb == b) {
                const int dw = cdw[e];
                if (r >= N1 && dw > 0) sum += min(dw, pba) - min(dw, pab);
            } else {
                sum += min(w, q16(Q, r, ld, b)) - min(w, q16(Q, r, ld, a));
            }
        }
        for (int e = cptr[b] + gl; e < cptr[b + 1]; e += 16) {
            const int r = cr[e], w = cw[e];
            const int nb = r >= N1 ? r - N1 : r;
            if (nb == a) {
                const int dw = cdw[e];
                if (r >= N1 && dw > 0) sum += min(dw, pab) - min(dw, pba);  // :400-405
            } else {
                sum += min(w, q16(Q, r, ld, a)) - min(w, q16(Q, r, ld, b));
            }
        }
    }
    sum = bm_wave::group_sum<16>(sum);
    if (pair < count && gl == 0) out[pair] = sum;
}

// ------------------------------------------------------------ one row

constexpr int kRowCh = 128;  // b columns per block

__device__ __forceinline__ unsigned long long key_first(int b, int d) {
    return ((unsigned long long)(unsigned)b << 32) | (unsigned)d;
}
__device__ __forceinline__ unsigned long long key_max(int b, int d) {
    return ((unsigned long long)((unsigned)d ^ 0x80000000u) << 32) | (0xFFFFFFFFu - (unsigned)b);
}

using bm_wave::wave_max_u64;
using bm_wave::wave_min_u64;


template <bool kLdsCol>
__global__ __launch_bounds__(256) void k_row_eval(int a, int lo, int hi, int mode, int N1, int ld,
                                                  const int16_t *__restrict__ Q, const int32_t *__restrict__ cptr,
                                                  const int32_t *__restrict__ cr, const int32_t *__restrict__ cw,
                                                  const int32_t *__restrict__ cdw, const int32_t *__restrict__ diag,
                                                  unsigned long long *__restrict__ result,
                                                  int32_t *__restrict__ store) {
    extern __shared__ int16_t colA[];  // colA[r] = Q[r][a], r < 2*N1
    __shared__ int push[256];
    __shared__ int pull[kRowCh];
    __shared__ int corr[kRowCh];
    __shared__ unsigned long long red[4];

    const int tid = threadIdx.x;
    const int b0 = lo + blockIdx.x * kRowCh;
    const int nbk = min(kRowCh, hi + 1 - b0);
    const int16_t *prow_a = Q + (size_t)a * ld;          // P[a][.]  == Q[N1+d][a]
    const int16_t *pcol_a = Q + (size_t)(N1 + a) * ld;   // PT[a][.] == Q[s][a]
    if (kLdsCol) {
        for (int i = tid; i < N1; i += 256) {
            colA[i] = pcol_a[i];
            colA[N1 + i] = prow_a[i];
        }
    }
    if (tid < kRowCh) {
        pull[tid] = 0;
        corr[tid] = 0;
    }
    __syncthreads();

    const int ca0 = cptr[a], ca1 = cptr[a + 1];
    // push: X(a,b) = sum_{e in comb(a)} min(w, Q[r][b]); two list halves
    {
        const int col = tid & (kRowCh - 1), split = tid / kRowCh;
        int acc = 0;
        if (col < nbk) {
            const int b = b0 + col;
            for (int e = ca0 + split; e < ca1; e += 2) acc += min(cw[e], q16(Q, cr[e], ld, b));
        }
        push[tid] = acc;
    }
    // pull: X(b,a) = sum_{e in comb(b)} min(w, Q[r][a]); 16 lanes per b
    {
        const int g = tid >> 4, gl = tid & 15;
        for (int bb = g; bb < nbk; bb += 16) {
            const int v = b0 + bb;
            int s = 0;
            for (int e = cptr[v] + gl; e < cptr[v + 1]; e += 16) {
                const int r = cr[e];
                int q;
                if (kLdsCol) q = colA[r];
                else q = r < N1 ? (int)pcol_a[r] : (int)prow_a[r - N1];
                s += min(cw[e], q);
            }
            s = bm_wave::group_sum<16>(s);
            if (gl == 0) pull[bb] = s;
        }
    }
    // corrections for b adjacent to a
    {
        const int paa = prow_a[a];
        for (int e = ca0 + tid; e < ca1; e += 256) {
            const int r = cr[e];
            const bool out = r >= N1;
            const int nb = out ? r - N1 : r;
            if (nb == a || nb < b0 || nb >= b0 + nbk) continue;
            const int pbb = q16(Q, nb, ld, nb), pab = prow_a[nb], pba = q16(Q, nb, ld, a);
            atomicAdd(&corr[nb - b0], pair_corr(out, cw[e], cdw[e], paa, pbb, pab, pba));
        }
    }
    __syncthreads();

    if (store) {  // whole row: every delta, no reduction
        if (tid < nbk) {
            const int b = b0 + tid;
            store[b] = b == a || b == 0 ? 0 : push[tid] + push[tid + kRowCh] + pull[tid] - diag[a] - diag[b] + corr[tid];
        }
        return;
    }
    unsigned long long key = mode == BM_ROW_FIRST_POSITIVE ? ~0ull : 0ull;
    if (tid < nbk) {
        const int b = b0 + tid;
        if (b != a) {
            const int d = push[tid] + push[tid + kRowCh] + pull[tid] - diag[a] - diag[b] + corr[tid];
            if (mode == BM_ROW_FIRST_POSITIVE) {
                if (d > 0) key = key_first(b, d);
            } else {
                key = key_max(b, d);
            }
        }
    }
    key = mode == BM_ROW_FIRST_POSITIVE ? wave_min_u64(key) : wave_max_u64(key);
    if ((tid & 63) == 0) red[tid >> 6] = key;
    __syncthreads();
    if (tid == 0) {
        unsigned long long k = red[0];
        for (int i = 1; i < 4; i++) k = mode == BM_ROW_FIRST_POSITIVE ? min(k, red[i]) : max(k, red[i]);
        if (mode == BM_ROW_FIRST_POSITIVE) {
            if (k != ~0ull) atomicMin(result, k);
        } else if (k != 0ull) {
            atomicMax(result, k);
        }
    }
}

// Sparse form of one row (taken when p is injective on mapped vertices and
// no G2 line names vertex 0, the preconditions of k_gain_sparse):
//   X(a,b) = C_a + sum over comb entries e of a, non-zero Q[r_e][b]:
//            min(w_e, Q[r_e][b]) - min(w_e, 0)
//   X(b,a) = C_b + sum over non-zero Q[r][a], comb entries (r, w) of b:
//            min(w, Q[r][a]) - min(w, 0)
// The non-zeros of column a of Q are the G2 neighbours of p(a) taken back
// through p^-1: Q[s][a] != 0 for s = p^-1(g), g an in-neighbour of p(a), and
// Q[n+1+d][a] != 0 for d = p^-1(g), g an out-neighbour; the comb entries
// reading row s are the G1 out-lines s->b, those reading row n+1+d the G1
// in-lines b->d.  k_row_final forms Delta_B(a,b) = X(a,b) + X(b,a) - X(a,a)
// - X(b,b) and reduces it to the first positive or arg-max key.
//
// Deltas of rows a0 + y (grid y), the first over [lo0, hi], the others over
// [lo_next, hi] ([r + 1, hi] for lo_next < 0).  A row's list heads -- the comb
// entries of a (their G2 lists give X(a, .)) and the G2 neighbours of p(a)
// (the G1 lists of their preimages give X(., a)) -- go round-robin to the 16
// waves of each of the row's `parts` workgroups (grid x = parts x column
// windows); a wave walks its heads' lists a word per lane and adds into the
// workgroup's LDS row, which is then stored whole as one part row.  A C2 row
// has about 2 x 10^5 list words: as device-scope atomics into one shared row
// they bound the row at about 24 us, in LDS they cost a few.  Rows wider than
// the LDS budget are cut into column windows, each walking all heads and
// keeping the words that land in it.  k_row_final sums the parts.
constexpr int kRowWin = 36 * 1024;  // ints of one column window (144 KB of LDS)

__global__ __launch_bounds__(1024) void k_row_parts(int a0, int lo0, int lo_next, int hi, int mode, int N1,
                                                    const int32_t *__restrict__ cptr,
                                                    const int32_t *__restrict__ cmid, const int2 *__restrict__ crw,
                                                    const int32_t *__restrict__ cdw, const int32_t *__restrict__ map,
                                                    const int32_t *__restrict__ g2optr,
                                                    const uint32_t *__restrict__ g2oy,
                                                    const int32_t *__restrict__ g2iptr,
                                                    const uint32_t *__restrict__ g2iy, const int16_t *__restrict__ Q,
                                                    int ld, int parts, int win, int32_t *__restrict__ part,
                                                    unsigned long long *__restrict__ result) {
    extern __shared__ __align__(16) int acc[];
    const int x = blockIdx.x % parts, wlo = (blockIdx.x / parts) * win, whi = min(N1, wlo + win);
    const int row = blockIdx.y, a = a0 + row;
    const int lo = max(row == 0 ? lo0 : (lo_next < 0 ? a + 1 : lo_next), wlo), top = min(hi, whi - 1);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int i = tid; i < whi - wlo; i += 1024) acc[i] = 0;
    if (blockIdx.x == 0 && tid == 0) result[row] = mode == BM_ROW_FIRST_POSITIVE ? ~0ull : 0ull;
    __syncthreads();
    const int ca0 = cptr[a], dega = cptr[a + 1] - ca0;
    const int pa = map[a];
    const int i0 = g2iptr[pa], nin = g2iptr[pa + 1] - i0;
    const int o0 = g2optr[pa], nout = g2optr[pa + 1] - o0;
    const int H = dega + nin + nout;
    if (x == 0) {  // Corr(a, b) of the lines between a and b, one comb entry each
        const int paa = (int)Q[(size_t)a * ld + a];
        for (int e = ca0 + tid; e < ca0 + dega; e += 1024) {
            const int2 rw = crw[e];
            const bool out = rw.x >= N1;
            const int b = out ? rw.x - N1 : rw.x;
            if (b == a || b < lo || b > top) continue;
            const int pbb = (int)Q[(size_t)b * ld + b], pab = (int)Q[(size_t)a * ld + b],
                      pba = (int)Q[(size_t)b * ld + a];
            atomicAdd(&acc[b - wlo], pair_corr(out, rw.y, cdw[e], paa, pbb, pab, pba));
        }
    }
    for (int h = x * 16 + wave; h < H; h += parts * 16) {
        if (h < dega) {  // X(a, b): comb entry of a, its G2 list
            const int2 rw = crw[ca0 + h];
            const bool out = rw.x >= N1;
            const int g = map[out ? rw.x - N1 : rw.x];
            const uint32_t *L = out ? g2iy : g2oy;
            const int b0 = out ? g2iptr[g] : g2optr[g], b1 = out ? g2iptr[g + 1] : g2optr[g + 1];
            const int w = rw.y, wneg = min(w, 0);
            int j = b0 + lane;
            for (; j + 192 < b1; j += 256) {  // four list words in flight per lane
                uint32_t yv[4];
#pragma unroll
                for (int u = 0; u < 4; u++) yv[u] = L[j + 64 * u];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int y = yv[u] >> 16;
                    if (y != 0xFFFF && y >= lo && y <= top)
                        atomicAdd(&acc[y - wlo], min(w, (int)(int16_t)(yv[u] & 0xffff)) - wneg);
                }
            }
            for (; j < b1; j += 64) {
                const uint32_t yv = L[j];
                const int y = yv >> 16;
                if (y != 0xFFFF && y >= lo && y <= top)
                    atomicAdd(&acc[y - wlo], min(w, (int)(int16_t)(yv & 0xffff)) - wneg);
            }
        } else {  // X(b, a): a G2 neighbour of p(a), the G1 lines of its preimage
            const bool in = h < dega + nin;
            const uint32_t yv = in ? g2iy[i0 + h - dega] : g2oy[o0 + h - dega - nin];
            const int v = yv >> 16;
            if (v == 0xFFFF) continue;
            const int val = (int)(int16_t)(yv & 0xffff);
            // in-neighbour: rows s = v of P, read by the out-lines v->b (in entries of b);
            // out-neighbour: rows n+1+v of P^T, read by the in-lines b->v (out entries of b)
            const int b0 = in ? cptr[v] : cmid[v], b1 = in ? cmid[v] : cptr[v + 1];
            int e = b0 + lane;
            for (; e + 192 < b1; e += 256) {
                int2 rw[4];
#pragma unroll
                for (int u = 0; u < 4; u++) rw[u] = crw[e + 64 * u];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int b = in ? rw[u].x - N1 : rw[u].x;
                    if (b >= lo && b <= top) atomicAdd(&acc[b - wlo], min(rw[u].y, val) - min(rw[u].y, 0));
                }
            }
            for (; e < b1; e += 64) {
                const int2 rw = crw[e];
                const int b = in ? rw.x - N1 : rw.x;
                if (b >= lo && b <= top) atomicAdd(&acc[b - wlo], min(rw.y, val) - min(rw.y, 0));
            }
        }
    }
    __syncthreads();
    int32_t *dst = part + ((int64_t)row * parts + x) * N1;
    for (int i = wlo + tid; i < whi; i += 1024) dst[i] = acc[i - wlo];
}

__global__ __launch_bounds__(256) void k_row_final(int a0, int lo0, int lo_next, int lo_base, int hi, int mode,
                                                   const int32_t *__restrict__ part, int parts, int N1,
                                                   const int32_t *__restrict__ cneg,
                                                   const int32_t *__restrict__ diag,
                                                   unsigned long long *__restrict__ result,
                                                   int32_t *__restrict__ store) {
    __shared__ unsigned long long red[4];
    // grid y: the rows of k_row_parts' grid y; the blocks cover [lo_base, hi], lo_base
    // the lowest of the rows' lower ends
    const int a = a0 + (int)blockIdx.y;
    const int lo = blockIdx.y == 0 ? lo0 : (lo_next < 0 ? a + 1 : lo_next);
    part += (int64_t)blockIdx.y * parts * N1;  // this row's part rows
    result += blockIdx.y;
    const int b = lo_base + blockIdx.x * 256 + threadIdx.x;
    auto sum_parts = [&]() {  // eight part rows' loads in flight
        int x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int p = 0;
        for (; p + 8 <= parts; p += 8) {
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] += part[(int64_t)(p + u) * N1 + b];
        }
        for (; p < parts; p++) x[0] += part[(int64_t)p * N1 + b];
        return ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]));
    };
    if (store) {  // whole row (one row per launch): every delta, no reduction
        if (b >= lo && b <= hi) {
            const int x = sum_parts();
            store[b] = b == a || b == 0 ? 0 : x + cneg[a] + cneg[b] - diag[a] - diag[b];
        }
        return;
    }
    unsigned long long key = mode == BM_ROW_FIRST_POSITIVE ? ~0ull : 0ull;
    if (b >= lo && b <= hi) {
        const int x = sum_parts();
        if (b != a) {
            const int d = x + cneg[a] + cneg[b] - diag[a] - diag[b];
            if (mode == BM_ROW_FIRST_POSITIVE) {
                if (d > 0) key = key_first(b, d);
            } else {
                key = key_max(b, d);
            }
        }
    }
    key = mode == BM_ROW_FIRST_POSITIVE ? wave_min_u64(key) : wave_max_u64(key);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long k = red[0];
        for (int i = 1; i < 4; i++) k = mode == BM_ROW_FIRST_POSITIVE ? min(k, red[i]) : max(k, red[i]);
        if (mode == BM_ROW_FIRST_POSITIVE) {
            if (k != ~0ull) atomicMin(result, k);
        } else if (k != 0ull) {
            atomicMax(result, k);
        }
    }
}

// ------------------------------------------------------------ all pairs

// X[a][c..c+W-1] = sum_{e in comb(a)} min(w_e, Q[r_e][c..c+W-1]) for every row a.
//
// Work items are (W-column slab of Q, block of 16 rows), W = 64 * kCols.
// Each wave owns one row at a time, so its list entries are wave-uniform and
// come in as scalar loads (8 entries per s_load); each lane covers kCols
// columns, so one entry costs one coalesced 128*kCols-byte vector load per
// wave.  Workgroups are handed to the 8 XCDs round-robin, so items are
// numbered per XCD: XCD x walks its own contiguous range of slabs, all row
// blocks of one slab before the next, and keeps that slab (2(n+1) rows x
// 2W bytes) in its L2 / the MALL while every row reads it.  The slab width
// trades L2 residency (narrow) against load instructions (wide); the
// autotuner in bm_sweep_all_pairs picks kCols per graph.  With non-negative
// weights the mins of column pairs are summed as packed uint16
// (v_pk_min_i16 / v_pk_add_u16) and widened to int32 every `flush` entries,
// chosen so that flush * max(w) < 65536.
typedef short bm_s2 __attribute__((ext_vector_type(2)));
typedef unsigned short bm_u2 __attribute__((ext_vector_type(2)));
typedef int bm_i4 __attribute__((ext_vector_type(4)));
typedef int bm_i2 __attribute__((ext_vector_type(2)));

constexpr int kRowsWave = 4;   // rows per wave per work item
constexpr int kRowsWG = 16;    // rows per work item

template <int kCols>
struct GainLoad;
template <>
struct GainLoad<1> {
    typedef int16_t T;
};
template <>
struct GainLoad<2> {
    typedef int T;
};
template <>
struct GainLoad<4> {
    typedef int2 T;
};

__device__ __forceinline__ bm_u2 pk_min(int x, bm_s2 wp) {
    return __builtin_bit_cast(bm_u2, __builtin_elementwise_min(__builtin_bit_cast(bm_s2, x), wp));
}

constexpr int kGainUnroll = 16;  // list entries (vector loads) in flight per wave

template <int kCols, bool kPacked>
__global__ __launch_bounds__(256) void k_gain(const int32_t *__restrict__ cptr, const int2 *__restrict__ crw,
                                              const int16_t *__restrict__ Q, int ld, int32_t *__restrict__ X,
                                              int64_t ldx, int N1, int nslab, int nrb, int per_xcd, int flush) {
    typedef typename GainLoad<kCols>::T LoadT;
    const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    if (k >= per_xcd) return;
    const int t = xcd * per_xcd + k;
    if (t >= nslab * nrb) return;
    const int slab = t / nrb, rb = t - slab * nrb;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = slab * 64 * kCols + (threadIdx.x & 63) * kCols;
    const bool live = col < ld;
    const unsigned boff = live ? (unsigned)col * 2u : 0u;  // byte offset of this lane in a Q row
    const char *qbase = reinterpret_cast<const char *>(Q);
    const size_t row_bytes = (size_t)ld * 2;
    // row r's columns for this lane: scalar row base + 32-bit lane offset
    auto load = [&](int r) { return *reinterpret_cast<const LoadT *>(qbase + (size_t)r * row_bytes + boff); };
    int s[4];
    bm_u2 acc[2];
    auto consume = [&](LoadT v, int w) {
        if constexpr (kCols == 1) {
            s[0] += min(w, (int)v);
        } else if constexpr (kCols == 2) {
            if (kPacked) {
                const bm_s2 wp = {(short)w, (short)w};
                acc[0] += pk_min(v, wp);
            } else {
                s[0] += min(w, (int)(int16_t)(v & 0xffff));
                s[1] += min(w, v >> 16);
            }
        } else {
            if (kPacked) {
                const bm_s2 wp = {(short)w, (short)w};
                acc[0] += pk_min(v.x, wp);
                acc[1] += pk_min(v.y, wp);
            } else {
                s[0] += min(w, (int)(int16_t)(v.x & 0xffff));
                s[1] += min(w, v.x >> 16);
                s[2] += min(w, (int)(int16_t)(v.y & 0xffff));
                s[3] += min(w, v.y >> 16);
            }
        }
    };
    auto widen = [&]() {
        if (kPacked && kCols >= 2) {
            s[0] += acc[0].x;
            s[1] += acc[0].y;
            s[2] += acc[1].x;
            s[3] += acc[1].y;
            acc[0] = bm_u2{0, 0};
            acc[1] = bm_u2{0, 0};
        }
    };
    for (int i = 0; i < kRowsWave; i++) {
        const int a = rb * kRowsWG + wave * kRowsWave + i;
        if (a >= N1) break;
        const int beg = cptr[a], end = cptr[a + 1];
        s[0] = s[1] = s[2] = s[3] = 0;
        acc[0] = bm_u2{0, 0};
        acc[1] = bm_u2{0, 0};
        // flush is a multiple of kGainUnroll, so widening falls on block edges
        int since = 0;
        int e = beg;
        if (e + kGainUnroll <= end) {
            int2 rw[kGainUnroll];
#pragma unroll
            for (int u = 0; u < kGainUnroll; u++) rw[u] = crw[e + u];
            for (;;) {
                LoadT v[kGainUnroll];
#pragma unroll
                for (int u = 0; u < kGainUnroll; u++) v[u] = load(rw[u].x);
                int w[kGainUnroll];
#pragma unroll
                for (int u = 0; u < kGainUnroll; u++) w[u] = rw[u].y;
                e += kGainUnroll;
                const bool more = e + kGainUnroll <= end;
                if (more) {  // next block's entries while this block's loads are in flight
#pragma unroll
                    for (int u = 0; u < kGainUnroll; u++) rw[u] = crw[e + u];
                }
#pragma unroll
                for (int u = 0; u < kGainUnroll; u++) consume(v[u], w[u]);
                since += kGainUnroll;
                if (kPacked && since >= flush) {
                    widen();
                    since = 0;
                }
                if (!more) break;
            }
        }
        for (; e < end; e++) {
            const int2 rw = crw[e];
            consume(load(rw.x), rw.y);
        }
        widen();
        if (live) {
            int32_t *dst = X + xidx(a, col, (int)(ldx >> 6));
            if constexpr (kCols == 4) {
                const bm_i4 out = {s[0], s[1], s[2], s[3]};
                __builtin_nontemporal_store(out, reinterpret_cast<bm_i4 *>(dst));
            } else if constexpr (kCols == 2) {
                const bm_i2 out = {s[0], s[1]};
                __builtin_nontemporal_store(out, reinterpret_cast<bm_i2 *>(dst));
            } else {
                __builtin_nontemporal_store(s[0], dst);
            }
        }
    }
}

// X (tile-major, row stride ld) of the lists (ptr, rw) for rows 0..N1-1;
// max_w bounds the weights and nonneg says all are >= 0.
template <int kCols>
static int launch_gain_on(bm_ctx *c, const int32_t *ptr, const int32_t *rw, int32_t *X, int max_w, bool nonneg) {
    const int N1 = c->N1;
    const int nslab = (c->ld + 64 * kCols - 1) / (64 * kCols), nrb = (N1 + kRowsWG - 1) / kRowsWG;
    const int per_xcd = (int)(((int64_t)nslab * nrb + 7) / 8);
    // packed uint16 partial sums need 0 <= min(w, P) and flush * max(w) < 65536
    // (the tail after the last full block adds < kGainUnroll more entries)
    const int flush = max_w > 0 ? ((65535 / max_w) - kGainUnroll) & ~(kGainUnroll - 1) : 1 << 20;
    if (kCols > 1 && nonneg && flush >= kGainUnroll) {
        BM_LAUNCH(c, "gain", hipLaunchKernelGGL((k_gain<kCols, true>), dim3(8 * per_xcd), dim3(256), 0, c->stream,
                                                ptr, (const int2 *)rw, c->Q, c->ld, X, (int64_t)c->ld, N1, nslab,
                                                nrb, per_xcd, flush));
    } else {
        BM_LAUNCH(c, "gain", hipLaunchKernelGGL((k_gain<kCols, false>), dim3(8 * per_xcd), dim3(256), 0, c->stream,
                                                ptr, (const int2 *)rw, c->Q, c->ld, X, (int64_t)c->ld, N1, nslab,
                                                nrb, per_xcd, flush));
    }
    return BM_OK;
}

template <int kCols>
static int launch_gain(bm_ctx *c) {
    return launch_gain_on<kCols>(c, c->cptr, c->crw, c->X, c->max_w1, c->nonneg);
}

static int launch_gain_cols(bm_ctx *c, int cols) {
    return cols == 1 ? launch_gain<1>(c) : cols == 2 ? launch_gain<2>(c) : launch_gain<4>(c);
}

int bm::launch_list_gain(bm_ctx *c, const int32_t *ptr, const int32_t *rw, int32_t *X, int max_w) {
    const int cols = c->gain_cols ? c->gain_cols : 2;
    const bool nonneg = c->g2nonneg;  // min(w, P) >= 0 needs P >= 0 as well
    return cols == 1   ? launch_gain_on<1>(c, ptr, rw, X, max_w, nonneg)
           : cols == 2 ? launch_gain_on<2>(c, ptr, rw, X, max_w, nonneg)
                       : launch_gain_on<4>(c, ptr, rw, X, max_w, nonneg);
}

// Slab-width autotuner: on the first sweep of a graph pair, time one launch
// of each width with HIP events on the context stream and keep the fastest
// (all widths produce identical X).  BM_GAIN_COLS = 1, 2 or 4 forces one.
static int tune_gain(bm_ctx *c) {
    if (c->gain_cols) return BM_OK;
    const char *force = getenv("BM_GAIN_COLS");
    if (force && (atoi(force) == 1 || atoi(force) == 2 || atoi(force) == 4)) {
        c->gain_cols = atoi(force);
        if (!getenv("BM_SWEEP_KERNEL")) c->sweep_kind = 1;  // a forced width means the dense kernel
        return BM_OK;
    }
    hipEvent_t ev[4];
    for (auto &e : ev) BM_HIP(hipEventCreate(&e));
    const int widths[3] = {1, 2, 4};
    float best = 1e30f;
    int best_w = 4;
    const bool was = c->timing;
    c->timing = false;
    for (int i = 0; i < 3; i++) {
        BM_HIP(hipEventRecord(ev[0], c->stream));
        int rc = launch_gain_cols(c, widths[i]);
        if (rc) return rc;
        BM_HIP(hipEventRecord(ev[1], c->stream));
        BM_HIP(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        BM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        c->gain_ms[i] = ms;
        if (ms < best) {
            best = ms;
            best_w = widths[i];
        }
    }
    c->timing = was;
    for (auto &e : ev) (void)hipEventDestroy(e);
    c->gain_cols = best_w;
    return BM_OK;
}

// Sparse form of the same row: P = W2 permuted is as sparse as G2, so
//   X(a,b) = C_a + sum over comb entries e of a and non-zero cells Q[r_e][b]
//            of (min(w_e, Q[r_e][b]) - min(w_e, 0)),   C_a = sum_e min(w_e, 0).
// The non-zero cells of row s of P are the G2 out-list of p(s) and those of
// row d of P^T the G2 in-list of p(d), with targets taken back through p^-1
// (the lists are kept translated, k_g2_to_p).  One 1024-thread workgroup per
// row a accumulates into an (n+1)-int LDS row.  The row's comb entries are
// taken in chunks; per chunk the G2 lists they select are laid end to end
// (prefix sums in LDS) and the 16 waves split that flat index space evenly,
// lanes of a wave reading consecutive 16-byte quads of 4 list words
// ((target << 16 | weight) words, coalesced) and ds_add-ing each word's
// contribution.  Work is the sum over G1 lines of the G2 degree instead of
// 2E * n.  Exact when p is injective on mapped vertices and no G2 line names
// vertex 0 (W2 row/column 0 empty); the caller checks that.
constexpr int kSparseChunk = 512;
constexpr int kSplitEntries = 2048;  // comb entries per block of the sparse kernel, at most
// Long G2 lists streamed as whole Q rows, per chunk of entries.  Kept small:
// two blocks per CU need the static LDS plus 4 (n+1) bytes under 80 KiB.
constexpr int kSparseHubs = 32;
// the sparse kernel is taken without timing when its list work is below 1/8 of the dense cell work
constexpr long long kSparseMargin = 8;

__global__ void k_g2_to_p(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src,
                          const int32_t *__restrict__ pinv, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t gv = src[i];
    const int y = pinv[gv >> 16];
    dst[i] = ((uint32_t)(y < 0 ? 0xFFFF : y) << 16) | (gv & 0xFFFFu);
}

// G2 lists with targets in P coordinates, rebuilt after the mapping changed
int bm::launch_refresh_g2y(bm_ctx *c) {
    if (c->ng2 > 0) {
        const unsigned blocks = (unsigned)((c->ng2 + 255) / 256);
        BM_LAUNCH(c, "g2_to_p", hipLaunchKernelGGL(k_g2_to_p, dim3(blocks), dim3(256), 0, c->stream, c->g2oy, c->g2o,
                                                   c->pinv, c->ng2));
        BM_LAUNCH(c, "g2_to_p", hipLaunchKernelGGL(k_g2_to_p, dim3(blocks), dim3(256), 0, c->stream, c->g2iy, c->g2i,
                                                   c->pinv, c->ng2));
    }
    c->g2y_dirty = false;
    return BM_OK;
}

// One block per work item (a, lo, hi, flags) of the row plan: entries
// [lo, hi) of comb(a), with the pair corrections those entries carry.  A row
// with one item (flags & kRowWhole) is finished here: X(a,a) into diag, the
// row stored.  The items of a split row add their partial rows into X (zeroed
// beforehand) with atomics; the first of them (flags & kRowFirst) adds cneg,
// and X(a,a) becomes diag[a] afterwards (k_diag_from_x).
constexpr int kRowFirst = 1, kRowWhole = 2;

// The translated lists copied into their quads (the padding words, N1 << 16,
// are written once when the quads are laid out): one thread per list word,
// an untranslated word (0xFFFF) becoming N1 << 16 too, a term
// min(w, 0) - min(w, 0) = 0 (or min(w, 0) = 0 when no G1 weight is negative)
// added to the spare accumulator slot N1.  Also P's diagonal.
__global__ __launch_bounds__(256) void k_pack_quads(uint32_t *__restrict__ g2q, const uint32_t *__restrict__ g2y,
                                                    const int32_t *__restrict__ pos, int64_t m,
                                                    const int16_t *__restrict__ Q, int ld, int32_t *__restrict__ pdg,
                                                    int N1, const int *__restrict__ run_flag) {
    if (run_flag && *run_flag == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < max(m, (int64_t)N1);
         i += (int64_t)gridDim.x * blockDim.x) {  // grid-stride: few blocks to retire when X is current
        if (i < N1) pdg[i] = Q[i * ld + i];  // P[i][i], for the pair corrections
        if (i < m) {
            const uint32_t w = g2y[i];
            g2q[pos[i]] = (w >> 16) == 0xFFFFu ? (uint32_t)N1 << 16 : w;
        }
    }
}

// The sparse row over the quad copy of the lists.  The flat index space of a
// chunk counts quads and a lane loads one 16-byte quad (4 list words) per
// step, so the entry bookkeeping (is the lane past its entry's end, the list
// start) and the address are paid once per 4 words.  The row accumulator is
// the first thing in LDS (byte address 0: the kernel has no static LDS), so
// under kSmall (every list weight in [0, 16383], whose word then has zero bits
// 14-15) the byte address of target y is the word shifted right by 14.  No
// word is skipped: untranslated targets were made zero terms by k_pack_quads.
// Each row walks its entries in increasing neighbour id (crws), one order for
// all rows, so that the workgroups resident on an XCD read the same G2 lists
// at about the same time (L2 hits 46 -> 53 %).  Scalar registers are capped:
// above about 72 the GPU holds one 1024-thread workgroup per CU instead of two
// (measured: 1.84 ms against 1.40 ms at n = 18,524), although the compiler's
// occupancy estimate is 8 waves per SIMD for both.
constexpr int kQuadUnroll = 2;  // quads in flight per lane (1: 1.61 ms, 2: 1.58, 3: 1.59, 4: 1.60 at C2)

static size_t sparse_lds_bytes(int N1) {
    // acc[N1 + 8 rounded to 4], pref[kSparseChunk + 2], lw[kSparseChunk] (int2),
    // hubr[kSparseHubs], hubw[kSparseHubs], wsum[16], nhub (+3 spare ints)
    return ((size_t)((N1 + 8 + 3) & ~3) + (kSparseChunk + 2) + 2 * kSparseChunk + 2 * kSparseHubs + 16 + 4) * 4;
}

template <bool kHubs, bool kNeg, bool kSmall>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_num_sgpr(80))) void k_gain_sparse(
    const int4 *__restrict__ plan, const int2 *__restrict__ crws, const int32_t *__restrict__ cdws,
    const int32_t *__restrict__ pdg, const int32_t *__restrict__ cnegv, const int32_t *__restrict__ map,
    const int32_t *__restrict__ qoptr, const int32_t *__restrict__ qiptr, const uint4 *__restrict__ g2q,
    const int16_t *__restrict__ Q, int ld, int32_t *__restrict__ X, int64_t ldx, int32_t *__restrict__ diag, int N1,
    int dense_len, const int *__restrict__ run_flag) {
    if (run_flag && *run_flag == 0) return;  // X already current (the batched step's incremental update)
    extern __shared__ __align__(16) int smem[];
    int *acc = smem;  // LDS byte address 0
    int *pref = smem + ((N1 + 8 + 3) & ~3);
    int2 *lw = (int2 *)(pref + kSparseChunk + 2);  // per entry: first quad of its list, weight
    int *hubr = (int *)(lw + kSparseChunk), *hubw = hubr + kSparseHubs;
    int *wsum = hubw + kSparseHubs;
    int *nhub = wsum + 16;
    const int4 item = plan[blockIdx.x];
    const int a = item.x, tid = threadIdx.x, flags = item.w;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    for (int i = tid; i < N1; i += 1024) acc[i] = 0;
    const int paa = (int)Q[(size_t)a * ld + a];
    // one list word: f(w, q) = min(w, q) - min(w, 0) into the target's slot
    auto add = [&](uint32_t yv, int w) {
        const int v = min(w, (int)(int16_t)(yv & 0xffff)) - (kNeg ? min(w, 0) : 0);
        if (kSmall)  // the LDS address itself: smem starts at byte 0
            __hip_atomic_fetch_add((__attribute__((address_space(3))) int *)(size_t)(yv >> 14), v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            atomicAdd(&acc[yv >> 16], v);
    };
    const int beg = item.y, end = item.z;
    for (int c0 = beg; c0 < end; c0 += kSparseChunk) {
        const int m = min(kSparseChunk, end - c0);
        if (tid == 0) *nhub = 0;
        __syncthreads();
        int len = 0;  // in quads
        if (tid < m) {
            const int2 rw = crws[c0 + tid];
            const bool out = rw.x >= N1;  // row of P^T (out entry a->d: the in-list of p(d)) or of P
            const int b = out ? rw.x - N1 : rw.x;
            // Corr(a, b) for b > a from this entry (each G1 line between a and b is one
            // entry of comb(a)); P[b][a] read as P^T[a][b], so the reads stay in two rows.
            // Its loads go first, so that they overlap the list lookups below.
            int pab = 0, pba = 0, pbb = 0, dw = 0;
            if (b > a) {
                pab = (int)Q[(size_t)a * ld + b];
                pba = (int)Q[(size_t)(N1 + a) * ld + b];
                pbb = pdg[b];
                dw = cdws[c0 + tid];
            }
            const int g = map[b];
            const int lo = out ? qiptr[g] : qoptr[g];
            len = (out ? qiptr[g + 1] : qoptr[g + 1]) - lo;
            lw[tid] = make_int2(lo, rw.y);
            if (kHubs && 4 * len > dense_len) {
                const int h = atomicAdd(nhub, 1);
                if (h < kSparseHubs) {  // beyond the cap the list is walked like the others
                    hubr[h] = rw.x;
                    hubw[h] = rw.y;
                    len = 0;
                }
            }
            if (b > a) atomicAdd(&acc[b], pair_corr(out, rw.y, dw, paa, pbb, pab, pba));
        }
        int x = len;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        if (tid < 16) {
            int v = wsum[tid];
            for (int o = 1; o < 16; o <<= 1) {
                const int y = __shfl_up(v, o, 16);
                if (tid >= o) v += y;
            }
            wsum[tid] = v;
        }
        __syncthreads();
        if (tid < m) pref[tid + 1] = x + (wave ? wsum[wave - 1] : 0);
        if (tid == 0) pref[0] = 0;
        __syncthreads();
        const int total = pref[m];
        const int span = (total + 15) / 16;
        const int w0 = wave * span, w1 = min(total, w0 + span);
        if (w0 < w1) {
            int lo_i = 0, hi_i = m - 1;  // entry holding flat quad w0
            while (lo_i < hi_i) {
                const int mid = (lo_i + hi_i + 1) >> 1;
                if (pref[mid] <= w0) lo_i = mid;
                else hi_i = mid - 1;
            }
            int i0 = lo_i, s0 = pref[i0], s1 = pref[i0 + 1];
            int2 cur = lw[i0];
            int q = w0 + lane;
            for (; q + 64 * (kQuadUnroll - 1) < w1; q += 64 * kQuadUnroll) {
                int at[kQuadUnroll], wu[kQuadUnroll];
#pragma unroll
                for (int u = 0; u < kQuadUnroll; u++) {
                    while (s1 <= q + 64 * u) {
                        i0++;
                        s0 = s1;
                        s1 = pref[i0 + 1];
                        cur = lw[i0];
                    }
                    at[u] = cur.x - s0;  // flat quad q + 64u is g2q[at + q + 64u]
                    wu[u] = cur.y;
                }
                uint4 v[kQuadUnroll];
#pragma unroll
                for (int u = 0; u < kQuadUnroll; u++) v[u] = g2q[(uint32_t)(at[u] + q) + 64u * u];
#pragma unroll
                for (int u = 0; u < kQuadUnroll; u++) {
                    add(v[u].x, wu[u]);
                    add(v[u].y, wu[u]);
                    add(v[u].z, wu[u]);
                    add(v[u].w, wu[u]);
                }
            }
            for (; q < w1; q += 64) {
                while (s1 <= q) {
                    i0++;
                    s0 = s1;
                    s1 = pref[i0 + 1];
                    cur = lw[i0];
                }
                const uint4 v = g2q[cur.x - s0 + q];
                add(v.x, cur.y);
                add(v.y, cur.y);
                add(v.z, cur.y);
                add(v.w, cur.y);
            }
        }
        __syncthreads();
        // long lists: whole Q rows, up to four folded per pass; each lane owns an adjacent
        // pair of columns (64-bit LDS accesses)
        const int nh = kHubs ? min(*nhub, kSparseHubs) : 0;
        for (int h0 = 0; h0 < nh; h0 += 4) {
            const int g = min(4, nh - h0);
            const int16_t *row[4];
            int w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                row[u] = Q + (size_t)hubr[h0 + min(u, g - 1)] * ld;
                w[u] = u < g ? hubw[h0 + u] : 0;
            }
            for (int y = tid * 2; y < N1; y += 2048) {
                int2 v = *(int2 *)&acc[y];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (u < g) {
                        const bm_s2 qq = *(const bm_s2 *)(row[u] + y);
                        const int base = -min(w[u], 0);
                        v.x += min(w[u], (int)qq[0]) + base;
                        v.y += min(w[u], (int)qq[1]) + base;
                    }
                }
                *(int2 *)&acc[y] = v;  // the spare slot N1 absorbs y + 1 = N1
            }
        }
        __syncthreads();
    }
    __syncthreads();  // also orders the zeroing of acc when the item has no entries
    // X(a,a) (no correction lands on slot a)
    const int ca = cnegv[a];
    if ((flags & kRowWhole) && tid == 0) diag[a] = acc[a] + ca;
    const int T = (int)(ldx >> 6);
    if (flags & kRowWhole) {
        for (int i = tid; i < ldx; i += 1024) __builtin_nontemporal_store(i < N1 ? acc[i] + ca : 0, X + xidx(a, i, T));
    } else {
        const int add0 = (flags & kRowFirst) ? ca : 0;
        for (int i = tid; i < N1; i += 1024) {
            const int v = acc[i] + add0;
            if (v) atomicAdd(X + xidx(a, i, T), v);
        }
    }
}

// The largest LDS row the sparse kernel may use (bytes).
constexpr size_t kSparseLdsMax = 144 * 1024 - 32;

// The sparse kernel's preconditions: the mapping (sparse_ok), the LDS row, and
// list offsets that fit the 32-bit byte offsets of its buffer loads.
static bool sparse_fits(const bm_ctx *c) {
    return c->sparse_ok && (size_t)c->N1 * 4 <= kSparseLdsMax && (uint64_t)c->ng2 * 8 < ((uint64_t)1 << 31);
}

// rows split over several plan items start from zero
__global__ void k_zero_rows(int32_t *__restrict__ X, int64_t ldx, const int32_t *__restrict__ rows,
                            const int *__restrict__ run_flag) {
    if (run_flag && *run_flag == 0) return;
    const int a = rows[blockIdx.x];
    const int T = (int)(ldx >> 6);
    for (int i = threadIdx.x; i < ldx; i += blockDim.x) X[xidx(a, i, T)] = 0;
}

__global__ void k_diag_from_x(const int32_t *__restrict__ X, int64_t ldx, const int32_t *__restrict__ rows, int count,
                              int32_t *__restrict__ diag, const int *__restrict__ run_flag) {
    if (run_flag && *run_flag == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int a = rows[i];
    diag[a] = X[xidx(a, a, (int)(ldx >> 6))];
}

// The row plan of the sparse kernel: rows with more than kSplitEntries comb
// entries become ceil(deg / kSplitEntries) items; items ordered by entry count,
// largest first, so that the longest blocks start first.
// The quad copy's layout, built on the device: every G2 word's place (word k
// of out-list g at 4 qo[g] + (k - optr[g]), in-list words after them), the
// padding words (N1 << 16, zero terms at the spare slot), and the comb
// entries of each row in increasing neighbour id (crws, cdws: a radix sort of
// (row << 16 | neighbour) keys; ties in any order, the row sums do not depend
// on it).
__global__ void k_quad_pos(int32_t *__restrict__ pos, const int32_t *__restrict__ optr,
                           const int32_t *__restrict__ iptr, const int32_t *__restrict__ qo,
                           const int32_t *__restrict__ qi, int N1, int64_t ng2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ng2) return;
    const bool in = t >= ng2;
    const int k = (int)(in ? t - ng2 : t);
    const int32_t *ptr = in ? iptr : optr;
    int lo = 0, hi = N1 - 1;  // the list holding word k: the last g with ptr[g] <= k
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ptr[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    pos[t] = 4 * (in ? qi[lo] : qo[lo]) + (k - ptr[lo]);
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, int64_t count, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// one wave per row: key (row << 16 | neighbour) and index of each comb entry
__global__ void k_row_nb_keys(uint32_t *__restrict__ keys, int32_t *__restrict__ vals,
                              const int32_t *__restrict__ cptr, const int2 *__restrict__ crw, int N1) {
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= N1) return;
    for (int e = cptr[a] + (threadIdx.x & 63); e < cptr[a + 1]; e += 64) {
        const int r = crw[e].x;
        keys[e] = ((uint32_t)a << 16) | (uint32_t)(r >= N1 ? r - N1 : r);
        vals[e] = e;
    }
}

__global__ void k_gather_entries(int2 *__restrict__ rws, int32_t *__restrict__ dws, const int2 *__restrict__ rw,
                                 const int32_t *__restrict__ dw, const int32_t *__restrict__ idx, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int32_t j = idx[i];
    rws[i] = rw[j];
    dws[i] = dw[j];
}

static int build_quads(bm_ctx *c) {
    if (c->g2q) return BM_OK;
    const int N1 = c->N1;
    std::vector<int32_t> qo(N1 + 1, 0), qi(N1 + 1, 0);
    for (int g = 0; g < N1; g++) qo[g + 1] = qo[g] + (c->hg2optr[g + 1] - c->hg2optr[g] + 3) / 4;
    qi[0] = qo[N1];
    for (int g = 0; g < N1; g++) qi[g + 1] = qi[g] + (c->hg2iptr[g + 1] - c->hg2iptr[g] + 3) / 4;
    int rc;
    const int64_t nwords = 2 * c->ng2;
    if (!c->qoptr && (rc = dev_alloc(c, (void **)&c->qoptr, (size_t)(N1 + 1) * 4))) return rc;
    if (!c->qiptr && (rc = dev_alloc(c, (void **)&c->qiptr, (size_t)(N1 + 1) * 4))) return rc;
    if (!c->g2qpos && (rc = dev_alloc(c, (void **)&c->g2qpos, (size_t)nwords * 4 + 4))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->g2q, (size_t)qi[N1] * 16))) return rc;
    c->nquads = qi[N1];
    BM_HIP(hipMemcpyAsync(c->qoptr, qo.data(), (size_t)(N1 + 1) * 4, hipMemcpyHostToDevice, c->stream));
    BM_HIP(hipMemcpyAsync(c->qiptr, qi.data(), (size_t)(N1 + 1) * 4, hipMemcpyHostToDevice, c->stream));
    if (nwords > 0)
        BM_LAUNCH(c, "quad_pos", hipLaunchKernelGGL(k_quad_pos, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0,
                                                    c->stream, c->g2qpos, c->g2optr, c->g2iptr, c->qoptr, c->qiptr,
                                                    N1, c->ng2));
    if (qi[N1] > 0)
        BM_LAUNCH(c, "quad_pad", hipLaunchKernelGGL(k_fill_u32, dim3(2048), dim3(256), 0, c->stream, c->g2q,
                                                    (int64_t)qi[N1] * 4, (uint32_t)N1 << 16));
    const size_t nc = (size_t)c->ncomb;
    if (!c->crws && (rc = dev_alloc(c, (void **)&c->crws, nc * 8 + 8))) return rc;
    if (!c->cdws && (rc = dev_alloc(c, (void **)&c->cdws, nc * 4 + 4))) return rc;
    if (!c->pdg && (rc = dev_alloc(c, (void **)&c->pdg, (size_t)N1 * 4))) return rc;
    if (nc > 0) {
        uint32_t *k0 = nullptr, *k1 = nullptr;
        int32_t *v0 = nullptr, *v1 = nullptr;
        void *tmp = nullptr;
        size_t tmp_bytes = 0;
        int end_bit = 16;
        while ((1 << (end_bit - 16)) < N1) end_bit++;
        BM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, v0, v1, (int)nc, 0, end_bit,
                                                  c->stream));
        if ((rc = dev_alloc(c, (void **)&k0, nc * 4)) || (rc = dev_alloc(c, (void **)&k1, nc * 4)) ||
            (rc = dev_alloc(c, (void **)&v0, nc * 4)) || (rc = dev_alloc(c, (void **)&v1, nc * 4)) ||
            (rc = dev_alloc(c, &tmp, tmp_bytes))) {
            dev_free(c, (void **)&k0, nc * 4), dev_free(c, (void **)&k1, nc * 4), dev_free(c, (void **)&v0, nc * 4),
                dev_free(c, (void **)&v1, nc * 4), dev_free(c, &tmp, tmp_bytes);
            return rc;
        }
        BM_LAUNCH(c, "row_nb_keys", hipLaunchKernelGGL(k_row_nb_keys, dim3((N1 + 3) / 4), dim3(256), 0, c->stream,
                                                       k0, v0, c->cptr, (const int2 *)c->crw, N1));
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, v0, v1, (int)nc, 0, end_bit,
                                                          c->stream);
        if (e == hipSuccess)
            hipLaunchKernelGGL(k_gather_entries, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream,
                               (int2 *)c->crws, c->cdws, (const int2 *)c->crw, c->cdw, v1, (int64_t)nc);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // before the buffers are freed
        dev_free(c, (void **)&k0, nc * 4), dev_free(c, (void **)&k1, nc * 4), dev_free(c, (void **)&v0, nc * 4),
            dev_free(c, (void **)&v1, nc * 4), dev_free(c, &tmp, tmp_bytes);
        if (e != hipSuccess) return hip_fail(e, "neighbour-ordered comb lists");
    }
    return BM_OK;
}

static int build_row_plan(bm_ctx *c) {
    int rc0 = build_quads(c);
    if (rc0) return rc0;
    if (c->plan) return BM_OK;
    int split = kSplitEntries;
    if (const char *e = getenv("BM_SPARSE_SPLIT")) split = std::max(4, atoi(e));
    // G2 lists longer than N1 / 2 stream their Q row densely (BM_SPARSE_DENSE_LEN overrides)
    c->dense_len = std::max(64, c->N1 / 2);
    if (const char *e = getenv("BM_SPARSE_DENSE_LEN")) c->dense_len = std::max(0, atoi(e));
    std::vector<int4> items;
    std::vector<int32_t> rows;
    std::vector<int32_t> cp(c->N1 + 1, 0);
    for (int v = 0; v < c->N1; v++) cp[v + 1] = cp[v] + c->hcdeg[v];
    for (int a = 0; a < c->N1; a++) {
        const int deg = c->hcdeg[a];
        if (deg <= split) {
            items.push_back(make_int4(a, cp[a], cp[a + 1], kRowFirst | kRowWhole));
            continue;
        }
        rows.push_back(a);
        for (int lo = cp[a]; lo < cp[a + 1]; lo += split)
            items.push_back(make_int4(a, lo, std::min(cp[a + 1], lo + split), lo == cp[a] ? kRowFirst : 0));
    }
    std::stable_sort(items.begin(), items.end(), [](const int4 &x, const int4 &y) { return x.z - x.y > y.z - y.y; });
    int rc;
    if ((rc = dev_alloc(c, (void **)&c->plan, items.size() * sizeof(int4)))) return rc;
    c->nplan = (int64_t)items.size();
    if ((rc = copy_sync(c->plan, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice))) return rc;
    c->nsplit = (int)rows.size();
    if (!rows.empty()) {
        if ((rc = dev_alloc(c, (void **)&c->split_rows, rows.size() * 4))) return rc;
        if ((rc = copy_sync(c->split_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice))) return rc;
    }
    return BM_OK;
}

static int launch_gain_sparse(bm_ctx *c, const int *run_flag = nullptr) {
    if (c->g2y_dirty) {
        int rc = launch_refresh_g2y(c);
        if (rc) return rc;
    }
    int rc = build_row_plan(c);
    if (rc) return rc;
    if (!c->sparse_attr) {  // once per context, so that captured launches make no attribute calls
        const void *fns[8] = {(const void *)k_gain_sparse<true, true, false>,
                              (const void *)k_gain_sparse<true, false, false>,
                              (const void *)k_gain_sparse<false, true, false>,
                              (const void *)k_gain_sparse<false, false, false>,
                              (const void *)k_gain_sparse<true, true, true>,
                              (const void *)k_gain_sparse<true, false, true>,
                              (const void *)k_gain_sparse<false, true, true>,
                              (const void *)k_gain_sparse<false, false, true>};
        for (const void *f : fns)
            BM_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)sparse_lds_bytes((int)(kSparseLdsMax / 4))));
        c->sparse_attr = true;
    }
    if (c->nsplit > 0)
        BM_LAUNCH(c, "zero_rows", hipLaunchKernelGGL(k_zero_rows, dim3(c->nsplit), dim3(256), 0, c->stream, c->X,
                                                     c->ldx, c->split_rows, run_flag));
    {
        const int64_t m = 2 * c->ng2, threads = std::max<int64_t>(m, c->N1);
        BM_LAUNCH(c, "pack_quads",
                  hipLaunchKernelGGL(k_pack_quads, dim3((unsigned)std::min<int64_t>((threads + 255) / 256, 2048)),
                                     dim3(256), 0, c->stream,
                                     c->g2q, c->g2oy, c->g2qpos, m, c->Q, c->ld, c->pdg, c->N1, run_flag));
    }
    // the long-list path costs registers and scheduling even when idle: only
    // the instantiation that needs it carries it; likewise the min(w, 0) term
    const bool hubs = c->g2_maxlen > c->dense_len, neg = !c->g1nonneg;
    const size_t lds = sparse_lds_bytes(c->N1);
#define BM_GAIN_SPARSE(H, NEG, SMALL)                                                                                  \
    BM_LAUNCH(c, "gain", hipLaunchKernelGGL((k_gain_sparse<H, NEG, SMALL>), dim3((unsigned)c->nplan), dim3(1024), lds, \
                                            c->stream, c->plan, (const int2 *)c->crws, c->cdws, c->pdg, c->cneg,       \
                                            c->map, c->qoptr, c->qiptr, (const uint4 *)c->g2q, c->Q, c->ld, c->X,      \
                                            c->ldx, c->diag, c->N1, c->dense_len, run_flag))
    if (c->g2q_small) {
        if (hubs && neg) BM_GAIN_SPARSE(true, true, true);
        else if (hubs) BM_GAIN_SPARSE(true, false, true);
        else if (neg) BM_GAIN_SPARSE(false, true, true);
        else BM_GAIN_SPARSE(false, false, true);
    } else {
        if (hubs && neg) BM_GAIN_SPARSE(true, true, false);
        else if (hubs) BM_GAIN_SPARSE(true, false, false);
        else if (neg) BM_GAIN_SPARSE(false, true, false);
        else BM_GAIN_SPARSE(false, false, false);
    }
#undef BM_GAIN_SPARSE
    if (c->nsplit > 0)
        BM_LAUNCH(c, "diag_from_x", hipLaunchKernelGGL(k_diag_from_x, dim3((c->nsplit + 255) / 256), dim3(256), 0,
                                                       c->stream, c->X, c->ldx, c->split_rows, c->nsplit, c->diag,
                                                       run_flag));
    c->diag_valid = true;
    return BM_OK;
}

// Chooses the all-pairs gain kernel on the first sweep of a graph pair: the
// sparse kernel when its preconditions hold, timed against the best dense
// slab width, the faster one kept (both produce identical X).
// BM_SWEEP_KERNEL = dense | sparse forces one.
// Choices per class of instance (log2 n, log2 mean comb length, sparse
// possible), shared by every context of the process: later graph pairs of a
// class skip the timing runs.
static std::mutex g_gain_mu;
static std::map<std::tuple<int, int, int>, std::pair<int, int>> g_gain_cache;

static int ilog2(long long x) {
    int r = 0;
    while ((1ll << r) < x) r++;
    return r;
}

static int choose_gain_timed(bm_ctx *c);

static int choose_gain(bm_ctx *c) {
    if (c->gain_cols && (c->sweep_kind || !(c->sparse_ok && (size_t)c->N1 * 4 <= kSparseLdsMax))) return BM_OK;
    const bool can_sparse = sparse_fits(c);
    const bool forced = getenv("BM_GAIN_COLS") || getenv("BM_SWEEP_KERNEL");
    const auto key = std::make_tuple(ilog2(c->N1), ilog2(std::max<long long>(1, c->ncomb / c->N1)), can_sparse ? 1 : 0);
    if (!forced) {
        std::lock_guard<std::mutex> g(g_gain_mu);
        auto it = g_gain_cache.find(key);
        if (it != g_gain_cache.end()) {
            c->gain_cols = it->second.second;
            if (can_sparse) c->sweep_kind = it->second.first;
            return BM_OK;
        }
    }
    int rc = choose_gain_timed(c);
    if (rc) return rc;
    if (!forced) {
        std::lock_guard<std::mutex> g(g_gain_mu);
        g_gain_cache[key] = std::make_pair(c->sweep_kind, c->gain_cols);
    }
    return BM_OK;
}

static int choose_gain_timed(bm_ctx *c) {
    const bool can_sparse = sparse_fits(c);
    if (can_sparse && !getenv("BM_SWEEP_KERNEL") && !getenv("BM_GAIN_COLS")) {
        // By the work counts: the sparse kernel walks about sum_v deg1(v) deg2(p(v)) list
        // words, the dense ones ncomb * (n+1) cells.  A clear margin decides without timing
        // (timings taken while other contexts load the device can mislead).
        long long ws = 0;
        for (int v = 0; v < c->N1; v++) ws += (long long)c->hcdeg[v] * c->hg2deg[c->hmap[v]];
        if (ws * kSparseMargin < (long long)c->ncomb * c->N1) {
            c->sweep_kind = 2;
            c->gain_cols = 4;  // unused while the sparse kernel runs
            return build_row_plan(c);
        }
    }
    int rc = tune_gain(c);
    if (rc) return rc;
    if (!can_sparse) return BM_OK;
    if ((rc = build_row_plan(c))) return rc;  // its host copies stay outside the timed launch
    if (c->sweep_kind) return BM_OK;
    const char *force = getenv("BM_SWEEP_KERNEL");
    if (force && !strcmp(force, "dense")) {
        c->sweep_kind = 1;
        return BM_OK;
    }
    if (force && !strcmp(force, "sparse")) {
        c->sweep_kind = 2;
        return BM_OK;
    }
    hipEvent_t ev[2];
    for (auto &e : ev) BM_HIP(hipEventCreate(&e));
    const bool was = c->timing;
    c->timing = false;
    BM_HIP(hipEventRecord(ev[0], c->stream));
    if ((rc = launch_gain_sparse(c))) return rc;
    BM_HIP(hipEventRecord(ev[1], c->stream));
    BM_HIP(hipEventSynchronize(ev[1]));
    c->timing = was;
    BM_HIP(hipEventElapsedTime(&c->sparse_ms, ev[0], ev[1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    const float dense_ms = c->gain_ms[c->gain_cols == 1 ? 0 : c->gain_cols == 2 ? 1 : 2];
    c->sweep_kind = (dense_ms > 0.f && dense_ms < c->sparse_ms) ? 1 : 2;
    return BM_OK;
}

// Corr(a,b) added into X[min(a,b)][max(a,b)], one thread per G1 line.
__global__ __launch_bounds__(256) void k_corr(const int32_t *__restrict__ es, const int32_t *__restrict__ ed,
                                              const int32_t *__restrict__ ew, const int32_t *__restrict__ edw,
                                              int64_t m, const int16_t *__restrict__ Q, int ld,
                                              int32_t *__restrict__ X, int64_t ldx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int x = es[i], y = ed[i];
    if (x == y) return;
    const int pxx = q16(Q, x, ld, x), pyy = q16(Q, y, ld, y), pxy = q16(Q, x, ld, y), pyx = q16(Q, y, ld, x);
    // the line x->y seen from the pair (a=x, b=y) as an out entry of a
    const int v = pair_corr(true, ew[i], edw[i], pxx, pyy, pxy, pyx);
    const int lo = min(x, y), hi = max(x, y);
    atomicAdd(&X[xidx(lo, hi, (int)(ldx >> 6))], v);
}

// Delta_B(a,b) = X(a,b) + X(b,a) - X(a,a) - X(b,b) for 1 <= a < b <= n, with
// Corr(a,b) already folded into X[a][b].  A block takes one 64-row tile I and
// a chunk of kCombChunk column tiles J >= I.  Each tile pair (I,J), (J,I) is
// two contiguous 16 KB blocks of the tile-major X, staged in LDS; lane l of
// every wave owns row a = 64I + l and wave w the 16 columns 16w..16w+15 of
// the tile, so each lane keeps its row's best (delta, -b) in a register
// across the whole chunk.  At the end the 4 waves' bests of a row meet in
// LDS: one 64-bit atomic max per row per block, and one for the block's
// global best (lexicographic (a, b) tie-break in the key).
constexpr int kCombChunk = 8;

__global__ __launch_bounds__(256) void k_combine(const int32_t *__restrict__ X, int64_t ldx,
                                                 const int32_t *__restrict__ diag, int n,
                                                 unsigned long long *__restrict__ rowkey,
                                                 unsigned long long *__restrict__ gkey,
                                                 int32_t *__restrict__ dmat, const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int I = blockIdx.x;
    const int T = (int)(ldx >> 6), Tn = (n + 64) / 64;
    const int J0 = I + blockIdx.y * kCombChunk, J1 = min(Tn, J0 + kCombChunk);
    if (J0 >= J1) return;
    __shared__ int XA[64][65], XB[64][65];
    __shared__ int dA[64], dB[64];
    __shared__ unsigned long long best[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a0 = I * 64, a = a0 + lane;
    const int N1 = n + 1;
    if (tid < 64) dA[lane] = (a <= n) ? diag[a] : 0;
    unsigned long long kb = 0;
    // the next tile pair's loads are issued before the current pair is scored
    typedef int i4v __attribute__((ext_vector_type(4)));
    i4v va[4], vb[4];
    auto load = [&](int J) {
        const i4v *ta = reinterpret_cast<const i4v *>(X + ((size_t)I * T + J) * 4096);
        const i4v *tb = reinterpret_cast<const i4v *>(X + ((size_t)J * T + I) * 4096);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            va[k] = __builtin_nontemporal_load(ta + tid + 256 * k);
            vb[k] = __builtin_nontemporal_load(tb + tid + 256 * k);
        }
    };
    load(J0);
    for (int J = J0; J < J1; J++) {
        const int b0 = J * 64;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int q = tid + 256 * k;  // int4 index in the 1024-int4 tile
            const int r = q >> 4, c = (q & 15) * 4;
            XA[r][c] = va[k].x; XA[r][c + 1] = va[k].y; XA[r][c + 2] = va[k].z; XA[r][c + 3] = va[k].w;
            XB[r][c] = vb[k].x; XB[r][c + 1] = vb[k].y; XB[r][c + 2] = vb[k].z; XB[r][c + 3] = vb[k].w;
        }
        if (tid < 64) dB[lane] = (b0 + lane <= n) ? diag[b0 + lane] : 0;
        __syncthreads();
        if (J + 1 < J1) load(J + 1);
        if (a >= 1 && a <= n) {
            const int da = dA[lane];
#pragma unroll 4
            for (int j = 0; j < 16; j++) {
                const int bl = wv * 16 + j, b = b0 + bl;
                if (a < b && b <= n) {
                    const int d = XA[lane][bl] + XB[bl][lane] - da - dB[bl];
                    const unsigned long long k =
                        ((unsigned long long)((unsigned)d ^ 0x80000000u) << 32) | (0xFFFFFFFFu - (unsigned)b);
                    kb = k > kb ? k : kb;
                    if (dmat) {
                        dmat[(size_t)a * N1 + b] = d;
                        dmat[(size_t)b * N1 + a] = d;
                    }
                }
            }
        }
        __syncthreads();
    }
    best[wv][lane] = kb;
    __syncthreads();
    if (tid < 64) {
        unsigned long long k = max(max(best[0][lane], best[1][lane]), max(best[2][lane], best[3][lane]));
        unsigned long long g = 0;
        if (k && a <= n) {
            atomicMax(&rowkey[a], k);
            const unsigned bb = 0xFFFFFFFFu - (unsigned)(k & 0xFFFFFFFFu);
            g = (k & 0xFFFFFFFF00000000ull) | (0xFFFFFFFFu - (((unsigned)a << 16) | bb));
        }
        g = wave_max_u64(g);
        if (lane == 0 && g) atomicMax(gkey, g);
    }
}

// ------------------------------------------------------------ C ABI

int bm::launch_eval_pairs(bm_ctx *c, const int32_t *dA, const int32_t *dB, int64_t m, int32_t *dD) {
    if (m <= 0) return BM_OK;
    const unsigned blocks = (unsigned)((m * 16 + 255) / 256);
    BM_LAUNCH(c, "eval_pairs", hipLaunchKernelGGL(k_eval_pairs, dim3(blocks), dim3(256), 0, c->stream, dA, dB, m, dD,
                                                  c->cptr, c->cr, c->cw, c->cdw, c->Q, c->N1, c->ld));
    return BM_OK;
}

static int check_pairs(bm_ctx *c, const int32_t *a, const int32_t *b, int64_t count) {
    for (int64_t i = 0; i < count; i++)
        if (a[i] < 0 || a[i] > c->n || b[i] < 0 || b[i] > c->n || a[i] == b[i])
            return set_error(BM_EINVAL, "pair %lld: (%d,%d) invalid", (long long)i, a[i], b[i]);
    return BM_OK;
}

extern "C" int bm_eval_swaps(bm_ctx *c, const int32_t *a, const int32_t *b, int64_t count, int32_t *delta) {
    if (!c || (count > 0 && (!a || !b || !delta))) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (count < 0) return set_error(BM_EINVAL, "count < 0");
    if (count == 0) return BM_OK;
    int rc = check_pairs(c, a, b, count);
    if (rc) return rc;
    BM_HIP(hipSetDevice(c->device));
    const int64_t chunk = 1 << 22;
    if ((rc = ensure_stage(c, (size_t)chunk * 12))) return rc;
    int32_t *dA = (int32_t *)c->stage, *dB = dA + chunk, *dD = dB + chunk;
    for (int64_t off = 0; off < count; off += chunk) {
        const int64_t m = std::min(chunk, count - off);
        BM_HIP(hipMemcpyAsync(dA, a + off, m * 4, hipMemcpyHostToDevice, c->stream));
        BM_HIP(hipMemcpyAsync(dB, b + off, m * 4, hipMemcpyHostToDevice, c->stream));
        if ((rc = launch_eval_pairs(c, dA, dB, m, dD))) return rc;
        BM_HIP(hipMemcpyAsync(delta + off, dD, m * 4, hipMemcpyDeviceToHost, c->stream));
        BM_HIP(hipStreamSynchronize(c->stream));
    }
    return BM_OK;
}

// ------------------------------------------------------------ no allocation

// bm_eval_swaps without any allocation: the pairs and their deltas are staged
// in a file-scope device array instead of the context's staging buffer, which
// is allocated on first use.  Each device holds its own copy of the array at
// its own address, taken once per device with hipGetSymbolAddress; host code
// uses only that address (the array's name is a host-side stand-in).  Calls
// on one device are serialised here: they share the array, and a call
// returns after its deltas are copied out.
namespace {
constexpr int64_t kNoAllocPairs = 1 << 20;  // pairs per call, at most (12 MiB of scratch)
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) int32_t scratch[3 * kNoAllocPairs];
std::mutex noalloc_mu[kNoAllocDevices];
int32_t *noalloc_ptr[kNoAllocDevices];  // per device, under noalloc_mu[device]
}  // namespace

extern "C" int bm_eval_swaps_noalloc(bm_ctx *c, const int32_t *a, const int32_t *b, int64_t count,
                                     int32_t *delta) {
    if (!c || (count > 0 && (!a || !b || !delta))) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (count < 0) return set_error(BM_EINVAL, "count < 0");
    if (count > kNoAllocPairs)
        return set_error(BM_EINVAL, "bm_eval_swaps_noalloc: %lld pairs exceed the scratch of %lld pairs per call",
                         (long long)count, (long long)kNoAllocPairs);
    if (c->device < 0 || c->device >= kNoAllocDevices) return set_error(BM_ENODEV, "device %d", c->device);
    if (count == 0) return BM_OK;
    int rc = check_pairs(c, a, b, count);
    if (rc) return rc;
    BM_HIP(hipSetDevice(c->device));
    std::lock_guard<std::mutex> lk(noalloc_mu[c->device]);
    if (!noalloc_ptr[c->device]) {  // first call on this device, with the device current
        void *p = nullptr;
        BM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        noalloc_ptr[c->device] = (int32_t *)p;
    }
    int32_t *dA = noalloc_ptr[c->device], *dB = dA + kNoAllocPairs, *dD = dB + kNoAllocPairs;
    BM_HIP(hipMemcpyAsync(dA, a, count * 4, hipMemcpyHostToDevice, c->stream));
    BM_HIP(hipMemcpyAsync(dB, b, count * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_eval_pairs, dim3((unsigned)((count * 16 + 255) / 256)), dim3(256), 0, c->stream, dA, dB,
                       count, dD, c->cptr, c->cr, c->cw, c->cdw, c->Q, c->N1, c->ld);
    if ((rc = launch_check("eval_pairs (no allocation)"))) return rc;
    BM_HIP(hipMemcpyAsync(delta, dD, count * 4, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    return BM_OK;
}

// The dense-matrix delta of quick.cu:48-250 through the same pair kernel over
// the dense-form lists (each entry owns its cell, so the dense-cell weight
// array is the weight array).
extern "C" int bm_eval_swaps_dense(bm_ctx *c, const int32_t *a, const int32_t *b, int64_t count, int32_t *delta) {
    if (!c || (count > 0 && (!a || !b || !delta))) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (count < 0) return set_error(BM_EINVAL, "count < 0");
    for (int64_t i = 0; i < count; i++)
        if (a[i] < 1 || a[i] > c->n || b[i] < 1 || b[i] > c->n || a[i] == b[i])
            return set_error(BM_EINVAL, "pair %lld: (%d,%d) invalid", (long long)i, a[i], b[i]);
    if (count == 0) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    int rc = ensure_dense_lists(c);
    if (rc) return rc;
    const int64_t chunk = 1 << 22;
    if ((rc = ensure_stage(c, (size_t)chunk * 12))) return rc;
    int32_t *dA = (int32_t *)c->stage, *dB = dA + chunk, *dD = dB + chunk;
    for (int64_t off = 0; off < count; off += chunk) {
        const int64_t m = std::min(chunk, count - off);
        BM_HIP(hipMemcpyAsync(dA, a + off, m * 4, hipMemcpyHostToDevice, c->stream));
        BM_HIP(hipMemcpyAsync(dB, b + off, m * 4, hipMemcpyHostToDevice, c->stream));
        BM_LAUNCH(c, "eval_pairs_dense",
                  hipLaunchKernelGGL(k_eval_pairs, dim3((unsigned)((m * 16 + 255) / 256)), dim3(256), 0, c->stream, dA,
                                     dB, m, dD, c->dfptr, c->dfr, c->dfw, c->dfw, c->Q, c->N1, c->ld));
        BM_HIP(hipMemcpyAsync(delta + off, dD, m * 4, hipMemcpyDeviceToHost, c->stream));
        BM_HIP(hipStreamSynchronize(c->stream));
    }
    return BM_OK;
}

// dense: the dense-form lists in place of the comb lists.

// rows a .. a + rows - 1 of k_row_parts; *parts_out = part rows per row, in c->rowpart
static int launch_row_parts(bm_ctx *c, bool dense, int a, int lo, int lo_next, int hi, int mode,
                            unsigned long long *res, int rows, int *parts_out) {
    int heads = 0;
    for (int i = 0; i < rows; i++)
        heads = std::max(heads, (dense ? c->hdfdeg[a + i] : c->hcdeg[a + i]) + c->hg2deg[c->hmap[a + i]]);
    const int N1 = c->N1;
    // about two heads per wave, at most 64 workgroups per row and window
    const int parts = std::min(64, std::max(1, (heads + 31) / 32));
    const int win = std::min(N1, kRowWin), wins = (N1 + win - 1) / win;
    const size_t need = (size_t)rows * parts * N1 * 4;
    if (c->rowpart_bytes < need) {
        dev_free(c, (void **)&c->rowpart, c->rowpart_bytes);
        c->rowpart_bytes = 0;
        if (int rc = dev_alloc(c, (void **)&c->rowpart, need)) return rc;
        c->rowpart_bytes = need;
    }
    const int32_t *lptr = dense ? c->dfptr : c->cptr, *lmid = dense ? c->dfmid : c->cmid;
    const int2 *lrw = (const int2 *)(dense ? c->dfrw : c->crw);
    const int32_t *ldw = dense ? c->dfw : c->cdw;
    BM_LAUNCH(c, "row_expand",
              hipLaunchKernelGGL(k_row_parts, dim3(parts * wins, rows), dim3(1024), (size_t)win * 4, c->stream, a, lo,
                                 lo_next, hi, mode, N1, lptr, lmid, lrw, ldw, c->map, c->g2optr, c->g2oy, c->g2iptr,
                                 c->g2iy, c->Q, c->ld, parts, win, c->rowpart, res));
    *parts_out = parts;
    return BM_OK;
}

static void decode_row_key(unsigned long long k, int mode, int32_t *best_b, int32_t *best_delta);

extern "C" int bm_eval_swap_row(bm_ctx *c, int32_t a, int32_t b_lo, int32_t b_hi, int mode, int32_t *best_b,
                                int32_t *best_delta) {
    if (!c || !best_b || !best_delta) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (mode != BM_ROW_FIRST_POSITIVE && mode != BM_ROW_ARGMAX) return set_error(BM_EINVAL, "bad mode %d", mode);
    if (a < 0 || a > c->n) return set_error(BM_EINVAL, "row %d outside [0,%d]", a, c->n);
    if (b_lo < 0) b_lo = 0;
    if (b_hi > c->n) b_hi = c->n;
    *best_b = 0;
    *best_delta = mode == BM_ROW_ARGMAX ? INT32_MIN : 0;
    if (b_lo > b_hi) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    int rc = ensure_diag(c);
    if (rc) return rc;
    unsigned long long *res = c->keys + c->N1 + 1;
    if (c->sparse_ok) {
        if (c->g2y_dirty) {
            if ((rc = launch_refresh_g2y(c))) return rc;
        }
        int parts;
        if ((rc = launch_row_parts(c, false, a, b_lo, -1, b_hi, mode, res, 1, &parts))) return rc;
        BM_LAUNCH(c, "row_final", hipLaunchKernelGGL(k_row_final, dim3((b_hi - b_lo + 256) / 256), dim3(256), 0,
                                                     c->stream, a, b_lo, -1, b_lo, b_hi, mode, c->rowpart, parts,
                                                     c->N1, c->cneg, c->diag, res, nullptr));
    } else {
        BM_HIP(hipMemsetAsync(res, mode == BM_ROW_FIRST_POSITIVE ? 0xFF : 0x00, 8, c->stream));
        const int blocks = (b_hi - b_lo + kRowCh) / kRowCh;
        const size_t lds = (size_t)2 * c->N1 * sizeof(int16_t);
        if (lds <= 96 * 1024) {
            BM_LAUNCH(c, "row_eval", hipLaunchKernelGGL(k_row_eval<true>, dim3(blocks), dim3(256), lds, c->stream, a,
                                                        b_lo, b_hi, mode, c->N1, c->ld, c->Q, c->cptr, c->cr, c->cw,
                                                        c->cdw, c->diag, res, nullptr));
        } else {
            BM_LAUNCH(c, "row_eval", hipLaunchKernelGGL(k_row_eval<false>, dim3(blocks), dim3(256), 0, c->stream, a,
                                                        b_lo, b_hi, mode, c->N1, c->ld, c->Q, c->cptr, c->cr, c->cw,
                                                        c->cdw, c->diag, res, nullptr));
        }
    }
    BM_HIP(hipMemcpyAsync(c->hres, res, 8, hipMemcpyDeviceToHost, c->stream));  // pinned: no staging copy
    BM_HIP(hipStreamSynchronize(c->stream));
    decode_row_key(c->hres[0], mode, best_b, best_delta);
    return BM_OK;
}

static void decode_row_key(unsigned long long k, int mode, int32_t *best_b, int32_t *best_delta) {
    *best_b = 0;
    *best_delta = mode == BM_ROW_ARGMAX ? INT32_MIN : 0;
    if (mode == BM_ROW_FIRST_POSITIVE) {
        if (k != ~0ull) {
            *best_b = (int32_t)(k >> 32);
            *best_delta = (int32_t)(uint32_t)(k & 0xFFFFFFFFu);
        }
    } else if (k != 0ull) {
        *best_delta = (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
        *best_b = (int32_t)(0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFu));
    }
}

// Consecutive rows in one launch of each row kernel (grid z / y: the row), each
// with its own zeroed accumulator and key; without the sparse preconditions the
// rows go one by one through bm_eval_swap_row.
extern "C" int bm_eval_swap_rows(bm_ctx *c, int32_t a, int32_t count, int32_t b_lo, int32_t b_lo_next, int mode,
                                 int32_t *best_b, int32_t *best_delta) {
    if (!c || !best_b || !best_delta) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (mode != BM_ROW_FIRST_POSITIVE && mode != BM_ROW_ARGMAX) return set_error(BM_EINVAL, "bad mode %d", mode);
    if (count < 1 || count > BM_ROW_GROUP_MAX)
        return set_error(BM_EINVAL, "count=%d outside [1,%d]", count, BM_ROW_GROUP_MAX);
    if (a < 0 || a + count - 1 > c->n)
        return set_error(BM_EINVAL, "rows %d..%d outside [0,%d]", a, a + count - 1, c->n);
    const int n = c->n;
    auto lo_of = [&](int i) { return std::max(0, i == 0 ? b_lo : (b_lo_next < 0 ? a + i + 1 : b_lo_next)); };
    if (!c->sparse_ok) {
        for (int i = 0; i < count; i++)
            if (int rc = bm_eval_swap_row(c, a + i, lo_of(i), n, mode, best_b + i, best_delta + i)) return rc;
        return BM_OK;
    }
    if (int prc = ensure_p(c)) return prc;
    BM_HIP(hipSetDevice(c->device));
    int rc = ensure_diag(c);
    if (rc) return rc;
    if (!c->grpres && (rc = dev_alloc(c, (void **)&c->grpres, (size_t)BM_ROW_GROUP_MAX * 8))) return rc;
    if (!c->hgrpres) {
        std::lock_guard<std::mutex> g(sync_mutex());  // host allocations may wait for the device
        BM_HIP(hipHostMalloc((void **)&c->hgrpres, (size_t)BM_ROW_GROUP_MAX * 8, hipHostMallocDefault));
    }
    if (c->g2y_dirty && (rc = launch_refresh_g2y(c))) return rc;
    // rows whose range is empty keep their neutral key (set by k_row_parts) and
    // add nothing: k_row_parts and k_row_final test b against the row's own range
    int lo_base = n + 1;
    for (int i = 0; i < count; i++) lo_base = std::min(lo_base, lo_of(i));
    const int first_lo = lo_of(0), next_lo = b_lo_next < 0 ? -1 : std::max(0, b_lo_next);
    int parts;
    if ((rc = launch_row_parts(c, false, a, first_lo, next_lo, n, mode, c->grpres, count, &parts))) return rc;
    if (lo_base <= n)
        BM_LAUNCH(c, "row_final", hipLaunchKernelGGL(k_row_final, dim3((n - lo_base + 256) / 256, count), dim3(256),
                                                     0, c->stream, a, first_lo, next_lo, lo_base, n, mode, c->rowpart,
                                                     parts, c->N1, c->cneg, c->diag, c->grpres, nullptr));
    BM_HIP(hipMemcpyAsync(c->hgrpres, c->grpres, (size_t)count * 8, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < count; i++) decode_row_key(c->hgrpres[i], mode, best_b + i, best_delta + i);
    return BM_OK;
}

// Whole rows of the changed vertices (quick.cu:141-250 evaluates exactly
// the pairs that touch them): the row kernels above with their reduction
// replaced by a store of every delta, rows one after another on the stream
// into the staging buffer, then one copy to the caller.
extern "C" int bm_eval_rows(bm_ctx *c, const int32_t *rows, int32_t count, int32_t kind, int32_t *deltas) {
    if (!c || (count > 0 && (!rows || !deltas))) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (count < 0 || count > 64) return set_error(BM_EINVAL, "count=%d outside [0,64]", count);
    if (kind != BM_DELTA_LISTS && kind != BM_DELTA_DENSE) return set_error(BM_EINVAL, "kind=%d unknown", kind);
    for (int i = 0; i < count; i++) {
        if (rows[i] < 1 || rows[i] > c->n) return set_error(BM_EINVAL, "rows[%d]=%d outside [1,%d]", i, rows[i], c->n);
        for (int j = 0; j < i; j++)
            if (rows[j] == rows[i]) return set_error(BM_EINVAL, "row %d repeated", rows[i]);
    }
    if (count == 0) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    const bool dense = kind == BM_DELTA_DENSE;
    int rc = dense ? ensure_dense_lists(c) : ensure_diag(c);
    if (rc) return rc;
    const size_t N1 = (size_t)c->N1;
    if ((rc = ensure_stage(c, (size_t)count * N1 * 4))) return rc;
    int32_t *out = (int32_t *)c->stage;
    unsigned long long *res = c->keys + c->N1 + 1;  // written by k_row_parts, not read here
    if (c->sparse_ok && c->g2y_dirty && (rc = launch_refresh_g2y(c))) return rc;
    // the dense form: X(v,v) over its own lists, from the current P (kept current by later moves)
    if (dense && !c->dfdiag_valid) {
        if ((rc = launch_list_diag(c, c->dfdiag, c->dfptr, c->dfr, c->dfw))) return rc;
        c->dfdiag_valid = true;
    }
    const int32_t *lptr = dense ? c->dfptr : c->cptr, *lr = dense ? c->dfr : c->cr;
    const int32_t *lw = dense ? c->dfw : c->cw, *ldw = dense ? c->dfw : c->cdw;
    const int32_t *ldiag = dense ? c->dfdiag : c->diag, *lneg = dense ? c->dfneg : c->cneg;
    const int n = c->n;
    for (int i = 0; i < count; i++) {
        const int a = rows[i];
        int32_t *row = out + (size_t)i * N1;
        if (c->sparse_ok) {
            int parts;
            if ((rc = launch_row_parts(c, dense, a, 0, -1, n, BM_ROW_ARGMAX, res, 1, &parts))) return rc;
            BM_LAUNCH(c, "row_store", hipLaunchKernelGGL(k_row_final, dim3((n + 256) / 256), dim3(256), 0, c->stream,
                                                         a, 0, -1, 0, n, BM_ROW_ARGMAX, c->rowpart, parts, c->N1,
                                                         lneg, ldiag, res, row));
        } else {
            const int blocks = (n + kRowCh) / kRowCh;
            const size_t lds = (size_t)2 * c->N1 * sizeof(int16_t);
            if (lds <= 96 * 1024) {
                BM_LAUNCH(c, "row_store", hipLaunchKernelGGL(k_row_eval<true>, dim3(blocks), dim3(256), lds, c->stream,
                                                             a, 0, n, BM_ROW_ARGMAX, c->N1, c->ld, c->Q, lptr, lr, lw,
                                                             ldw, ldiag, res, row));
            } else {
                BM_LAUNCH(c, "row_store", hipLaunchKernelGGL(k_row_eval<false>, dim3(blocks), dim3(256), 0, c->stream,
                                                             a, 0, n, BM_ROW_ARGMAX, c->N1, c->ld, c->Q, lptr, lr, lw,
                                                             ldw, ldiag, res, row));
            }
        }
    }
    // through a pinned buffer: a copy into pageable memory is staged by the
    // runtime in pieces, several times slower for rows of this size
    const size_t bytes = (size_t)count * N1 * 4;
    if (c->hrows_bytes < bytes) {
        BM_HIP(hipStreamSynchronize(c->stream));
        std::lock_guard<std::mutex> g(sync_mutex());  // host allocations may wait for the device
        if (c->hrows) (void)hipHostFree(c->hrows);
        c->hrows = nullptr;
        c->hrows_bytes = 0;
        const size_t want = std::max(bytes, (size_t)8 * N1 * 4);
        BM_HIP(hipHostMalloc(&c->hrows, want, hipHostMallocDefault));
        c->hrows_bytes = want;
    }
    BM_HIP(hipMemcpyAsync(c->hrows, out, bytes, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    memcpy(deltas, c->hrows, bytes);
    return BM_OK;
}

__global__ void k_fill_keys(unsigned long long *__restrict__ p, int64_t count, const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0ull;
}

int bm::sweep_prepare(bm_ctx *c) {
    int rc;
    if (!c->X) {
        c->ldx = c->ld;
        // tile-major: whole 64-row tiles, the last one partly unused
        if ((rc = dev_alloc(c, (void **)&c->X, (size_t)((c->N1 + 63) / 64 * 64) * c->ldx * 4))) return rc;
    }
    if ((rc = choose_gain(c))) return rc;
    if (c->sparse_ok && (size_t)c->N1 * 4 <= kSparseLdsMax && (rc = build_row_plan(c))) return rc;
    if (c->g2y_dirty && c->sparse_ok) {
        if ((rc = launch_refresh_g2y(c))) return rc;
    }
    return BM_OK;
}

// Graph-capturable: kernels only (the choice of gain kernel and all
// allocations are made by sweep_prepare).
bool bm::sweep_is_sparse(const bm_ctx *c) { return c->sweep_kind == 2 && sparse_fits(c); }

int bm::sweep_enqueue(bm_ctx *c, int32_t *dmat, const int *run_flag, const int *skip_flag) {
    const int N1 = c->N1;
    int rc;
    BM_LAUNCH(c, "fill", hipLaunchKernelGGL(k_fill_keys, dim3((N1 + 256) / 256), dim3(256), 0, c->stream, c->keys,
                                            (int64_t)N1 + 1, skip_flag));
    const bool sparse = c->sweep_kind == 2 && c->sparse_ok && (size_t)N1 * 4 <= kSparseLdsMax;
    if (sparse) {
        if ((rc = launch_gain_sparse(c, run_flag))) return rc;  // also writes X(a,a) and folds Corr in
    } else {
        if ((rc = launch_diag(c))) return rc;
        if ((rc = launch_gain_cols(c, c->gain_cols))) return rc;
        if (c->e1 > 0)
            BM_LAUNCH(c, "corr", hipLaunchKernelGGL(k_corr, dim3((unsigned)((c->e1 + 255) / 256)), dim3(256), 0,
                                                    c->stream, c->esrc, c->edst, c->ew, c->edw, c->e1, c->Q, c->ld,
                                                    c->X, c->ldx));
    }
    c->diag_valid = true;
    const int T = (N1 + 63) / 64;
    BM_LAUNCH(c, "combine", hipLaunchKernelGGL(k_combine, dim3(T, (T + kCombChunk - 1) / kCombChunk), dim3(256), 0,
                                               c->stream, c->X, c->ldx, c->diag, c->n, c->keys, c->keys + N1, dmat,
                                               skip_flag));
    return BM_OK;
}

static int sweep_results(bm_ctx *c, int32_t *dmat, int32_t *best_a, int32_t *best_b, int32_t *best_delta,
                         int32_t *row_best_b, int32_t *row_best_delta, int32_t *delta_matrix);

// The dense-matrix delta of calculateSwapDeltasKernel (quick.cu:48-138,
// graph-re-alignment-gpu.cu:60-128) for every pair: Xd over the dense-form
// lists with the dense all-pairs kernel, their X(v,v), Corr of their lines
// (each its own dense cell) folded in, then the same combine and arg-max.
static int sweep_dense_enqueue(bm_ctx *c, int32_t *dmat) {
    const int N1 = c->N1;
    int rc = ensure_dense_lists(c);
    if (rc) return rc;
    if (!c->dfX && (rc = dev_alloc(c, (void **)&c->dfX, (size_t)(N1 + 63) / 64 * 64 * c->ld * 4))) return rc;
    BM_LAUNCH(c, "fill", hipLaunchKernelGGL(k_fill_keys, dim3((N1 + 256) / 256), dim3(256), 0, c->stream, c->keys,
                                            (int64_t)N1 + 1, nullptr));
    if ((rc = launch_list_diag(c, c->dfdiag, c->dfptr, c->dfr, c->dfw))) return rc;
    c->dfdiag_valid = true;
    if ((rc = launch_list_gain(c, c->dfptr, c->dfrw, c->dfX, c->dfmaxw))) return rc;
    if (c->ndfout > 0)
        BM_LAUNCH(c, "corr", hipLaunchKernelGGL(k_corr, dim3((unsigned)((c->ndfout + 255) / 256)), dim3(256), 0,
                                                c->stream, c->dfes, c->dfed, c->dfew, c->dfew, c->ndfout, c->Q,
                                                c->ld, c->dfX, (int64_t)c->ld));
    const int T = (N1 + 63) / 64;
    BM_LAUNCH(c, "combine", hipLaunchKernelGGL(k_combine, dim3(T, (T + kCombChunk - 1) / kCombChunk), dim3(256), 0,
                                               c->stream, c->dfX, (int64_t)c->ld, c->dfdiag, c->n, c->keys,
                                               c->keys + N1, dmat, nullptr));
    return BM_OK;
}

extern "C" int bm_sweep_all_pairs_dense(bm_ctx *c, int32_t *best_a, int32_t *best_b, int32_t *best_delta,
                                        int32_t *row_best_b, int32_t *row_best_delta, int32_t *delta_matrix) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    BM_HIP(hipSetDevice(c->device));
    const int N1 = c->N1;
    int rc;
    int32_t *dmat = nullptr;
    if (delta_matrix) {
        if ((rc = ensure_stage(c, (size_t)N1 * N1 * 4))) return rc;
        dmat = (int32_t *)c->stage;
        BM_HIP(hipMemsetAsync(dmat, 0, (size_t)N1 * N1 * 4, c->stream));
    }
    // when every G1 line is its own dense cell the comb lists are the dense-form lists and the two
    // deltas are one: the pair sweep's own (sparse) kernels serve
    if (c->dense_equal) {
        if ((rc = sweep_prepare(c)) || (rc = sweep_enqueue(c, dmat))) return rc;
    } else if ((rc = sweep_dense_enqueue(c, dmat))) {
        return rc;
    }
    return sweep_results(c, dmat, best_a, best_b, best_delta, row_best_b, row_best_delta, delta_matrix);
}

extern "C" int bm_sweep_all_pairs(bm_ctx *c, int32_t *best_a, int32_t *best_b, int32_t *best_delta,
                                  int32_t *row_best_b, int32_t *row_best_delta, int32_t *delta_matrix) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    BM_HIP(hipSetDevice(c->device));
    const int N1 = c->N1;
    int rc;
    int32_t *dmat = nullptr;
    if (delta_matrix) {
        if ((rc = ensure_stage(c, (size_t)N1 * N1 * 4))) return rc;
        dmat = (int32_t *)c->stage;
        BM_HIP(hipMemsetAsync(dmat, 0, (size_t)N1 * N1 * 4, c->stream));
    }
    if ((rc = sweep_prepare(c))) return rc;
    if ((rc = sweep_enqueue(c, dmat))) return rc;
    return sweep_results(c, dmat, best_a, best_b, best_delta, row_best_b, row_best_delta, delta_matrix);
}

// Row-best and global keys (and the matrix) of the last combine back to the host.
static int sweep_results(bm_ctx *c, int32_t *dmat, int32_t *best_a, int32_t *best_b, int32_t *best_delta,
                         int32_t *row_best_b, int32_t *row_best_delta, int32_t *delta_matrix) {
    const int N1 = c->N1;
    unsigned long long *gkey = c->keys + N1;
    std::vector<unsigned long long> keys;
    unsigned long long g = 0;
    if (row_best_b || row_best_delta) {
        keys.resize(N1);
        BM_HIP(hipMemcpyAsync(keys.data(), c->keys, (size_t)N1 * 8, hipMemcpyDeviceToHost, c->stream));
    }
    BM_HIP(hipMemcpyAsync(&g, gkey, 8, hipMemcpyDeviceToHost, c->stream));
    if (delta_matrix)
        BM_HIP(hipMemcpyAsync(delta_matrix, dmat, (size_t)N1 * N1 * 4, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    for (int v = 0; v < (int)keys.size(); v++) {
        int32_t bb = 0, dd = INT32_MIN;
        if (keys[v]) {
            dd = (int32_t)((uint32_t)(keys[v] >> 32) ^ 0x80000000u);
            bb = (int32_t)(0xFFFFFFFFu - (uint32_t)(keys[v] & 0xFFFFFFFFu));
        }
        if (row_best_b) row_best_b[v] = bb;
        if (row_best_delta) row_best_delta[v] = dd;
    }
    int32_t ba = 0, bb = 0, bd = 0;
    if (g) {
        const int32_t d = (int32_t)((uint32_t)(g >> 32) ^ 0x80000000u);
        const uint32_t ab = 0xFFFFFFFFu - (uint32_t)(g & 0xFFFFFFFFu);
        if (d > 0) {
            bd = d;
            ba = (int32_t)(ab >> 16);
            bb = (int32_t)(ab & 0xFFFF);
        }
    }
    if (best_a) *best_a = ba;
    if (best_b) *best_b = bb;
    if (best_delta) *best_delta = bd;
    return BM_OK;
}

extern "C" int bm_sweep_config(bm_ctx *c, int *kernel, int *dense_cols) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    const bool sparse = c->sweep_kind == 2 && c->sparse_ok && (size_t)c->N1 * 4 <= kSparseLdsMax;
    if (kernel) *kernel = c->gain_cols == 0 ? 0 : sparse ? 2 : 1;
    if (dense_cols) *dense_cols = c->gain_cols;
    return BM_OK;
}
