// Kernels for the 3-vertex, k-set and hill-climb move kinds (gfx950).
//
//   k_rot3     : Delta_G3 of triple.cu:85-214 for a triple list, kLanes
//                (8, 16, 32 or 64) lanes per triple.  Each neighbour entry of m_k
//                loads the three values Q[r][m_0..2] once and feeds all five
//                permutations.
//   k_permset  : Delta_P of graph-alignment-permset-parallel.c:304-339 for
//                every permutation of a k-set, one workgroup (64 or 256
//                threads) per set: an LDS table C[i][j] (vertex i takes j's
//                image) plus the intra-set lines (counted twice, as the
//                reference does, or once for the exact score change), then
//                k! evaluations in next_permutation order, strict-> arg-max.
//   k_hillclimb: Delta_A of refine-matching-hillclimb.c:244-296.
#include <limits.h>

#include <stdlib.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "bm_device.h"
#include "bm_internal.h"
#include "bm_wave.h"

using namespace bm;

__device__ __forceinline__ int qv(const int16_t *__restrict__ Q, int row, int ld, int col) {
    return (int)Q[(size_t)row * (size_t)ld + col];
}

// Q[r][m] read from row m of the other half: Q[n+1+d][m] = P^T[d][m] = P[m][d]
// and Q[s][m] = P[s][m] = P^T[m][s].  A move kernel reads column m_j for many
// neighbour rows r; this way those reads stay inside the rows m_j (a few KB
// per move vertex) instead of touching one cache line per neighbour.
__device__ __forceinline__ int qcol(const int16_t *__restrict__ Q, int r, int N1, int ld, int m) {
    return r >= N1 ? (int)Q[(size_t)m * (size_t)ld + (r - N1)] : (int)Q[((size_t)N1 + m) * (size_t)ld + r];
}

// permutation table of triple.cu:85-92: new image of m_k = old image of m_tab[p][k]
__constant__ int kG3Tab[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
// the same table as a compile-time constant, for indices the unrolled loops fix
constexpr int kG3Perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
constexpr int kRot3Unroll = 4;

template <int kLanes>
__global__ __launch_bounds__(256) void k_rot3(const int32_t *__restrict__ T, int64_t count, int mode,
                                              int32_t *__restrict__ best_d, int32_t *__restrict__ best_p,
                                              const int32_t *__restrict__ ddptr, const int32_t *__restrict__ ddr,
                                              const int32_t *__restrict__ ddw, const int16_t *__restrict__ Q, int N1,
                                              int ld) {
    const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanes;
    const int gl = threadIdx.x & (kLanes - 1);
    int s[6] = {0, 0, 0, 0, 0, 0};
    int m[3] = {0, 0, 0};
    if (t < count) {
        m[0] = T[3 * t];
        m[1] = T[3 * t + 1];
        m[2] = T[3 * t + 2];
        // rows m_j of P (out entries) and of P^T (in entries): every column read stays in them
        const int16_t *rp[3], *rt[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            rp[j] = Q + (size_t)m[j] * ld;
            rt[j] = Q + ((size_t)N1 + m[j]) * ld;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int v = m[k];
            const int e1 = ddptr[v + 1];
            // kRot3Unroll entries per lane in flight: their list words, then their 3 column reads each
            for (int eb = ddptr[v] + gl; eb < e1; eb += kLanes * kRot3Unroll) {
                int r[kRot3Unroll], w[kRot3Unroll], q[kRot3Unroll][3];
#pragma unroll
                for (int u = 0; u < kRot3Unroll; u++) {
                    const int e = eb + u * kLanes;
                    const bool in = e < e1;
                    r[u] = in ? ddr[e] : N1;  // past the list: reads column 0, contributes nothing
                    w[u] = in ? (int)(int16_t)ddw[e] : 0;  // the dense G1 cell
                }
#pragma unroll
                for (int u = 0; u < kRot3Unroll; u++) {
                    const bool out = r[u] >= N1;
                    const int col = out ? r[u] - N1 : r[u];
#pragma unroll
                    for (int j = 0; j < 3; j++) q[u][j] = (int)(out ? rp[j] : rt[j])[col];
                }
#pragma unroll
                for (int u = 0; u < kRot3Unroll; u++) {
                    const int wu = w[u];
                    if (wu <= 0) continue;
                    const bool out = r[u] >= N1;
                    const int nb = out ? r[u] - N1 : r[u];
                    if (nb == 0) continue;  // the dense scan runs over i = 1..n
                    const int j = nb == m[0] ? 0 : nb == m[1] ? 1 : nb == m[2] ? 2 : -1;
                    if (j >= 0) {
                        // intra-triple cell m_k -> m_j (:164-207), from the out entry only
                        if (out && j != k) {
                            const int old = min(wu, qv(Q, m[k], ld, m[j]));
#pragma unroll
                            for (int p = 1; p < 6; p++)
                                s[p] += min(wu, qv(Q, m[kG3Tab[p][k]], ld, m[kG3Tab[p][j]])) - old;
                        }
                        continue;
                    }
                    const int xs[3] = {min(wu, q[u][0]), min(wu, q[u][1]), min(wu, q[u][2])};
                    const int old = xs[k];
#pragma unroll
                    for (int p = 1; p < 6; p++) s[p] += xs[kG3Perm[p][k]] - old;
                }
            }
        }
    }
#pragma unroll
    for (int p = 1; p < 6; p++) s[p] = bm_wave::group_sum<kLanes>(s[p]);
    if (t < count && gl == 0) {
        int bd = 0, bp = 0;
        for (int p = 1; p < 6; p++) {
            if (mode == BM_ROT3_ROTATIONS && p != 3 && p != 4) continue;
            if (s[p] > bd) {
                bd = s[p];
                bp = p;
            }
        }
        best_d[t] = bd;
        best_p[t] = bp;
    }
}

// Delta_G3 of every permutation q of a triple from its 3x3 blocks: x = Xd
// (the dense-form list sums, x[i][j] = Xd(m_i, m_j)), p = P at the triple,
// c = the G1 cells between triple vertices (0 on the diagonal and where
// absent).  Moving m_k to the image of m_s(k) changes its list terms by
// x[k][s(k)] - x[k][k]; the entries between triple vertices, which Xd
// counted, are taken out and the intra-triple cells (triple.cu:164-207)
// added instead.  Best strict > from 0 over the allowed q (:94-214).
__device__ __forceinline__ void rot3_score(const int (&x)[3][3], const int (&p)[3][3], const int (&c)[3][3],
                                           int mode, int32_t &best_d, int32_t &best_p) {
    int bd = 0, bp = 0;
#pragma unroll
    for (int q = 1; q < 6; q++) {
        if (mode == BM_ROT3_ROTATIONS && q != 3 && q != 4) continue;
        // m_k takes the image of m_s(k), s(k) = kG3Perm[q][k]
        int d = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) d += x[k][kG3Perm[q][k]] - x[k][k];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int w = c[i][j];
                if (i == j || w <= 0) continue;
                const int si = kG3Perm[q][i], sj = kG3Perm[q][j];
                const int old = min(w, p[i][j]);
                // the cell's out entry in m_i's list and in entry in m_j's list, as counted in Xd
                d -= min(w, p[si][j]) - old;
                d -= min(w, p[i][sj]) - old;
                d += min(w, p[si][sj]) - old;  // its own term
            }
        if (d > bd) {
            bd = d;
            bp = q;
        }
    }
    best_d = bd;
    best_p = bp;
}

// Table form of the same Delta_G3, Xd(u,v) being the sum over the dense-form
// list of u (the entries triple.cu's scan visits: cells > 0 as short, no
// self-loop, no vertex 0) of min(w, Q[r][v]).  rot3_score needs only the 3x3
// blocks of Xd, of P and of the G1 cells at the triple, which one table holds
// side by side: XP[u][v] = (Xd(u,v), W1[u][v] << 16 | (uint16)P[u][v]),
// 8 bytes, so one thread per triple makes nine loads instead of three per
// list entry.
__global__ __launch_bounds__(256) void k_rot3_table(const int32_t *__restrict__ T, int64_t count, int mode,
                                                    int32_t *__restrict__ best_d, int32_t *__restrict__ best_p,
                                                    const int2 *__restrict__ XP, int ld) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int m[3];
#pragma unroll
    for (int i = 0; i < 3; i++) m[i] = T[3 * t + i];
    int p[3][3], x[3][3], c[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int2 v = XP[(size_t)m[i] * ld + m[j]];
            x[i][j] = v.x;
            p[i][j] = (int)(int16_t)(v.y & 0xffff);  // P[m_i][m_j] = W2[f_i][f_j]
            c[i][j] = i == j ? 0 : v.y >> 16;         // the G1 cell m_i -> m_j
        }
    rot3_score(x, p, c, mode, best_d[t], best_p[t]);
}

// XP[u][v] for u < N1, v < ld from the tile-major Xd, P (rows 0..n of Q) and
// the cell matrix; one thread per pair of columns.
__global__ __launch_bounds__(256) void k_table_pack(int2 *__restrict__ XP, const int32_t *__restrict__ X,
                                                    const int16_t *__restrict__ Q, const int16_t *__restrict__ W1,
                                                    int N1, int ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t half = (int64_t)ld / 2;
    if (i >= (int64_t)N1 * half) return;
    const int u = (int)(i / half), v = (int)(i - (int64_t)u * half) * 2;
    const size_t at = (size_t)u * ld + v;
    const size_t xa = xidx(u, v, ld >> 6);  // v even: v and v + 1 share the tile row
    const int2 xs = *reinterpret_cast<const int2 *>(X + xa);
    const int ps = *reinterpret_cast<const int *>(Q + at), ws = *reinterpret_cast<const int *>(W1 + at);
    int4 out;
    out.x = xs.x;
    out.y = (int)(((uint32_t)ws << 16) | ((uint32_t)ps & 0xffffu));
    out.z = xs.y;
    out.w = (int)(((uint32_t)ws & 0xffff0000u) | ((uint32_t)ps >> 16));
    *reinterpret_cast<int4 *>(XP + at) = out;
}

// The table form inside the batched step, on the step's own all-pairs gain
// matrix X (tile-major, kept current by the step).  When every G1 line is
// its own dense cell (no self-loop, no repeated pair, no line at vertex 0,
// weights in [1, 32767]), the comb lists are the dense-form lists and X = Xd,
// except that X[min(a,b)][max(a,b)] also holds Corr(a,b) of the pair sweep
// (pair_corr over the lines between a and b, each its own dense cell); that
// is taken out again for the triple's pairs.
// With pmap, Q is W2 and P[x][y] is read as W2[p(x)][p(y)].
__global__ __launch_bounds__(256) void k_rot3_x(const int32_t *__restrict__ T, int64_t count, int mode,
                                                int32_t *__restrict__ best_d, int32_t *__restrict__ best_p,
                                                const int32_t *__restrict__ X, int tiles,
                                                const int16_t *__restrict__ Q, const int16_t *__restrict__ W1,
                                                int ld, const int32_t *__restrict__ pmap) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int m[3];
#pragma unroll
    for (int i = 0; i < 3; i++) m[i] = T[3 * t + i];
    int g[3];
#pragma unroll
    for (int i = 0; i < 3; i++) g[i] = pmap ? pmap[m[i]] : m[i];
    int p[3][3], x[3][3], c[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            p[i][j] = qv(Q, g[i], ld, g[j]);
            x[i][j] = X[xidx(m[i], m[j], tiles)];
            c[i][j] = i == j ? 0 : (int)W1[(size_t)m[i] * ld + m[j]];
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (i == j || m[i] > m[j]) continue;
            // Corr(m_i, m_j) from m_i's entries: the line m_i -> m_j (out) and m_j -> m_i (in)
            int corr = 0;
            if (c[i][j] > 0) corr += pair_corr(true, c[i][j], c[i][j], p[i][i], p[j][j], p[i][j], p[j][i]);
            if (c[j][i] > 0) corr += pair_corr(false, c[j][i], c[j][i], p[i][i], p[j][j], p[i][j], p[j][i]);
            x[i][j] -= corr;
        }
    rot3_score(x, p, c, mode, best_d[t], best_p[t]);
}

// dfW1[v][d] = the dense-form cell of the out entry v -> d (0 elsewhere).
__global__ __launch_bounds__(256) void k_df_cells(int16_t *__restrict__ W1, const int32_t *__restrict__ ptr,
                                                  const int32_t *__restrict__ mid, const int2 *__restrict__ rw,
                                                  int N1, int ld) {
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= N1) return;
    for (int e = ptr[v] + (threadIdx.x & 63); e < mid[v]; e += 64) {
        const int2 x = rw[e];
        W1[(size_t)v * ld + (x.x - N1)] = (int16_t)x.y;
    }
}

__global__ void k_zero16(int16_t *__restrict__ p, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

constexpr int kMaxIntra = 256;
constexpr int kTuneMinTriples = 4096;  // lists shorter than this use the current variant untimed
constexpr int kTuneMinSets = 256;

// kFixedK > 0 compiles the kernel for that k (the loops over set positions
// unroll and index the per-thread permutation statically); 0 takes k at run time.
// With X (the batched step, when every G1 line is its own dense cell, W1 the
// cell matrix) the table is read from the all-pairs gain matrix instead of
// the lists: C[i][j] = X(S_i, S_j) - X(S_i, S_i) without the list entries
// of S_i's lines to the set (they are the intra lines), and without the
// Corr(S_i, S_j) that the pair sweep folds into X[min][max].
template <int kThreads, int kFixedK>
__global__ __launch_bounds__(kThreads) void k_permset(const int32_t *__restrict__ sets, int64_t count, int k_arg,
                                                      int32_t *__restrict__ best_d, int32_t *__restrict__ best_rank,
                                                      const int32_t *__restrict__ cptr, const int32_t *__restrict__ cr,
                                                      const int32_t *__restrict__ cw, const int16_t *__restrict__ Q,
                                                      int N1, int ld, int *__restrict__ overflow, int mult,
                                                      const int32_t *__restrict__ X, int tiles,
                                                      const int16_t *__restrict__ W1, const int16_t *__restrict__ W2,
                                                      const int32_t *__restrict__ pmap) {
    constexpr int kWaves = kThreads / 64;
    const int k = kFixedK ? kFixedK : k_arg;
    __shared__ int S[8];
    __shared__ int C[8][8];
    __shared__ int PS[8][8];
    __shared__ int CW[8][8];  // G1 cells between set members (the table from X)
    __shared__ int ni;
    __shared__ int iw[kMaxIntra];
    __shared__ unsigned char ix[kMaxIntra], il[kMaxIntra];
    __shared__ int lbase;
    __shared__ unsigned long long red[kWaves];
    const int64_t set = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < k) S[tid] = sets[set * k + tid];
    if (tid < 64) C[tid >> 3][tid & 7] = 0;
    if (tid == 0) {
        ni = 0;
        lbase = 0;
    }
    __syncthreads();
    if (tid < k * k) {
        // P among the set; with pmap (the table from X only) read as W2[p(x)][p(y)]
        PS[tid / k][tid % k] = pmap ? qv(W2, pmap[S[tid / k]], ld, pmap[S[tid % k]]) : qv(Q, S[tid / k], ld, S[tid % k]);
        if (X) CW[tid / k][tid % k] = tid / k == tid % k ? 0 : (int)W1[(size_t)S[tid / k] * ld + S[tid % k]];
    }
    if (X) {
        __syncthreads();
        if (tid < k * k) {
            const int i = tid / k, j = tid % k;
            if (i != j && CW[i][j] > 0) {  // the intra line S_i -> S_j (k * k <= kMaxIntra slots)
                const int slot = atomicAdd(&ni, 1);
                iw[slot] = CW[i][j];
                ix[slot] = (unsigned char)i;
                il[slot] = (unsigned char)j;
            }
            if (i != j) {
                const int a = S[i], b = S[j];
                int x = X[xidx(a, b, tiles)] - X[xidx(a, a, tiles)];
                if (a < b) {  // Corr(a, b), folded into X[a][b]: the lines a -> b (out) and b -> a (in)
                    if (CW[i][j] > 0) x -= pair_corr(true, CW[i][j], CW[i][j], PS[i][i], PS[j][j], PS[i][j], PS[j][i]);
                    if (CW[j][i] > 0) x -= pair_corr(false, CW[j][i], CW[j][i], PS[i][i], PS[j][j], PS[i][j], PS[j][i]);
                }
                // the entries of S_i's lines to the set: out S_i -> S_l reads P[v][S_l], in S_l -> S_i reads P[S_l][v]
                for (int l = 0; l < k; l++) {
                    if (l == i) continue;
                    const int wo = CW[i][l], wi = CW[l][i];
                    if (wo > 0) x -= min(wo, PS[j][l]) - min(wo, PS[i][l]);
                    if (wi > 0) x -= min(wi, PS[l][j]) - min(wi, PS[l][i]);
                }
                C[i][j] = x;
            }
        }
    }
    // table C over all entries of the k lists: per-thread partial sums of row i
    // in registers, one wave reduction and one LDS atomic per (wave, j)
    for (int i = 0; i < (X ? 0 : k); i++) {
        const int v = S[i];
        int part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int e = cptr[v] + tid; e < cptr[v + 1]; e += kThreads) {
            const int r = cr[e], w = cw[e];
            const bool out = r >= N1;
            const int nb = out ? r - N1 : r;
            int l = -1;
            for (int q = 0; q < k; q++)
                if (S[q] == nb) l = q;
            if (l >= 0) {
                // intra line: both of its list entries change with sigma;
                // keep it once (from the out entry), weighted by mult when evaluated
                if (out) {
                    const int slot = atomicAdd(&ni, 1);
                    if (slot < kMaxIntra) {
                        iw[slot] = w;
                        ix[slot] = (unsigned char)i;
                        il[slot] = (unsigned char)l;
                    }
                }
                continue;
            }
            const int old = min(w, qcol(Q, r, N1, ld, v));
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (j < k && j != i) part[j] += min(w, qcol(Q, r, N1, ld, S[j])) - old;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (j >= k || j == i) continue;
            int x = part[j];
            x = bm_wave::wave_sum(x);
            if ((tid & 63) == 0 && x) atomicAdd(&C[i][j], x);
        }
    }
    __syncthreads();
    const int nint = min(ni, kMaxIntra);
    if (ni > kMaxIntra && tid == 0) atomicOr(overflow, 1);
    for (int q = tid; q < nint; q += kThreads) atomicAdd(&lbase, -mult * min(iw[q], PS[ix[q]][il[q]]));
    __syncthreads();
    // k! permutations by lexicographic rank (the order of next_permutation,
    // :342-363): rank = prefix * s! + suffix, where a prefix fixes the first
    // k - s positions (s = min(k, 3)) and its s! suffixes are the orderings of
    // the s values left, themselves in lexicographic order.  A thread takes
    // whole prefixes (ranks ascending within the thread); a permutation is
    // kept packed, 3 bits per position, so indexing it costs a shift.
    int total = 1;
    for (int i = 2; i <= k; i++) total *= i;
    const int sl = k < 3 ? k : 3, sfact = sl == 3 ? 6 : 2;
    const int nprefix = total / sfact, plen = k - sl;
    int bestd = 0, bestr = -1;
    for (int t = tid; t < nprefix; t += kThreads) {
        // Lehmer digits of the prefix: position i takes the d-th smallest unused value,
        // d = (t / ((k-1-i)! / s!)) mod (k - i)
        int rem = t, used = 0, pk = 0, psum = 0;
        int f = nprefix;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if (i >= plen) break;
            f /= k - i;
            int d = rem / f;
            rem -= d * f;
            int v = 0;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (j >= k || ((used >> j) & 1)) continue;
                if (d == 0) {
                    v = j;
                    d = -1;
                } else if (d > 0) {
                    d--;
                }
            }
            used |= 1 << v;
            pk |= v << (3 * i);
            psum += C[i][v];
        }
        // the s values left, ascending
        int rest[3] = {0, 0, 0}, nr = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (j < k && !((used >> j) & 1)) {
                if (nr < 3) rest[nr] = j;
                nr++;
            }
        const int base = t * sfact;
        auto eval = [&](int u, int sum, int pku) {
            int d = lbase + psum + sum;
            for (int q = 0; q < nint; q++)
                d += mult * min(iw[q], PS[(pku >> (3 * ix[q])) & 7][(pku >> (3 * il[q])) & 7]);
            if (d > bestd) {
                bestd = d;
                bestr = base + u;
            }
        };
        const int p0 = plen, p1 = plen + 1;
        const int a = rest[0], b = rest[1];
        if (sl == 2) {
            eval(0, C[p0][a] + C[p1][b], pk | a << (3 * p0) | b << (3 * p1));
            eval(1, C[p0][b] + C[p1][a], pk | b << (3 * p0) | a << (3 * p1));
        } else {
            const int p2 = plen + 2, cc = rest[2];
            const int a0 = C[p0][a], b0 = C[p0][b], c0 = C[p0][cc];
            const int a1 = C[p1][a], b1 = C[p1][b], c1 = C[p1][cc];
            const int a2 = C[p2][a], b2 = C[p2][b], c2 = C[p2][cc];
            const int sa = a << (3 * p0), sb = b << (3 * p0), sc = cc << (3 * p0);
            const int ta = a << (3 * p1), tb = b << (3 * p1), tc = cc << (3 * p1);
            const int ua = a << (3 * p2), ub = b << (3 * p2), uc = cc << (3 * p2);
            eval(0, a0 + b1 + c2, pk | sa | tb | uc);
            eval(1, a0 + c1 + b2, pk | sa | tc | ub);
            eval(2, b0 + a1 + c2, pk | sb | ta | uc);
            eval(3, b0 + c1 + a2, pk | sb | tc | ua);
            eval(4, c0 + a1 + b2, pk | sc | ta | ub);
            eval(5, c0 + b1 + a2, pk | sc | tb | ua);
        }
    }
    // block arg-max: larger delta, then smaller rank
    unsigned long long key = 0;
    if (bestr >= 0)
        key = ((unsigned long long)((unsigned)bestd ^ 0x80000000u) << 32) | (0xFFFFFFFFu - (unsigned)bestr);
    key = bm_wave::wave_max_u64(key);
    if (kWaves > 1) {
        if ((tid & 63) == 0) red[tid >> 6] = key;
        __syncthreads();
        if (tid == 0)
            for (int q = 1; q < kWaves; q++) key = max(key, red[q]);
    }
    if (tid == 0) {
        const unsigned long long kk = key;
        if (kk) {
            best_d[set] = (int32_t)((uint32_t)(kk >> 32) ^ 0x80000000u);
            best_rank[set] = (int32_t)(0xFFFFFFFFu - (uint32_t)(kk & 0xFFFFFFFFu));
        } else {
            best_d[set] = 0;
            best_rank[set] = -1;
        }
    }
}

// Delta_A: out entries of the de-duplicated G1 only, a<->b skipped.
__global__ __launch_bounds__(256) void k_hillclimb(const int32_t *__restrict__ A, const int32_t *__restrict__ B,
                                                   int64_t count, int32_t *__restrict__ out,
                                                   const int32_t *__restrict__ ddptr, const int32_t *__restrict__ ddmid,
                                                   const int32_t *__restrict__ ddr, const int32_t *__restrict__ ddw,
                                                   const int16_t *__restrict__ Q, int N1, int ld) {
    const int64_t pair = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int gl = threadIdx.x & 15;
    int sum = 0;
    if (pair < count) {
        const int a = A[pair], b = B[pair];
        for (int side = 0; side < 2; side++) {
            const int x = side ? b : a, y = side ? a : b;
            for (int e = ddptr[x] + gl; e < ddmid[x]; e += 16) {
                const int t = ddr[e] - N1, w = ddw[e];
                if (t == y) continue;
                const int oldv = min(w, qv(Q, x, ld, t));
                const int newv = min(w, t == x ? qv(Q, y, ld, y) : qv(Q, y, ld, t));
                sum += newv - oldv;
            }
        }
    }
    sum = bm_wave::group_sum<16>(sum);
    if (pair < count && gl == 0) out[pair] = sum;
}

// ------------------------------------------------------------ launchers

int bm::launch_rot3(bm_ctx *c, const int32_t *dT, int64_t m, int mode, int32_t *dD, int32_t *dP) {
    if (m <= 0) return BM_OK;
    // the de-duplicated lists on the device (a no-op after the first call; a batched step's
    // configuration makes them before its capture)
    if (int rc = ensure_dd(c)) return rc;
    const int lanes = c->rot3_lanes;
    const unsigned blocks = (unsigned)((m * lanes + 255) / 256);
#define BM_ROT3(L)                                                                                                 \
    BM_LAUNCH(c, "rot3", hipLaunchKernelGGL(k_rot3<L>, dim3(blocks), dim3(256), 0, c->stream, dT, m, mode, dD, dP, \
                                            c->ddptr, c->ddr, c->ddw, c->Q, c->N1, c->ld))
    if (lanes == 8) BM_ROT3(8);
    else if (lanes == 32) BM_ROT3(32);
    else if (lanes == 16) BM_ROT3(16);
    else if (lanes == 64) BM_ROT3(64);
    else return set_error(BM_EINVAL, "rot3: %d lanes per triple", lanes);
#undef BM_ROT3
    return BM_OK;
}

// The table form: Xd for the current mapping (one dense all-pairs pass over
// the dense-form lists, made for every evaluation call since the mapping may
// have changed), then one thread per triple.  The G1 cell matrix is built once
// per graph pair.
// The dense-form lists and their G1 cell matrix (once per graph pair).
// The G1 cell matrix dfW1 of the dense-form lists.  When every G1 line is
// its own dense cell those lists are the comb lists, which are on the device
// already, so the cells come from them without building the dense-form lists.
int bm::ensure_df_cells(bm_ctx *c) {
    if (c->dfW1) return BM_OK;
    int rc;
    const int32_t *ptr = c->cptr, *mid = c->cmid;
    const int2 *rw = (const int2 *)c->crw;
    if (!c->dense_equal) {
        if ((rc = ensure_dense_lists(c))) return rc;
        ptr = c->dfptr, mid = c->dfmid, rw = (const int2 *)c->dfrw;
    }
    const size_t N1 = (size_t)c->N1, ld = (size_t)c->ld;
    if ((rc = dev_alloc(c, (void **)&c->dfW1, N1 * ld * 2))) return rc;
    const int64_t cells = (int64_t)(N1 * ld);
    BM_LAUNCH(c, "df_cells", hipLaunchKernelGGL(k_zero16, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0,
                                                c->stream, c->dfW1, cells));
    BM_LAUNCH(c, "df_cells", hipLaunchKernelGGL(k_df_cells, dim3((unsigned)((N1 + 3) / 4)), dim3(256), 0, c->stream,
                                                c->dfW1, ptr, mid, rw, c->N1, c->ld));
    return BM_OK;
}

static int rot3_table_prepare(bm_ctx *c) {
    int rc = ensure_dense_lists(c);
    if (rc || (rc = ensure_df_cells(c))) return rc;
    const size_t N1 = (size_t)c->N1, ld = (size_t)c->ld;
    if (!c->dfX && (rc = dev_alloc(c, (void **)&c->dfX, (N1 + 63) / 64 * 64 * ld * 4))) return rc;
    if (!c->dfXP && (rc = dev_alloc(c, (void **)&c->dfXP, N1 * ld * 8))) return rc;
    if ((rc = launch_list_gain(c, c->dfptr, c->dfrw, c->dfX, c->dfmaxw))) return rc;
    const int64_t pairs = (int64_t)N1 * (int64_t)(ld / 2);
    BM_LAUNCH(c, "table_pack", hipLaunchKernelGGL(k_table_pack, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0,
                                                  c->stream, (int2 *)c->dfXP, c->dfX, c->Q, c->dfW1, c->N1, c->ld));
    return BM_OK;
}

int bm::launch_rot3_x(bm_ctx *c, const int32_t *dT, int64_t m, int mode, int32_t *dD, int32_t *dP, bool via_map) {
    if (m <= 0) return BM_OK;
    BM_LAUNCH(c, "rot3_x", hipLaunchKernelGGL(k_rot3_x, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream,
                                              dT, m, mode, dD, dP, c->X, (int)(c->ldx >> 6), via_map ? c->W2 : c->Q,
                                              c->dfW1, c->ld, via_map ? c->map : nullptr));
    return BM_OK;
}

static int launch_rot3_table(bm_ctx *c, const int32_t *dT, int64_t m, int mode, int32_t *dD, int32_t *dP) {
    if (m <= 0) return BM_OK;
    BM_LAUNCH(c, "rot3_table", hipLaunchKernelGGL(k_rot3_table, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                                                  c->stream, dT, m, mode, dD, dP, (const int2 *)c->dfXP, c->ld));
    return BM_OK;
}

int bm::launch_permset(bm_ctx *c, int k, const int32_t *dS, int64_t m, int32_t *dD, int32_t *dR, int *dOv, int mult,
                       bool from_x, bool via_map) {
    if (m <= 0) return BM_OK;
    const int32_t *xm = from_x ? c->X : nullptr;
    const int32_t *pm = from_x && via_map ? c->map : nullptr;
#define BM_PERMSET(T, K)                                                                                        \
    BM_LAUNCH(c, "permset", hipLaunchKernelGGL((k_permset<T, K>), dim3((unsigned)m), dim3(T), 0, c->stream, dS, m, \
                                               k, dD, dR, c->cptr, c->cr, c->cw, c->Q, c->N1, c->ld, dOv, mult, xm, \
                                               (int)(c->ldx >> 6), c->dfW1, c->W2, pm))
    if (c->permset_threads != 64 && c->permset_threads != 256)
        return set_error(BM_EINVAL, "permset: %d threads per set", c->permset_threads);
    if (k == 8) {
        if (c->permset_threads == 64) BM_PERMSET(64, 8);
        else BM_PERMSET(256, 8);
    } else {
        if (c->permset_threads == 64) BM_PERMSET(64, 0);
        else BM_PERMSET(256, 0);
    }
#undef BM_PERMSET
    return BM_OK;
}

// ------------------------------------------------------------ autotuner
//
// Lanes per triple (8 / 16 / 32 / 64, i.e. 8, 4, 2 or 1 triples per wavefront) and
// threads per set (64 / 256) are timed with HIP events on the caller's own
// candidate list (one warm-up and three timed launches per variant) and the
// fastest is kept per instance class (log2 n, log2 mean list length, k) for
// the whole process.  Every variant computes the same integers.  The
// environment variables BM_ROT3_LANES / BM_PERMSET_THREADS force a variant.

static std::mutex g_mv_mu;
static std::map<std::tuple<int, int, int, int>, int> g_mv_tune;

static int log2c(long long x) {
    int r = 0;
    while ((1ll << r) < x) r++;
    return r;
}

template <class F>
static int time_launches(bm_ctx *c, F launch, float *ms) {
    hipEvent_t e0, e1;
    BM_HIP(hipEventCreate(&e0));
    BM_HIP(hipEventCreate(&e1));
    const bool was = c->timing;
    c->timing = false;
    int rc = launch();
    if (!rc) {
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < 3 && !rc; r++) rc = launch();
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(ms, e0, e1);
    }
    c->timing = was;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

template <class Set, class Launch>
static int tune_variant(bm_ctx *c, const char *env, std::tuple<int, int, int, int> key, const int *choices, int nch,
                        Set set, Launch launch) {
    if (const char *f = getenv(env)) {
        set(atoi(f));
        return BM_OK;
    }
    {
        std::lock_guard<std::mutex> g(g_mv_mu);
        auto it = g_mv_tune.find(key);
        if (it != g_mv_tune.end()) {
            set(it->second);
            return BM_OK;
        }
    }
    float best_ms = 1e30f;
    int best = choices[0];
    for (int i = 0; i < nch; i++) {
        set(choices[i]);
        float ms = 0.f;
        int rc = time_launches(c, launch, &ms);
        if (rc) return rc;
        if (ms < best_ms) {
            best_ms = ms;
            best = choices[i];
        }
    }
    set(best);
    std::lock_guard<std::mutex> g(g_mv_mu);
    g_mv_tune[key] = best;
    return BM_OK;
}

int bm::tune_rot3(bm_ctx *c, const int32_t *dT, int64_t m, int32_t *dD, int32_t *dP) {
    static const int ch[4] = {8, 16, 32, 64};
    const auto key = std::make_tuple(1, log2c(c->n), log2c(std::max<long long>(1, c->ndd / c->N1)), 3);
    int rc = tune_variant(c, "BM_ROT3_LANES", key, ch, 4, [&](int v) { c->rot3_lanes = v; },
                          [&] { return launch_rot3(c, dT, m, BM_ROT3_ALL, dD, dP); });
    if (!rc) c->rot3_tuned = true;
    return rc;
}

// Direct kernel or table form for a triple list of m entries: the table form
// pays one all-pairs pass per call and then 9 loads per triple, the direct
// kernel three loads per list entry, so long lists and large batches favour
// the table.  On the first list of each size class of an instance class the
// direct kernel, the pass and the table kernel are timed (the kernels on the
// first kTuneSample triples, then scaled to m) and the cheaper total kept;
// BM_ROT3_TABLE = 0 / 1 forces one.  The table needs (n+1) x ld x 14 bytes,
// so it is only offered up to kTableMaxBytes.
constexpr size_t kTableMaxBytes = (size_t)8 << 30;
constexpr int64_t kTuneSample = 1 << 17;

int bm::choose_rot3_table(bm_ctx *c, const int32_t *dT, int64_t m, int32_t *dD, int32_t *dP, int *table) {
    *table = 0;
    int rc = BM_OK;
    const int64_t ms = std::min(m, kTuneSample);
    if (!c->rot3_tuned && m >= kTuneMinTriples && (rc = tune_rot3(c, dT, ms, dD, dP))) return rc;
    if ((size_t)c->N1 * c->ld * 14 > kTableMaxBytes) return BM_OK;
    const char *f = getenv("BM_ROT3_TABLE");
    if (f || m < kTuneMinTriples) {
        *table = f && atoi(f) == 1;
        return BM_OK;
    }
    const auto key = std::make_tuple(4, log2c(c->n) * 64 + log2c(m), log2c(std::max<long long>(1, c->ndd / c->N1)), 3);
    {
        std::lock_guard<std::mutex> g(g_mv_mu);
        auto it = g_mv_tune.find(key);
        if (it != g_mv_tune.end()) {
            *table = it->second;
            return BM_OK;
        }
    }
    float t_dir = 0.f, t_pass = 0.f, t_tab = 0.f;
    if ((rc = time_launches(c, [&] { return launch_rot3(c, dT, ms, BM_ROT3_ALL, dD, dP); }, &t_dir)) ||
        (rc = time_launches(c, [&] { return rot3_table_prepare(c); }, &t_pass)) ||
        (rc = time_launches(c, [&] { return launch_rot3_table(c, dT, ms, BM_ROT3_ALL, dD, dP); }, &t_tab)))
        return rc;
    const double scale = (double)m / (double)ms;
    *table = (double)t_pass + t_tab * scale < t_dir * scale;
    std::lock_guard<std::mutex> g(g_mv_mu);
    g_mv_tune[key] = *table;
    return BM_OK;
}

int bm::launch_rot3_moves(bm_ctx *c, const int32_t *dT, int64_t m, int mode, int32_t *dD, int32_t *dP, int table,
                          bool prepare) {
    if (!table) return launch_rot3(c, dT, m, mode, dD, dP);
    int rc = prepare ? rot3_table_prepare(c) : BM_OK;
    return rc ? rc : launch_rot3_table(c, dT, m, mode, dD, dP);
}

int bm::tune_permset(bm_ctx *c, int k, const int32_t *dS, int64_t m, int32_t *dD, int32_t *dR, int *dOv, int mult,
                     bool from_x) {
    static const int ch[2] = {64, 256};
    const auto key = std::make_tuple(2, log2c(c->n), log2c(std::max<long long>(1, c->ncomb / c->N1)), k);
    int rc = tune_variant(c, "BM_PERMSET_THREADS", key, ch, 2, [&](int v) { c->permset_threads = v; },
                          [&] { return launch_permset(c, k, dS, m, dD, dR, dOv, mult, from_x); });
    if (!rc) c->permset_threads_k = k;
    return rc;
}

// ------------------------------------------------------------ C ABI

extern "C" int bm_eval_rot3(bm_ctx *c, const int32_t *triples, int64_t count, int mode, int32_t *best_delta,
                            int32_t *best_perm) {
    if (!c || (count > 0 && (!triples || !best_delta || !best_perm))) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (mode != BM_ROT3_ROTATIONS && mode != BM_ROT3_ALL) return set_error(BM_EINVAL, "bad mode %d", mode);
    if (count < 0) return set_error(BM_EINVAL, "count < 0");
    for (int64_t i = 0; i < count; i++) {
        const int32_t *t = triples + 3 * i;
        for (int j = 0; j < 3; j++)
            if (t[j] < 1 || t[j] > c->n) return set_error(BM_EINVAL, "triple %lld: vertex %d", (long long)i, t[j]);
        if (t[0] == t[1] || t[0] == t[2] || t[1] == t[2])
            return set_error(BM_EINVAL, "triple %lld has repeated vertices", (long long)i);
    }
    if (count == 0) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    const int64_t chunk = 1 << 22;
    int rc = ensure_stage(c, (size_t)chunk * 20);
    if (rc) return rc;
    int32_t *dT = (int32_t *)c->stage, *dD = dT + 3 * chunk, *dP = dD + chunk;
    int table = 0;
    for (int64_t off = 0; off < count; off += chunk) {
        const int64_t m = std::min(chunk, count - off);
        BM_HIP(hipMemcpyAsync(dT, triples + 3 * off, m * 12, hipMemcpyHostToDevice, c->stream));
        if (off == 0 && (rc = choose_rot3_table(c, dT, m, dD, dP, &table))) return rc;
        // the mapping is the same for every chunk: the table form's pass runs once
        if ((rc = launch_rot3_moves(c, dT, m, mode, dD, dP, table, off == 0))) return rc;
        BM_HIP(hipMemcpyAsync(best_delta + off, dD, m * 4, hipMemcpyDeviceToHost, c->stream));
        BM_HIP(hipMemcpyAsync(best_perm + off, dP, m * 4, hipMemcpyDeviceToHost, c->stream));
        BM_HIP(hipStreamSynchronize(c->stream));
    }
    return BM_OK;
}

extern "C" int bm_eval_permset(bm_ctx *c, int32_t k, const int32_t *sets, int64_t count, int32_t *best_delta,
                               int32_t *best_rank) {
    if (!c || (count > 0 && (!sets || !best_delta || !best_rank))) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (k < 2 || k > 8) return set_error(BM_EINVAL, "k=%d outside [2,8]", k);
    if (count < 0) return set_error(BM_EINVAL, "count < 0");
    for (int64_t i = 0; i < count; i++) {
        const int32_t *s = sets + (size_t)k * i;
        for (int j = 0; j < k; j++) {
            if (s[j] < 1 || s[j] > c->n) return set_error(BM_EINVAL, "set %lld: vertex %d", (long long)i, s[j]);
            for (int q = 0; q < j; q++)
                if (s[q] == s[j]) return set_error(BM_EINVAL, "set %lld repeats vertex %d", (long long)i, s[j]);
        }
    }
    if (count == 0) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    const int64_t chunk = 1 << 20;
    int rc = ensure_stage(c, (size_t)chunk * (k + 2) * 4 + 64);
    if (rc) return rc;
    int32_t *dS = (int32_t *)c->stage, *dD = dS + (size_t)k * chunk, *dR = dD + chunk;
    int *dOv = dR + chunk;
    BM_HIP(hipMemsetAsync(dOv, 0, 4, c->stream));
    for (int64_t off = 0; off < count; off += chunk) {
        const int64_t m = std::min(chunk, count - off);
        BM_HIP(hipMemcpyAsync(dS, sets + (size_t)k * off, (size_t)m * k * 4, hipMemcpyHostToDevice, c->stream));
        if (c->permset_threads_k != k && m >= kTuneMinSets && (rc = tune_permset(c, k, dS, m, dD, dR, dOv, 2)))
            return rc;
        if ((rc = launch_permset(c, k, dS, m, dD, dR, dOv))) return rc;
        BM_HIP(hipMemcpyAsync(best_delta + off, dD, m * 4, hipMemcpyDeviceToHost, c->stream));
        BM_HIP(hipMemcpyAsync(best_rank + off, dR, m * 4, hipMemcpyDeviceToHost, c->stream));
    }
    int ov = 0;
    BM_HIP(hipMemcpyAsync(&ov, dOv, 4, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    if (ov) return set_error(BM_ERANGE, "a set has more than %d intra-set G1 lines", kMaxIntra);
    return BM_OK;
}

extern "C" int bm_perm_from_rank(int32_t k, int32_t rank, int32_t *perm) {
    if (k < 1 || k > 8 || !perm) return set_error(BM_EINVAL, "bad k or NULL perm");
    int total = 1;
    for (int i = 2; i <= k; i++) total *= i;
    if (rank < 0 || rank >= total) return set_error(BM_EINVAL, "rank %d outside [0,%d)", rank, total);
    int pool[8];
    for (int i = 0; i < k; i++) pool[i] = i;
    int left = k, rem = rank;
    for (int i = 0; i < k; i++) {
        int f = 1;
        for (int q = 2; q <= k - 1 - i; q++) f *= q;
        const int idx = rem / f;
        rem -= idx * f;
        perm[i] = pool[idx];
        for (int q = idx; q < left - 1; q++) pool[q] = pool[q + 1];
        left--;
    }
    return BM_OK;
}

extern "C" int bm_eval_swaps_hillclimb(bm_ctx *c, const int32_t *a, const int32_t *b, int64_t count,
                                       int32_t *delta) {
    if (!c || (count > 0 && (!a || !b || !delta))) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    if (count < 0) return set_error(BM_EINVAL, "count < 0");
    for (int64_t i = 0; i < count; i++)
        if (a[i] < 0 || a[i] > c->n || b[i] < 0 || b[i] > c->n || a[i] == b[i])
            return set_error(BM_EINVAL, "pair %lld: (%d,%d) invalid", (long long)i, a[i], b[i]);
    if (count == 0) return BM_OK;
    BM_HIP(hipSetDevice(c->device));
    if (int drc = ensure_dd(c)) return drc;
    const int64_t chunk = 1 << 22;
    int rc = ensure_stage(c, (size_t)chunk * 12);
    if (rc) return rc;
    int32_t *dA = (int32_t *)c->stage, *dB = dA + chunk, *dD = dB + chunk;
    for (int64_t off = 0; off < count; off += chunk) {
        const int64_t m = std::min(chunk, count - off);
        BM_HIP(hipMemcpyAsync(dA, a + off, m * 4, hipMemcpyHostToDevice, c->stream));
        BM_HIP(hipMemcpyAsync(dB, b + off, m * 4, hipMemcpyHostToDevice, c->stream));
        BM_LAUNCH(c, "hillclimb", hipLaunchKernelGGL(k_hillclimb, dim3((unsigned)((m * 16 + 255) / 256)), dim3(256), 0,
                                                     c->stream, dA, dB, m, dD, c->ddptr, c->ddmid, c->ddr, c->ddw,
                                                     c->Q, c->N1, c->ld));
        BM_HIP(hipMemcpyAsync(delta + off, dD, m * 4, hipMemcpyDeviceToHost, c->stream));
        BM_HIP(hipStreamSynchronize(c->stream));
    }
    return BM_OK;
}
