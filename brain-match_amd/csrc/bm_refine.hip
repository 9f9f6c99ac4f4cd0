// Batched refinement: many exact improving moves per step, on the device.
//
// One step (all kernels on the context stream, captured once in a hipGraph
// and replayed):
//   1. candidates: the row-best swap of every row a (b > a) from the
//      all-pairs Delta_B sweep; `triples` 3-vertex moves drawn from 2-hop
//      G1 neighbourhoods, scored with Delta_G3 over p = 1..5; `sets` k-sets
//      grown along G1 lines, scored over all k! permutations with the intra
//      lines counted once (the exact score change);
//   2. keys: an improving candidate whose vertices all lie in [1, n] and are
//      "simple" (no self-loop, no repeated line touches them; there every
//      delta kind equals the exact change of S_opt) gets the priority key
//      delta << 32 | ~index;
//   3. selection, `rounds` times: each live candidate claims its vertices
//      (atomicMax of its key); a candidate whose closed G1 neighbourhood N[V]
//      holds no larger claim is scored exactly given the winners of the
//      earlier rounds (its lines to the vertices they moved corrected) and
//      wins if that change is positive; winners take their vertices, which
//      retires every candidate sharing one.  Two winners of one round are
//      never equal or G1-adjacent, so each round adds up exactly on top of
//      the earlier ones;
//   4. apply: every winner permutes the images of its vertices (and p^-1);
//      P and P^T are rebuilt from W2 and the new mapping, the G2 lists
//      re-expressed through p^-1;
//   5. verify: S_opt is recomputed from scratch on the device and must equal
//      the previous score plus the sum of the applied deltas (and rise when a
//      move was applied); the first failing step is recorded.
// No host round trip inside a step.  Selection and results do not depend on
// timing: the autotuned choices (gain kernel, lanes per triple, threads per
// set) only pick among variants that compute identical integers.
//
// Reference analogues: the per-step evaluate / host arg-max / apply loops of
// triple.cu:1082-1126, quick.cu:1175-1217 and
// graph-alignment-permset-parallel.c:651-730, which apply one move per step.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "bm_device.h"
#include "bm_internal.h"
#include "bm_wave.h"

using namespace bm;

namespace {

constexpr int kHist = 4096;  // per-step history ring
constexpr int kMaxRounds = 64;  // selection rounds per step
constexpr int kXcolBlocks = 512;  // workgroups of the column recompute (grid-stride over its chunks)
constexpr int kXcolLines = 256;   // G2 lines of a column's new image per chunk (at least)
constexpr int kXcolSpare = 512;   // chunk slots beyond one per moved vertex
constexpr int kXdeltaQual = 256;  // entries with a moved neighbour a row update keeps in LDS

struct RefineState {
    unsigned long long step;   // steps completed in this run
    long long score;           // exact S_opt after the last step
    long long moves;           // moves applied in this run
    long long kind_moves[3];   // of which pairs, triples, sets
    long long gain_total;      // sum of applied deltas in this run
    unsigned long long gain;   // this step
    unsigned int applied;      // this step
    unsigned int nwin;         // winner-list length this step
    unsigned int bad_step;     // 1 + first step whose exact check failed (0: none)
    unsigned int log_count;
    unsigned long long score_slot;
    long long idle;            // consecutive steps without a move
    long long initial_score;
    unsigned long long gen_step;  // steps since bm_refine_configure: seeds the candidate generators
    // incremental update of the all-pairs gain matrix X between steps
    int xstale;   // 1: X does not match the mapping, the next sweep recomputes it whole
    int inc;      // this step updates X incrementally (few moved vertices)
    int pinc;     // this step patches the moved rows and columns of P / P^T only
    int nu;       // moved vertices this step
    int npairs;   // their comb entries (prefix ubase)
    int full;     // this step began with a whole sweep
    int xchunks;  // work items of the column recompute this step (prefix cbase)
    int xlines;   // G2 lines per work item of the column recompute
    // steps that read W2 through the mapping leave P / P^T behind it
    int pstale;    // P / P^T do not match the mapping
    int prebuild;  // this step rebuilds P / P^T before its whole sweep (X and P both stale)
    unsigned long long score_change;  // exact S_opt change of this step over the lines at U
    int nhash, ndense;  // the row updates by kind
    int keep_keys;      // the previous step of this run moved nothing: the row-best keys still hold
    int nlive[kMaxRounds];  // per selection round: candidates in its list (each has claimed for the round)
};

struct MoveRec {
    int32_t step, kind, k, delta, round, base;
    int32_t v[8];
    int32_t img[8];  // images after the move
};

// what the candidate index space holds: [0, N1) rows, then triples, then sets
struct CandView {
    const unsigned long long *rowkeys;
    const int32_t *tri;
    const int32_t *sets;
    const int32_t *tri_d, *tri_p, *set_d, *set_r;
    int N1, n, T, S, k, pairs;
};

__constant__ int kTab3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ unsigned long long rnd(unsigned long long seed, unsigned long long step, long long i,
                                                  int j) {
    return mix64(seed ^ mix64(step * 0x100000001B3ull + 0x51ED27ull) ^ mix64((unsigned long long)i * 16 + j));
}

__device__ __forceinline__ int nbr(const int32_t *__restrict__ cptr, const int32_t *__restrict__ cr, int N1, int x,
                                   unsigned long long h) {
    const int deg = cptr[x + 1] - cptr[x];
    if (deg == 0) return -1;
    const int r = cr[cptr[x] + (int)(h % (unsigned long long)deg)];
    return r >= N1 ? r - N1 : r;
}

// vertices, delta and kind of candidate i; k = 0 when the slot is empty
__device__ __forceinline__ int cand_get(const CandView &cv, long long i, int *v, int &delta, int &kind) {
    if (i < cv.N1) {
        kind = 0;
        if (!cv.pairs || i == 0) return 0;
        const unsigned long long rk = cv.rowkeys[i];
        if (!rk) return 0;
        delta = (int)((unsigned)(rk >> 32) ^ 0x80000000u);
        v[0] = (int)i;
        v[1] = (int)(0xFFFFFFFFu - (unsigned)rk);
        return 2;
    }
    i -= cv.N1;
    if (i < cv.T) {
        kind = 1;
        delta = cv.tri_d[i];
        v[0] = cv.tri[3 * i];
        v[1] = cv.tri[3 * i + 1];
        v[2] = cv.tri[3 * i + 2];
        return 3;
    }
    i -= cv.T;
    kind = 2;
    delta = cv.set_d[i];
    for (int q = 0; q < cv.k; q++) v[q] = cv.sets[(long long)cv.k * i + q];
    return cv.k;
}

// The step's opening in one launch: the state fields reset (thread 0), with
// the P rebuild gate when the step reads W2 through the mapping (pgate: P and
// P^T are rebuilt before a whole sweep when both X and P are stale), and
// three word ranges zeroed (the claim keys, the taken words, the byte flags
// of the incremental update; a count of 0 skips one).  Volatile reads as in
// k_step_end.
__global__ void k_step_prologue(RefineState *st, int pgate, uint32_t *__restrict__ f0, long long n0,
                                uint32_t *__restrict__ f1, long long n1, uint32_t *__restrict__ f2, long long n2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        const volatile RefineState *vs = st;
        const int rebuild = pgate && vs->xstale && vs->pstale ? 1 : 0;
        st->keep_keys = vs->step > 0 && vs->applied == 0 ? 1 : 0;
        for (int r = 0; r < kMaxRounds; r++) st->nlive[r] = 0;
        st->gain = 0;
        st->applied = 0;
        st->nwin = 0;
        st->score_slot = 0;
        st->score_change = 0;
        st->nhash = 0;
        st->ndense = 0;
        st->nu = 0;
        st->full = 1;  // cleared at the candidate keys when the step's sweep had nothing to do
        if (pgate) {
            st->prebuild = rebuild;
            if (rebuild) st->pstale = 0;
        }
    }
    if (i < n0) f0[i] = 0;
    if (i < n1) f1[i] = 0;
    if (i < n2) f2[i] = 0;
}

// 3-vertex candidates: u uniform, v a G1 neighbour of u, w a neighbour of v or u
__global__ void k_gen_triples(int32_t *__restrict__ tri, int T, const RefineState *__restrict__ st,
                              unsigned long long seed, const int32_t *__restrict__ cptr,
                              const int32_t *__restrict__ cr, int N1, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    const unsigned long long step = st->gen_step;
    const unsigned long long h0 = rnd(seed, step, i, 0), h1 = rnd(seed, step, i, 1), h2 = rnd(seed, step, i, 2);
    const int u = 1 + (int)(h0 % (unsigned long long)n);
    int v = nbr(cptr, cr, N1, u, h1);
    if (v < 0) v = 1 + (int)(h1 % (unsigned long long)n);
    const int x = (h2 & 1) ? v : u;
    int w = nbr(cptr, cr, N1, x, h2 >> 1);
    if (w < 0) w = 1 + (int)((h2 >> 1) % (unsigned long long)n);
    tri[3 * i] = u;
    tri[3 * i + 1] = v;
    tri[3 * i + 2] = w;
}

// k-set candidates from a uniform seed vertex, grown along G1 lines
// (uniform = 0) or drawn uniformly (uniform = 1)
__global__ void k_gen_sets(int32_t *__restrict__ sets, int S, int k, int uniform, const RefineState *__restrict__ st,
                           unsigned long long seed, const int32_t *__restrict__ cptr, const int32_t *__restrict__ cr,
                           int N1, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const unsigned long long step = st->gen_step;
    int s[8];
    int j = 0, draw = 0;
    s[j++] = 1 + (int)(rnd(seed ^ 0xA5A5ull, step, i, draw++) % (unsigned long long)n);
    for (int tries = 0; j < k && tries < 6 * k; tries++) {
        const unsigned long long h = rnd(seed ^ 0xA5A5ull, step, i, draw++);
        const int m = s[(int)(h % (unsigned long long)j)];
        int c = !uniform && tries < 4 * k ? nbr(cptr, cr, N1, m, h >> 8) : -1;
        if (c < 0) c = 1 + (int)((h >> 8) % (unsigned long long)n);
        bool fresh = c >= 1 && c <= n;
        for (int q = 0; q < j && fresh; q++) fresh = s[q] != c;
        if (fresh) s[j++] = c;
    }
    for (; j < k; j++) s[j] = s[0];  // left invalid: repeated vertex
    for (int q = 0; q < k; q++) sets[(long long)k * i + q] = s[q];
}

// Each live candidate also makes its first-round claim: the atomicMax of its
// key on each of its vertices (vmax0).  With inc (the incremental update),
// thread 0 also records whether the step's sweep ran whole (X was stale) and
// marks X current.
__global__ void k_cand_keys(unsigned long long *__restrict__ ckey, long long M, CandView cv,
                            const uint8_t *__restrict__ vsimple, RefineState *__restrict__ st, int inc,
                            unsigned long long *__restrict__ vmax0, int32_t *__restrict__ list0) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (inc && i == 0) {
        const int stale = ((const volatile RefineState *)st)->xstale;  // as in k_step_end
        st->full = stale;  // the gain kernels ran whole iff X was stale
        st->xstale = 0;
    }
    int v[8], delta = 0, kind = 0;
    const int k = i < M ? cand_get(cv, i, v, delta, kind) : 0;
    bool ok = k > 0 && delta > 0;
    for (int q = 0; q < k && ok; q++) {
        ok = v[q] >= 1 && v[q] <= cv.n && vsimple[v[q]];
        for (int r = 0; r < q && ok; r++) ok = v[r] != v[q];
    }
    const unsigned long long key = ok ? ((unsigned long long)(unsigned)delta << 32) | (0xFFFFFFFFu - (unsigned)i) : 0ull;
    if (i < M) ckey[i] = key;
    // the first round's list: one atomic per wave
    const unsigned long long m = __ballot(ok);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(&st->nlive[0], __popcll(m));
    base = __shfl(base, 0);
    if (!ok) return;
    list0[base + __popcll(m & ((1ull << lane) - 1))] = (int32_t)i;
    for (int q = 0; q < k; q++) atomicMax(&vmax0[v[q]], key);
}

__device__ __forceinline__ void unrank(int k, int rank, int *perm) {
    int pool[8];
    for (int q = 0; q < k; q++) pool[q] = q;
    int left = k;
    for (int q = 0; q < k; q++) {
        int f = 1;
        for (int t = 2; t <= k - 1 - q; t++) f *= t;
        const int idx = rank / f;
        rank -= idx * f;
        perm[q] = pool[idx];
        for (int t = idx; t < left - 1; t++) pool[t] = pool[t + 1];
        left--;
    }
}

// the move of candidate i: new p(v[q]) = old p(v[src[q]])
__device__ __forceinline__ void move_src(const CandView &cv, long long i, int kind, int k, int *src) {
    if (kind == 0) {
        src[0] = 1;
        src[1] = 0;
    } else if (kind == 1) {
        const int p = cv.tri_p[i - cv.N1];
        for (int q = 0; q < 3; q++) src[q] = kTab3[p][q];
    } else {
        unrank(k, cv.set_r[i - cv.N1 - cv.T], src);
    }
}

// Selection round `round`, one wave per live candidate (grid-stride); every
// live candidate has claimed its vertices in vm (the round's claims, made at
// the candidate keys for the first round and by the losers of the previous
// round for the others).  A candidate sharing a vertex with an earlier
// winner (taken) retires.  Its claim holds when no vertex of N[V] carries a
// larger claim: the own vertices are checked first (one lane each), then
// their G1 lists in one pass with kDecideUnroll entries in flight per lane;
// the wave stops at the first larger claim it sees, and the loser claims its
// vertices for the next round in vmn.  (A claim made there by a candidate
// that the next round retires may make a neighbour lose that round too:
// losing is always safe, the neighbour claims again.)  The same pass scores
// the candidate given the winners of the earlier rounds of the step: its
// delta assumed every other vertex at its image before the step, so each G1
// line from V to a vertex y an earlier winner moved (taken[y] = (y's new
// image + 1, its old image | the round << 20); never a vertex of V) changes
// it by
//   t(p'x, p'y) - t(px, p'y) - t(p'x, py) + t(px, py),
// t(a, b) = min(w, W2[a][b]) for a line x -> y, min(w, W2[b][a]) for y -> x.
// It wins when that exact change is positive: its vertices are taken with
// their new images, and (candidate, change, round) join the winner list
// (win, win + max_win, win + 2 max_win).  Two winners of one round are
// never equal or G1-adjacent (the smaller key would have seen the larger
// claim), so the winners of a round add up exactly on top of the earlier
// rounds, and the step's gain is the sum of the changes.
constexpr int kDecideUnroll = 8;       // list entries in flight per lane in k_decide's scan
constexpr int kDecideLoserCap = 256;   // losers a k_decide block holds in LDS before it appends them
constexpr int kDecideWave = 1024;      // list entries a wave scans; a candidate with more takes the workgroup

struct DecideArgs {
    unsigned long long *ckey;
    CandView cv;
    const unsigned long long *vmax;
    unsigned long long *vmn;
    int2 *taken;
    int32_t *win;
    int max_win, round;
    RefineState *st;
    const int32_t *cptr;
    const int2 *crw;
    const int32_t *map;
    const int16_t *W2;
    int ld, N1;
};

// One candidate decided by a group of kG threads (a wave, 64, or the
// workgroup, 256): votes by ballot or __syncthreads_or, the correction
// summed over the group.  A wave hands a candidate with more than
// kDecideWave list entries to its workgroup (bigq) instead.  Losers are
// queued in lq for the next round.
template <int kG>
__device__ __forceinline__ void decide_one(const DecideArgs &A, long long i, int g, int *lq, int *lqn, long long *bigq,
                                           int *nbig, int *red) {
    const bool block = kG == 256;
    auto vote = [&](bool x) { return block ? __syncthreads_or(x) != 0 : __ballot(x) != 0ull; };
    const unsigned long long key = A.ckey[i];
    int v[8], delta, kind;
    const int k = cand_get(A.cv, i, v, delta, kind);
    if (!block) {
        int ent = 0;
        for (int q = 0; q < k; q++) ent += A.cptr[v[q] + 1] - A.cptr[v[q]];
        if (ent > kDecideWave) {
            if (g == 0) bigq[atomicAdd(nbig, 1)] = i;
            return;
        }
    }
    int mine = -1;
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (q == g && q < k) mine = v[q];
    // shares a vertex with a winner of an earlier round (this round's winners are left to the claims,
    // so that what a group sees does not depend on timing)
    if (vote(mine >= 0 && A.taken[mine].x && (A.taken[mine].y >> 20) < A.round)) {
        if (g == 0) A.ckey[i] = 0;
        return;
    }
    auto claim_next = [&]() {
        if (A.vmn && mine >= 0) atomicMax(&A.vmn[mine], key);
        if (A.vmn && g == 0) lq[atomicAdd(lqn, 1)] = (int32_t)i;
    };
    if (vote(mine >= 0 && A.vmax[mine] > key)) {
        claim_next();
        return;
    }
    int src[8];
    move_src(A.cv, i, kind, k, src);
    // one pass over N(V): the claims, and the lines to vertices moved in earlier rounds
    int corr = 0;
    bool lost = false;
    const int N1 = A.N1;
    for (int q = 0; q < k && !lost; q++) {
        const int x = v[q], px = A.map[x], pn = A.map[v[src[q]]], e1 = A.cptr[x + 1];
        for (int eb = A.cptr[x]; eb < e1; eb += kG * kDecideUnroll) {  // uniform trip count: the vote sees the group
            const int e = eb + g;
            int2 rw[kDecideUnroll];
#pragma unroll
            for (int j = 0; j < kDecideUnroll; j++) rw[j] = e + kG * j < e1 ? A.crw[e + kG * j] : make_int2(-1, 0);
            bool lose = false;
            int2 t[kDecideUnroll];
#pragma unroll
            for (int j = 0; j < kDecideUnroll; j++) {
                t[j] = make_int2(0, 0);
                if (rw[j].x < 0) continue;
                const int y = rw[j].x >= N1 ? rw[j].x - N1 : rw[j].x;
                lose |= A.vmax[y] > key;
                t[j] = A.taken[y];
                if ((t[j].y >> 20) >= A.round) t[j].x = 0;  // only earlier rounds
            }
            if (vote(lose)) {
                lost = true;
                break;
            }
#pragma unroll
            for (int j = 0; j < kDecideUnroll; j++) {
                if (!t[j].x) continue;
                const bool out = rw[j].x >= N1;
                const int yn = t[j].x - 1, py = t[j].y & 0xFFFFF, w = rw[j].y;
                auto tw = [&](int a, int b) {
                    return min(w, (int)(out ? A.W2[(size_t)a * A.ld + b] : A.W2[(size_t)b * A.ld + a]));
                };
                corr += tw(pn, yn) - tw(px, yn) - tw(pn, py) + tw(px, py);
            }
        }
    }
    if (lost) {
        claim_next();
        return;
    }
    corr = bm_wave::wave_sum(corr);
    if (block) {
        if ((g & 63) == 0) red[g >> 6] = corr;
        __syncthreads();
        corr = red[0] + red[1] + red[2] + red[3];
        __syncthreads();
    }
    const int change = delta + corr;
    if (g == 0) {
        A.ckey[i] = 0;
        if (change > 0) {
            for (int q = 0; q < k; q++) A.taken[v[q]] = make_int2(A.map[v[src[q]]] + 1, A.map[v[q]] | (A.round << 20));
            const unsigned s = atomicAdd(&A.st->nwin, 1u);
            A.win[s] = (int32_t)i;
            A.win[A.max_win + s] = change;
            A.win[2 * A.max_win + s] = A.round;
        }
    }
}

__global__ __launch_bounds__(256) void k_decide(DecideArgs A, const int32_t *__restrict__ list,
                                                int32_t *__restrict__ next) {
    // the losers of the round are gathered in LDS and appended to the next
    // round's list with one atomic per block and flush
    __shared__ int32_t lq[kDecideLoserCap];
    __shared__ int lqn, lqbase, nbig, red[4];
    __shared__ long long bigq[4];
    RefineState *st = A.st;
    const int round = A.round;
    auto flush = [&]() {  // block-uniform
        __syncthreads();
        const int cnt = lqn;
        if (threadIdx.x == 0 && cnt) lqbase = atomicAdd(&st->nlive[round + 1], cnt);
        __syncthreads();
        for (int j = threadIdx.x; j < cnt; j += 256) next[lqbase + j] = lq[j];
        __syncthreads();
        if (threadIdx.x == 0) lqn = 0;
        __syncthreads();
    };
    if (threadIdx.x == 0) lqn = nbig = 0;
    __syncthreads();
    const int nl = st->nlive[round];
    for (int t0 = blockIdx.x * 4; t0 < nl; t0 += gridDim.x * 4) {  // block-uniform trip count
        const int t = t0 + (int)(threadIdx.x >> 6);
        if (t < nl) decide_one<64>(A, list[t], threadIdx.x & 63, lq, &lqn, bigq, &nbig, red);
        __syncthreads();
        const int nb = nbig;
        for (int b = 0; b < nb; b++) decide_one<256>(A, bigq[b], threadIdx.x, lq, &lqn, bigq, &nbig, red);
        __syncthreads();  // every thread has read nbig
        if (threadIdx.x == 0) nbig = 0;
        __syncthreads();
        if (A.vmn && lqn > kDecideLoserCap - 8) flush();
    }
    if (A.vmn) flush();
}


// thread per winner: new p(v[q]) = old p(v[src[q]]); the sets are disjoint.
// The step's gain adds each winner's exact change (not its candidate delta).
// With ulists set (the incremental update of X), each vertex whose image
// changes is flagged in moved[] and gets its two G2 lists in ulists: entry
// 2x the out-list, 2x + 1 the in-list (its words at ng2 + g2iptr in the
// back-to-back translated lists), each as (start at the new image, length,
// start at the old image, length).
__global__ void k_apply(const int32_t *__restrict__ win, int max_win, CandView cv, int32_t *__restrict__ map,
                        int32_t *__restrict__ pinv, int with_pinv, RefineState *__restrict__ st,
                        MoveRec *__restrict__ log, int log_cap, uint8_t *__restrict__ moved,
                        int4 *__restrict__ ulists, const int32_t *__restrict__ g2optr,
                        const int32_t *__restrict__ g2iptr, int ng2) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= max_win || s >= (int)st->nwin) return;
    const long long i = win[s];
    int v[8], src[8], old[8], base = 0, kind = 0;
    const int k = cand_get(cv, i, v, base, kind);
    const int delta = win[max_win + s];  // the exact change given the earlier rounds' winners
    move_src(cv, i, kind, k, src);
    for (int q = 0; q < k; q++) old[q] = map[v[src[q]]];
    if (ulists) {
        for (int q = 0; q < k; q++) {
            const int gnew = old[q], gold = map[v[q]];  // map is still the image before the move
            if (gnew == gold) continue;
            moved[v[q]] = 1;
            ulists[2 * v[q]] = make_int4(g2optr[gnew], g2optr[gnew + 1] - g2optr[gnew], g2optr[gold],
                                         g2optr[gold + 1] - g2optr[gold]);
            ulists[2 * v[q] + 1] = make_int4(ng2 + g2iptr[gnew], g2iptr[gnew + 1] - g2iptr[gnew],
                                             ng2 + g2iptr[gold], g2iptr[gold + 1] - g2iptr[gold]);
        }
    }
    for (int q = 0; q < k; q++) {
        map[v[q]] = old[q];
        if (with_pinv && old[q] != 0) pinv[old[q]] = v[q];
    }
    atomicAdd(&st->gain, (unsigned long long)(unsigned)delta);
    atomicAdd(&st->applied, 1u);
    atomicAdd((unsigned long long *)&st->kind_moves[kind], 1ull);
    if (log_cap > 0) {
        const unsigned slot = atomicAdd(&st->log_count, 1u);
        if (slot < (unsigned)log_cap) {
            MoveRec &r = log[slot];
            r.step = (int32_t)st->step;
            r.kind = kind;
            r.k = k;
            r.delta = delta;
            r.round = win[2 * max_win + s];
            r.base = base;
            for (int q = 0; q < 8; q++) {
                r.v[q] = q < k ? v[q] : 0;
                r.img[q] = q < k ? old[q] : 0;
            }
        }
    }
}

// The single-thread kernels that rewrite the state read its fields through a
// volatile pointer: that makes the reads vector loads, which stay ordered with
// this thread's later vector stores to the same words.  A uniform plain load
// may be issued through the scalar cache, and a vector store that overwrites
// the word with a value that does not depend on the load can then land before
// the load has returned.
// by_change: the step's score is the previous one plus the exact change over
// the lines at the moved vertices (k_score_change), else the whole rescore.
__global__ void k_step_end(RefineState *st, bm_refine_step *hist, int by_change) {
    if (threadIdx.x != 0) return;
    const volatile RefineState *vs = st;
    const long long now = by_change ? vs->score + (long long)vs->score_change : (long long)vs->score_slot;
    const long long gain = (long long)vs->gain;
    const long long prev = vs->score;
    const unsigned applied = vs->applied;
    const unsigned long long step = vs->step;
    const bool ok = now == prev + gain && (applied == 0 || now > prev);
    bm_refine_step &h = hist[step % kHist];
    h.score = now;
    h.gain = gain;
    h.moves = (int32_t)applied;
    h.ok = ok ? 1 : 0;
    h.moved = vs->nu;
    h.full = vs->full;
    if (!ok && !vs->bad_step) st->bad_step = (unsigned)(step + 1);
    st->score = now;
    st->moves = vs->moves + applied;
    st->gain_total = vs->gain_total + gain;
    st->idle = applied ? 0 : vs->idle + 1;
    st->step = step + 1;
    st->gen_step = vs->gen_step + 1;
}

// keep_x: the mapping is the one the previous run ended with, so X (current
// unless that run's last step marked it stale) carries over.  Volatile reads
// as in k_step_end.
// pstale: P / P^T lag the mapping at the start of the run.
__global__ void k_state_init(RefineState *st, int keep_x, int pstale) {
    if (threadIdx.x != 0) return;
    const volatile RefineState *vs = st;
    const long long s = (long long)vs->score_slot;
    RefineState z = {};
    z.score = s;
    z.initial_score = s;
    z.gen_step = vs->gen_step;
    z.xstale = keep_x ? vs->xstale : 1;
    z.pstale = pstale;
    *st = z;
}

// ---------------------------------------------------------------- incremental X
//
// X(u,v) = C_u + sum over e in comb(u) of the G2 list of e's neighbour's
// image, in P coordinates (the sparse form of the all-pairs sweep), plus
// Corr(u,v) at u < v.  After a step that moved the vertex set U (images
// permuted within U):
//   - an entry of comb(u) whose neighbour x is in U now reads the list of
//     p'(x) instead of p(x): row u gains f over the new list and loses f over
//     the old one, in the columns outside U (k_xdelta_rows, per row; or
//     k_xdelta, per entry);
//   - a list target in p(U) now translates to another vertex of U, so the
//     columns of U are recomputed whole, pulled from the G2 neighbours of
//     their new images (k_xcol_plan, k_xcol, k_xcol_sum);
//   - Corr changes on the lines at U: subtracted before the move with the old
//     P, added after it with the new P (k_corr_inc);
//   - X(v,v) is read back into diag (k_score_change).
// f(w, q) = min(w, q) - min(w, 0).  Costs about 3 |U| / n of a sweep; steps
// that move more than inc_max vertices mark X stale instead.

__device__ __forceinline__ int block_scan_excl(int v, int *wsum, int &total) {  // 1024 threads
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (threadIdx.x < 16) {
        int t = wsum[threadIdx.x];
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(t, o, 16);
            if ((int)threadIdx.x >= o) t += y;
        }
        wsum[threadIdx.x] = t;
    }
    __syncthreads();
    const int incl = x + (wave ? wsum[wave - 1] : 0);
    total = wsum[15];
    __syncthreads();
    return incl - v;
}

// largest i < count with base[i] <= p (base ascending, base[0] = 0)
__device__ __forceinline__ int find_base(const int32_t *__restrict__ base, int count, int p) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One workgroup: the moved vertices (winners' vertices) with their old
// images, the U flags, the prefix of their comb-list lengths, the decision.
// The update is taken when U is small (nu <= inc_max) and its list work,
// estimated as sum over x in U of deg1(x) * 2 deg2(p(x)) + deg2(p(x)) * dbar
// (list swaps of the rows next to x, the pull of x's column), is at most
// inc_budget; otherwise X is marked stale for the next sweep.
__global__ __launch_bounds__(1024) void k_collect(const int32_t *__restrict__ win, CandView cv,
                                                  const int32_t *__restrict__ map, const int32_t *__restrict__ cptr,
                                                  const int32_t *__restrict__ g2optr,
                                                  const int32_t *__restrict__ g2iptr, int32_t *__restrict__ ulist,
                                                  int32_t *__restrict__ uold, int32_t *__restrict__ ubase,
                                                  uint8_t *__restrict__ inU, int32_t *__restrict__ uoldv,
                                                  RefineState *__restrict__ st, int inc_max, int pinc_max,
                                                  long long inc_budget, int dbar, int pfree, int N1) {
    __shared__ int cnt, carry;
    __shared__ int wsum[16];
    __shared__ unsigned long long work;
    if (threadIdx.x == 0) {
        cnt = 0;
        carry = 0;
        work = 0;
    }
    __syncthreads();
    const int nwin = (int)st->nwin;
    for (int w = threadIdx.x; w < nwin; w += 1024) {
        int v[8], d, kind;
        const int k = cand_get(cv, win[w], v, d, kind);
        atomicAdd(&cnt, k);
        unsigned long long wk = 0;
        for (int q = 0; q < k; q++) {
            const int g = map[v[q]];
            uoldv[v[q]] = g;
            inU[v[q]] = 1;
            const unsigned long long d2 = (g2iptr[g + 1] - g2iptr[g]) + (g2optr[g + 1] - g2optr[g]);
            wk += (unsigned long long)(cptr[v[q] + 1] - cptr[v[q]]) * 2 * d2 + d2 * (unsigned long long)dbar;
        }
        atomicAdd(&work, wk);
    }
    __syncthreads();
    const int nu = cnt;
    {
        // U in increasing vertex order, from the flags: each thread a contiguous range of vertices
        // (the column pass then writes neighbouring columns of a row together)
        const int per = (N1 + 1023) / 1024, v0 = min(N1, (int)threadIdx.x * per), v1 = min(N1, v0 + per);
        int mine = 0;
        for (int v = v0; v < v1; v++) mine += inU[v];
        int total;
        int k = block_scan_excl(mine, wsum, total);
        for (int v = v0; v < v1; v++)
            if (inU[v]) {
                ulist[k] = v;
                uold[k] = uoldv[v];
                k++;
            }
        __syncthreads();
    }
    for (int c0 = 0; c0 < nu; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const int deg = i < nu ? cptr[ulist[i] + 1] - cptr[ulist[i]] : 0;
        int total;
        const int ex = block_scan_excl(deg, wsum, total);
        if (i < nu) ubase[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ubase[nu] = carry;
        st->nu = nu;
        st->npairs = carry;
        const bool inc = nu <= inc_max && work <= (unsigned long long)inc_budget;
        st->inc = inc ? 1 : 0;
        if (pfree && nu > 0) st->pstale = 1;  // the step moves vertices and leaves P / P^T as they are
        st->pinc = nu <= pinc_max ? 1 : 0;
        if (!inc) st->xstale = 1;
    }
}


// Corr of every G1 line at U into X(min, max), times sign; a line between two
// vertices of U is taken from the smaller one's list only.  With rowmark set
// (the pass before the move), every neighbour u of a vertex of U is marked:
// row u of X is the one whose list entry mirrors this one (k_xdelta_rows).
__global__ void k_corr_inc(int sign, const RefineState *__restrict__ st, const int32_t *__restrict__ ulist,
                           const int32_t *__restrict__ ubase, const uint8_t *__restrict__ inU,
                           const int32_t *__restrict__ cptr, const int2 *__restrict__ crw,
                           const int32_t *__restrict__ cdw, const int16_t *__restrict__ Q, int ld, int N1,
                           int32_t *__restrict__ X, int64_t ldx, uint8_t *__restrict__ rowmark,
                           const int32_t *__restrict__ pmap) {
    if (!st->inc) return;
    const int nu = st->nu, total = st->npairs, T = (int)(ldx >> 6);
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const int xi = find_base(ubase, nu, p);
        const int x = ulist[xi];
        const int e = cptr[x] + (p - ubase[xi]);
        const int2 rw = crw[e];
        const bool out = rw.x >= N1;
        const int u = out ? rw.x - N1 : rw.x;
        if (rowmark) rowmark[u] = 1;
        if (u == x || (inU[u] && u < x)) continue;
        const int a = min(x, u), b = max(x, u);
        const bool out_a = x == a ? out : !out;
        // P[x][y], or W2[p(x)][p(y)] with pmap (Q is W2 then)
        const int ga = pmap ? pmap[a] : a, gb = pmap ? pmap[b] : b;
        const int corr = pair_corr(out_a, rw.y, cdw[e], (int)Q[(size_t)ga * ld + ga], (int)Q[(size_t)gb * ld + gb],
                                   (int)Q[(size_t)ga * ld + gb], (int)Q[(size_t)gb * ld + ga]);
        if (corr) atomicAdd(X + xidx(a, b, T), sign * corr);
    }
}

// The exact S_opt change of the step, from W2 and the images before and after
// it (not from the deltas or X), and X(v,v) read back into diag when X was
// updated in place: every G1 line with an end in U once (an in
// entry whose source is in U is that source's out entry), lines leaving
// vertex 0 skipped as the score skips them, min(w, W2[p'(s)][p'(d)]) -
// min(w, W2[p(s)][p(d)]).  Runs after the apply; uoldv holds p on U.
__global__ __launch_bounds__(256) void k_score_change(RefineState *__restrict__ st,
                                                      const int32_t *__restrict__ ulist,
                                                      const int32_t *__restrict__ ubase,
                                                      const uint8_t *__restrict__ inU,
                                                      const int32_t *__restrict__ uoldv,
                                                      const int32_t *__restrict__ map,
                                                      const int32_t *__restrict__ cptr, const int2 *__restrict__ crw,
                                                      const int16_t *__restrict__ W2, int ld, int N1,
                                                      const int32_t *__restrict__ X, int64_t ldx,
                                                      int32_t *__restrict__ diag) {
    __shared__ long long part[4];
    const int nu = st->nu, total = st->npairs;
    if (st->inc)  // X(v,v) back into diag after the incremental update
        for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < N1; v += gridDim.x * blockDim.x)
            diag[v] = X[xidx(v, v, (int)(ldx >> 6))];
    long long s = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const int xi = find_base(ubase, nu, p);
        const int x = ulist[xi];
        const int2 rw = crw[cptr[x] + (p - ubase[xi])];
        const bool out = rw.x >= N1;
        const int u = out ? rw.x - N1 : rw.x;
        if (!out && inU[u]) continue;  // counted from u's out entry
        const int src = out ? x : u, dst = out ? u : x;
        if (src == 0) continue;
        const int ns = map[src], nd = map[dst];
        const int os = inU[src] ? uoldv[src] : ns, od = inU[dst] ? uoldv[dst] : nd;
        s += min(rw.y, (int)W2[(size_t)ns * ld + nd]) - min(rw.y, (int)W2[(size_t)os * ld + od]);
    }
    s = bm_wave::wave_sum64(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const long long t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(&st->score_change, (unsigned long long)t);
    }
}

// The translated G2 words at the images of U after the apply: every line at
// the image g of a vertex x of U now translates g to x (the images are
// permuted within U, so no other word changes).  One workgroup per vertex of
// U: a line s -> g is an entry of s's out-list, g -> v one of v's in-list.
// With rowq set, each comb entry (x, u) of a moved x also counts, for row u
// of X, one moved-neighbour entry in rowq[u] and the words of its two G2
// lists in rowq[N1 + u] (u's mirrored entry: an out entry x -> u is the in
// entry of u, which reads the out-lists, ulists[2x]).
__global__ __launch_bounds__(256) void k_g2y_patch(const RefineState *__restrict__ st,
                                                   const int32_t *__restrict__ ulist,
                                                   const int32_t *__restrict__ map, uint32_t *__restrict__ g2oy,
                                                   uint32_t *__restrict__ g2iy, const uint32_t *__restrict__ g2o,
                                                   const uint32_t *__restrict__ g2i,
                                                   const int32_t *__restrict__ g2optr,
                                                   const int32_t *__restrict__ g2iptr,
                                                   const int32_t *__restrict__ g2o_at_in,
                                                   const int32_t *__restrict__ g2i_at_out,
                                                   const uint8_t *__restrict__ moved, const int4 *__restrict__ ulists,
                                                   const int32_t *__restrict__ cptr, const int2 *__restrict__ crw,
                                                   int N1, int32_t *__restrict__ rowq) {
    const int nu = st->nu;
    for (int i = blockIdx.x; i < nu; i += gridDim.x) {
        const int x = ulist[i], g = map[x];
        if (rowq && moved[x]) {
            const int4 Lo = ulists[2 * x], Li = ulists[2 * x + 1];
            const int wo = Lo.y + Lo.w, wi = Li.y + Li.w;
            for (int e = cptr[x] + threadIdx.x; e < cptr[x + 1]; e += 256) {
                const int r = crw[e].x;
                const int u = r >= N1 ? r - N1 : r;
                atomicAdd(rowq + u, 1);
                atomicAdd(rowq + N1 + u, r >= N1 ? wo : wi);
            }
        }
        const uint32_t hi = (uint32_t)x << 16;
        const int i0 = g2iptr[g], nin = g2iptr[g + 1] - i0;
        const int o0 = g2optr[g], nout = g2optr[g + 1] - o0;
        for (int l = threadIdx.x; l < nin + nout; l += 256) {
            if (l < nin) {
                const int j = g2o_at_in[i0 + l];
                g2oy[j] = hi | (g2o[j] & 0xFFFFu);
            } else {
                const int j = g2i_at_out[o0 + l - nin];
                g2iy[j] = hi | (g2i[j] & 0xFFFFu);
            }
        }
    }
}

// one wave per comb entry of U: the mirrored entry of the neighbour's row
// swaps the list of the old image for that of the new one
__global__ __launch_bounds__(256) void k_xdelta(const RefineState *__restrict__ st, const int32_t *__restrict__ ulist,
                                                const int32_t *__restrict__ uold, const int32_t *__restrict__ ubase,
                                                const uint8_t *__restrict__ inU, const int32_t *__restrict__ map,
                                                const int32_t *__restrict__ cptr, const int2 *__restrict__ crw,
                                                const int32_t *__restrict__ g2optr, const uint32_t *__restrict__ g2oy,
                                                const int32_t *__restrict__ g2iptr, const uint32_t *__restrict__ g2iy,
                                                int N1, int32_t *__restrict__ X, int64_t ldx) {
    if (!st->inc) return;
    const int nu = st->nu, total = st->npairs, T = (int)(ldx >> 6);
    const int lane = threadIdx.x & 63;
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    for (int p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < total; p += nwaves) {
        const int xi = find_base(ubase, nu, p);
        const int x = ulist[xi];
        const int gold = uold[xi], gnew = map[x];
        if (gold == gnew) continue;
        const int2 rw = crw[cptr[x] + (p - ubase[xi])];
        const bool out = rw.x >= N1;  // x -> u: u reads row x of P (G2 out-lists); u -> x: row x of P^T (in-lists)
        const int u = out ? rw.x - N1 : rw.x;
        const int w = rw.y, wneg = min(w, 0);
        const int32_t *ptr = out ? g2optr : g2iptr;
        const uint32_t *L = out ? g2oy : g2iy;
        for (int side = 0; side < 2; side++) {
            const int g = side ? gold : gnew, sg = side ? -1 : 1;
            const int j1 = ptr[g + 1];
            for (int j = ptr[g] + lane; j < j1; j += 256) {  // four list words in flight per lane
                uint32_t yv[4];
#pragma unroll
                for (int q = 0; q < 4; q++) yv[q] = j + 64 * q < j1 ? L[j + 64 * q] : 0xFFFF0000u;
                uint8_t in[4];
#pragma unroll
                for (int q = 0; q < 4; q++) in[q] = (yv[q] >> 16) == 0xFFFF ? 1 : inU[yv[q] >> 16];
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (!in[q])
                        atomicAdd(X + xidx(u, (int)(yv[q] >> 16), T),
                                  sg * (min(w, (int)(int16_t)(yv[q] & 0xffff)) - wneg));
            }
        }
    }
}

// The row updates split by size.  k_g2y_patch counts, from the lists of the
// moved vertices, each row's moved-neighbour entries and the words of their
// lists (rowq, roww); k_xrow_class (a thread per row) lists the rows with at
// most kHashQual such entries and kHashWords words, which k_xdelta_hash
// accumulates in an LDS hash table of kHashSlots (column, sum) pairs (256
// threads and about 22 KB of LDS per row, so seven rows are resident per CU,
// and a row touches only its hashed columns), and the other rows, which keep
// the (n+1)-int LDS row of k_xdelta_rows.
constexpr int kHashQual = 256;
constexpr int kHashWords = 1536;
constexpr int kHashSlots = 2048;  // more than kHashWords: probing always finds a slot

__global__ void k_xrow_class(RefineState *__restrict__ st, const int32_t *__restrict__ rowq,
                             const int32_t *__restrict__ roww, int N1, int32_t *__restrict__ hrows,
                             int32_t *__restrict__ drows) {
    if (!st->inc) return;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = u < N1 ? rowq[u] : 0;
    const bool small = q <= kHashQual && (unsigned)roww[u < N1 ? u : 0] <= (unsigned)kHashWords;
    const bool h = q > 0 && small, d = q > 0 && !small;
    // one atomic per wave and list on the two counters (thousands of rows a step)
    const unsigned long long mh = __ballot(h), md = __ballot(d);
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1;
    int bh = 0, bd = 0;
    if (lane == 0) {
        if (mh) bh = atomicAdd(&st->nhash, __popcll(mh));
        if (md) bd = atomicAdd(&st->ndense, __popcll(md));
    }
    bh = __shfl(bh, 0);
    bd = __shfl(bd, 0);
    if (h) hrows[bh + __popcll(mh & below)] = u;
    if (d) drows[bd + __popcll(md & below)] = u;
}

__global__ __launch_bounds__(256) void k_xdelta_hash(const RefineState *__restrict__ st,
                                                     const int32_t *__restrict__ hrows,
                                                     const uint8_t *__restrict__ moved,
                                                     const int4 *__restrict__ ulists, const int32_t *__restrict__ cptr,
                                                     const int2 *__restrict__ crw, const uint32_t *__restrict__ g2y,
                                                     int N1, int32_t *__restrict__ X, int64_t ldx) {
    __shared__ int keys[kHashSlots], vals[kHashSlots];
    __shared__ int2 qual[kHashQual];
    __shared__ int lstart[2 * kHashQual], pref[2 * kHashQual + 1], wsum[4], nq;
    if (!st->inc) return;
    const int nrows = st->nhash, T = (int)(ldx >> 6);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int u = hrows[r];
        for (int i = tid; i < kHashSlots; i += 256) {
            keys[i] = -1;
            vals[i] = 0;
        }
        if (tid == 0) nq = 0;
        __syncthreads();
        for (int e = cptr[u] + tid; e < cptr[u + 1]; e += 256) {  // the entries whose neighbour moved
            const int2 rw = crw[e];
            if (moved[rw.x >= N1 ? rw.x - N1 : rw.x]) {
                const int slot = atomicAdd(&nq, 1);  // fewer than kHashQual + 1 (rowq)
                if (slot < kHashQual) qual[slot] = rw;
            }
        }
        __syncthreads();
        const int m = min(nq, kHashQual);
        int len = 0, lnew = 0;
        if (tid < m) {
            const int2 rw = qual[tid];
            const int x = rw.x >= N1 ? rw.x - N1 : rw.x;
            // an out entry u -> x reads the in-list of x's image, an in entry x -> u its out-list
            const int4 L = ulists[2 * x + (rw.x >= N1 ? 1 : 0)];
            lstart[2 * tid] = L.x;
            lstart[2 * tid + 1] = L.z;
            lnew = L.y;
            len = L.y + L.w;
        }
        int x = len;  // inclusive scan over the 4 waves
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        const int before = (wave > 0 ? wsum[0] : 0) + (wave > 1 ? wsum[1] : 0) + (wave > 2 ? wsum[2] : 0);
        const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (tid < m) {
            pref[2 * tid] = before + x - len;
            pref[2 * tid + 1] = before + x - len + lnew;
        }
        if (tid == 0) pref[2 * m] = total;
        __syncthreads();
        const int span = (total + 3) / 4;
        const int w0 = wave * span, w1 = min(total, w0 + span);
        if (w0 < w1) {
            int lo = 0, hi = 2 * m - 1;  // the list holding flat word w0
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pref[mid] <= w0) lo = mid;
                else hi = mid - 1;
            }
            int l = lo, s0 = pref[l], s1 = pref[l + 1];
            int at = lstart[l] - s0, w = qual[l >> 1].y, sg = (l & 1) ? -1 : 1;
            for (int q = w0 + lane; q < w1; q += 64) {
                while (s1 <= q) {
                    l++;
                    s0 = s1;
                    s1 = pref[l + 1];
                    at = lstart[l] - s0;
                    w = qual[l >> 1].y;
                    sg = (l & 1) ? -1 : 1;
                }
                const uint32_t yv = g2y[at + q];
                const int y = yv >> 16;
                if (y == 0xFFFF) continue;
                const int v = sg * (min(w, (int)(int16_t)(yv & 0xffff)) - min(w, 0));
                unsigned h = ((unsigned)y * 2654435761u) >> 21;  // 11 bits: kHashSlots
                for (int probe = 0; probe < kHashSlots; probe++) {
                    const int old = atomicCAS(&keys[h], -1, y);
                    if (old == -1 || old == y) {
                        atomicAdd(&vals[h], v);
                        break;
                    }
                    h = (h + 1) & (kHashSlots - 1);
                }
            }
        }
        __syncthreads();
        // X(u, y) += the sums: the reads of a thread's eight slots issued together
        int32_t *xr = X + (size_t)(u >> 6) * T * 4096 + ((u & 63) << 6);
        int ky[kHashSlots / 256], dv[kHashSlots / 256], old[kHashSlots / 256];
#pragma unroll
        for (int k = 0; k < kHashSlots / 256; k++) {
            ky[k] = keys[tid + 256 * k];
            dv[k] = ky[k] < 0 ? 0 : vals[tid + 256 * k];
        }
#pragma unroll
        for (int k = 0; k < kHashSlots / 256; k++) old[k] = dv[k] ? xr[((ky[k] >> 6) << 12) + (ky[k] & 63)] : 0;
#pragma unroll
        for (int k = 0; k < kHashSlots / 256; k++)
            if (dv[k]) xr[((ky[k] >> 6) << 12) + (ky[k] & 63)] = old[k] + dv[k];
        __syncthreads();
    }
}

// The same row updates as k_xdelta, gathered per row: the workgroup of a
// row u of the dense kind (k_xrow_class) finds the entries of comb(u) whose neighbour x
// moved (an in entry x -> u reads the G2 out-list of x's image, an out entry
// u -> x its in-list), lays the two lists of each such entry end to end (the
// list at the new image, added, and the one at the old image, subtracted) and
// splits that flat word range evenly over its 16 waves, so that a row whose
// one moved neighbour is a hub keeps every wave busy.  The words accumulate
// f(w, q) in an (n+1)-int LDS row, whose non-zero entries are then added to X
// with plain read-modify-writes: one workgroup owns the row, so the accesses
// are coalesced along it.  Targets inside U are accumulated too: the columns
// of U are rewritten whole afterwards (k_xcol_sum).  Entries are taken in
// passes of at most kXdeltaQual: the whole list when it holds that few moved
// neighbours, else windows of kXdeltaQual consecutive entries.
constexpr int kXdeltaUnroll = 4;  // list words in flight per lane
constexpr int kXdeltaRmw = 8;     // row entries per thread whose X reads are issued together

__global__ __launch_bounds__(1024) __attribute__((amdgpu_num_sgpr(80))) void k_xdelta_rows(
    const RefineState *__restrict__ st, const int32_t *__restrict__ drows, const uint8_t *__restrict__ moved,
    const int4 *__restrict__ ulists, const int32_t *__restrict__ cptr, const int2 *__restrict__ crw,
    const uint32_t *__restrict__ g2y, int N1, int32_t *__restrict__ X, int64_t ldx) {
    extern __shared__ __align__(16) int row[];
    __shared__ int nq;
    __shared__ int2 qual[kXdeltaQual];        // (row, weight) of the entries of this pass
    __shared__ int lstart[2 * kXdeltaQual];   // per list of the pass: its first word in g2y
    __shared__ int pref[2 * kXdeltaQual + 1];  // flat word offset of each list
    __shared__ int wsum[16];
    if (!st->inc) return;
    const int T = (int)(ldx >> 6);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto neighbour = [&](int r) { return r >= N1 ? r - N1 : r; };
    // the entries qual[0, m): their lists into the LDS row
    auto walk = [&](int m) {
        int len = 0, lnew = 0;
        if ((int)threadIdx.x < m) {
            const int r = qual[threadIdx.x].x;
            // an out entry u -> x reads the in-list of x's image, an in entry x -> u its out-list
            const int4 L = ulists[2 * neighbour(r) + (r >= N1 ? 1 : 0)];
            lstart[2 * threadIdx.x] = L.x;
            lstart[2 * threadIdx.x + 1] = L.z;
            lnew = L.y;
            len = L.y + L.w;
        }
        int total;
        const int ex = block_scan_excl(len, wsum, total);
        if ((int)threadIdx.x < m) {
            pref[2 * threadIdx.x] = ex;
            pref[2 * threadIdx.x + 1] = ex + lnew;
        }
        if (threadIdx.x == 0) pref[2 * m] = total;
        __syncthreads();
        const int span = (total + 15) / 16;
        const int w0 = wave * span, w1 = min(total, w0 + span);
        if (w0 < w1) {
            int lo = 0, hi = 2 * m - 1;  // the list holding flat word w0
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pref[mid] <= w0) lo = mid;
                else hi = mid - 1;
            }
            int l = lo, s0 = pref[l], s1 = pref[l + 1];
            int at = lstart[l] - s0, w = qual[l >> 1].y, sg = (l & 1) ? -1 : 1;
            for (int q0 = w0 + lane; q0 < w1; q0 += 64 * kXdeltaUnroll) {
                int ad[kXdeltaUnroll], wu[kXdeltaUnroll], su[kXdeltaUnroll];
#pragma unroll
                for (int k = 0; k < kXdeltaUnroll; k++) {
                    const int q = q0 + 64 * k;
                    ad[k] = -1;
                    wu[k] = 0;
                    su[k] = 1;
                    if (q < w1) {
                        while (s1 <= q) {
                            l++;
                            s0 = s1;
                            s1 = pref[l + 1];
                            at = lstart[l] - s0;
                            w = qual[l >> 1].y;
                            sg = (l & 1) ? -1 : 1;
                        }
                        ad[k] = at + q;
                        wu[k] = w;
                        su[k] = sg;
                    }
                }
                uint32_t yv[kXdeltaUnroll];
#pragma unroll
                for (int k = 0; k < kXdeltaUnroll; k++) yv[k] = ad[k] >= 0 ? g2y[ad[k]] : 0xFFFF0000u;
#pragma unroll
                for (int k = 0; k < kXdeltaUnroll; k++) {
                    if ((yv[k] >> 16) != 0xFFFFu)
                        atomicAdd(&row[yv[k] >> 16], su[k] * (min(wu[k], (int)(int16_t)(yv[k] & 0xffff)) - min(wu[k], 0)));
                }
            }
        }
        __syncthreads();
    };
    const int nd = st->ndense;
    for (int i = blockIdx.x; i < nd; i += gridDim.x) {
        const int u = drows[i];
        if (threadIdx.x == 0) nq = 0;
        for (int y = threadIdx.x; y < N1; y += 1024) row[y] = 0;
        __syncthreads();
        const int e0 = cptr[u], e1 = cptr[u + 1];
        for (int e = e0 + threadIdx.x; e < e1; e += 1024) {
            const int2 rw = crw[e];
            if (moved[neighbour(rw.x)]) {
                const int slot = atomicAdd(&nq, 1);
                if (slot < kXdeltaQual) qual[slot] = rw;
            }
        }
        __syncthreads();
        const int nall = nq;
        __syncthreads();  // every thread has read nq before it is reset
        if (nall == 0) continue;  // the moved neighbours kept their images (uniform)
        if (nall <= kXdeltaQual) {
            walk(nall);
        } else {
            for (int c0 = e0; c0 < e1; c0 += kXdeltaQual) {
                if (threadIdx.x == 0) nq = 0;
                __syncthreads();
                if ((int)threadIdx.x < min(kXdeltaQual, e1 - c0)) {
                    const int2 rw = crw[c0 + threadIdx.x];
                    if (moved[neighbour(rw.x)]) qual[atomicAdd(&nq, 1)] = rw;
                }
                __syncthreads();
                walk(nq);
            }
        }
        // kXdeltaRmw reads of X in flight per thread before their writes; X(u, y) is
        // xr[(y >> 6) * 4096 + (y & 63)] in the tile-major layout
        int32_t *xr = X + (size_t)(u >> 6) * T * 4096 + ((u & 63) << 6);
        for (int y0 = threadIdx.x; y0 < N1; y0 += 1024 * kXdeltaRmw) {
            int d[kXdeltaRmw], v[kXdeltaRmw];
#pragma unroll
            for (int k = 0; k < kXdeltaRmw; k++) d[k] = y0 + 1024 * k < N1 ? row[y0 + 1024 * k] : 0;
#pragma unroll
            for (int k = 0; k < kXdeltaRmw; k++) {
                const int y = y0 + 1024 * k;
                v[k] = d[k] ? xr[((y >> 6) << 12) + (y & 63)] : 0;
            }
#pragma unroll
            for (int k = 0; k < kXdeltaRmw; k++) {
                const int y = y0 + 1024 * k;
                if (d[k]) xr[((y >> 6) << 12) + (y & 63)] = v[k] + d[k];
            }
        }
        __syncthreads();
    }
}

// Column v of X for each v in U: C_u for every row u, plus, for each G2
// line at the new image p'(v) (G2 in-neighbour g -> p'(v): the G1 lines
// s -> u of s = p'^-1(g); out-neighbour p'(v) -> g: the lines u -> s),
// f(w, that line's G2 weight).  The G2 lines of each column are cut into
// chunks (k_xcol_plan, one workgroup, after the move: at least kXcolLines
// lines per chunk, more when the chunks would not fit the cap slots of the
// scratch), so a hub image's column is spread over many workgroups.  k_xcol
// accumulates a chunk in an (n+1)-int LDS column and stores it to its
// scratch row; k_xcol_sum adds C_u and the column's chunk rows and writes
// the column of X once.
__global__ __launch_bounds__(1024) void k_xcol_plan(RefineState *__restrict__ st, const int32_t *__restrict__ ulist,
                                                    const int32_t *__restrict__ map,
                                                    const int32_t *__restrict__ g2optr,
                                                    const int32_t *__restrict__ g2iptr, int32_t *__restrict__ cbase,
                                                    int cap) {
    __shared__ int wsum[16];
    __shared__ long long all;
    __shared__ int carry;
    if (!st->inc) return;
    const int nu = st->nu;
    auto lines_of = [&](int i) {
        const int g = map[ulist[i]];
        return (g2iptr[g + 1] - g2iptr[g]) + (g2optr[g + 1] - g2optr[g]);
    };
    if (threadIdx.x == 0) all = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < nu; i += 1024) atomicAdd((unsigned long long *)&all, (unsigned long long)lines_of(i));
    __syncthreads();
    // chunks <= total / L + nu <= cap
    const long long room = max(1, cap - nu);
    const int L = (int)max((long long)kXcolLines, (all + room - 1) / room);
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nu; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const int chunks = i < nu ? max(1, (lines_of(i) + L - 1) / L) : 0;
        int total;
        const int ex = block_scan_excl(chunks, wsum, total);
        if (i < nu) cbase[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cbase[nu] = carry;
        st->xchunks = carry;
        st->xlines = L;
    }
}

// A work item's G2 lines are taken kXcolHeads at a time; their G1 sub-lists
// are laid end to end and the flat entry range is split evenly over the 16
// waves (a hub preimage's long sub-list no longer falls to one group of lanes).
constexpr int kXcolHeads = 512;

__global__ __launch_bounds__(1024) __attribute__((amdgpu_num_sgpr(80))) void k_xcol(
    const RefineState *__restrict__ st, const int32_t *__restrict__ ulist, const int32_t *__restrict__ cbase,
    const int32_t *__restrict__ map, const int32_t *__restrict__ cptr, const int32_t *__restrict__ cmid,
    const int2 *__restrict__ crw, const int32_t *__restrict__ g2optr, const uint32_t *__restrict__ g2oy,
    const int32_t *__restrict__ g2iptr, const uint32_t *__restrict__ g2iy, int N1, int32_t *__restrict__ scratch) {
    extern __shared__ __align__(16) int col[];
    __shared__ int ls[kXcolHeads], lv[kXcolHeads], pref[kXcolHeads + 1], wsum[16];
    if (!st->inc) return;
    const int nu = st->nu, items = st->xchunks, L = st->xlines;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int it = blockIdx.x; it < items; it += gridDim.x) {
        const int vi = find_base(cbase, nu, it);
        const int g0 = map[ulist[vi]];
        for (int u = threadIdx.x; u < N1; u += 1024) col[u] = 0;
        const int i0 = g2iptr[g0], nin = g2iptr[g0 + 1] - i0;
        const int o0 = g2optr[g0], nout = g2optr[g0 + 1] - o0;
        const int h0 = (it - cbase[vi]) * L, h1 = min(h0 + L, nin + nout);
        for (int c0 = h0; c0 < h1; c0 += kXcolHeads) {
            const int m = min(kXcolHeads, h1 - c0);
            int len = 0;
            if ((int)threadIdx.x < m) {
                // in-line s -> p'(v): the out-lines s -> u of s (in entries of u); out-line p'(v) -> s: the
                // in-lines u -> s (out entries of u)
                const int h = c0 + threadIdx.x;
                const bool in = h < nin;
                const uint32_t yv = in ? g2iy[i0 + h] : g2oy[o0 + h - nin];
                const int sv = yv >> 16;
                if (sv != 0xFFFF) {
                    const int l0 = in ? cptr[sv] : cmid[sv];
                    len = (in ? cmid[sv] : cptr[sv + 1]) - l0;
                    ls[threadIdx.x] = l0;
                    lv[threadIdx.x] = (int)(int16_t)(yv & 0xffff) * 2 + (in ? 1 : 0);  // G2 weight, and the side
                }
            }
            int total;
            const int ex = block_scan_excl(len, wsum, total);  // also orders the zeroing of col
            if ((int)threadIdx.x < m) pref[threadIdx.x] = ex;
            if (threadIdx.x == 0) pref[m] = total;
            __syncthreads();
            const int span = (total + 15) / 16;
            const int w0 = wave * span, w1 = min(total, w0 + span);
            if (w0 < w1) {
                int lo = 0, hi = m - 1;  // the line holding flat entry w0
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (pref[mid] <= w0) lo = mid;
                    else hi = mid - 1;
                }
                int l = lo, s0 = pref[l], s1 = pref[l + 1], at = ls[l] - s0, hv = lv[l];
                for (int q0 = w0 + lane; q0 < w1; q0 += 256) {  // four entries in flight per lane
                    int ad[4], vu[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int q = q0 + 64 * k;
                        ad[k] = -1;
                        vu[k] = 0;
                        if (q < w1) {
                            while (s1 <= q) {
                                l++;
                                s0 = s1;
                                s1 = pref[l + 1];
                                at = ls[l] - s0;
                                hv = lv[l];
                            }
                            ad[k] = at + q;
                            vu[k] = hv;
                        }
                    }
                    int2 rw[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) rw[k] = ad[k] >= 0 ? crw[ad[k]] : make_int2(-1, 0);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        if (rw[k].x < 0) continue;
                        const int val = vu[k] >> 1;
                        atomicAdd(&col[(vu[k] & 1) ? rw[k].x - N1 : rw[k].x], min(rw[k].y, val) - min(rw[k].y, 0));
                    }
                }
            }
            __syncthreads();
        }
        __syncthreads();  // an item without lines: the zeroing before the store
        int32_t *row = scratch + (size_t)it * N1;
        for (int u = threadIdx.x; u < N1; u += 1024) row[u] = col[u];
        __syncthreads();
    }
}

// Column v of X for the moved set, summed and written a 64 x 64 block at a
// time: a block takes a tile row I (64 rows u) and 64 consecutive entries of
// U (kept in increasing vertex order), sums C_u and each column's chunk rows
// into an LDS tile (lanes along u: coalesced chunk-row reads), then writes it
// with lanes along the columns, so the columns of U that share a 64-column
// tile land in the same 256-byte piece of row u together.
__global__ __launch_bounds__(256) void k_xcol_sum(const RefineState *__restrict__ st,
                                                  const int32_t *__restrict__ ulist, const int32_t *__restrict__ cbase,
                                                  const int32_t *__restrict__ cneg,
                                                  const int32_t *__restrict__ scratch, int N1,
                                                  int32_t *__restrict__ X, int64_t ldx) {
    __shared__ int tile[64][65];
    __shared__ int vcol[64];
    if (!st->inc) return;
    const int nu = st->nu, T = (int)(ldx >> 6);
    const int rows = (N1 + 63) / 64, groups = (nu + 63) / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int w = blockIdx.x; w < rows * groups; w += gridDim.x) {
        const int I = w % rows, c0 = (w / rows) * 64, nc = min(64, nu - c0);
        const int u = I * 64 + lane;
        for (int c = wave; c < nc; c += 4) {
            const int vi = c0 + c;
            int x = 0;
            if (u < N1) {
                x = cneg[u];
                for (int it = cbase[vi]; it < cbase[vi + 1]; it++) x += scratch[(size_t)it * N1 + u];
            }
            tile[c][lane] = x;
        }
        if ((int)threadIdx.x < nc) vcol[threadIdx.x] = ulist[c0 + threadIdx.x];
        __syncthreads();
        if (lane < nc) {
            const int v = vcol[lane];
            for (int r = wave; r < 64; r += 4) {
                const int ur = I * 64 + r;
                if (ur < N1) X[xidx(ur, v, T)] = tile[lane][r];
            }
        }
        __syncthreads();
    }
}

}  // namespace

struct bm_batch_state {
    bm_refine_params prm{};
    bool configured = false;
    int captured_kind = -1;  // sparse_ok | all_mapped << 1 at the capture (they select the sweep kernels)
    int32_t *tri = nullptr, *tri_d = nullptr, *tri_p = nullptr;
    int32_t *sets = nullptr, *set_d = nullptr, *set_r = nullptr;
    int *ovf = nullptr;
    unsigned long long *ckey = nullptr, *vmax = nullptr;
    int2 *taken = nullptr;  // per vertex: 0, or (new image + 1, old image | round << 20) once a winner of this step moves it
    int32_t *win = nullptr;
    int32_t *clist = nullptr;  // the live candidates of a selection round, two lists of M (rounds alternate)
    int32_t *ulist = nullptr, *uold = nullptr, *ubase = nullptr, *cbase = nullptr;
    int32_t *xscratch = nullptr;  // chunk rows of the column recompute, xcap x (n+1)
    int xcap = 0;
    uint8_t *inU = nullptr;
    int4 *ulists = nullptr;    // per moved vertex: its G2 out- and in-lists at the new and the old image
    int32_t *hrows = nullptr;  // rows updated through the LDS hash table
    int32_t *drows = nullptr;  // rows updated through the (n+1)-int LDS row
    bool incremental = false;  // the captured step updates X incrementally
    bool rot3_x = false;       // triples and sets scored from the step's X (k_rot3_x, k_permset) instead of their lists
    bool pfree = false;        // the captured step reads W2 through the mapping and leaves P / P^T behind it
    bool xdelta_rows = true;   // row updates gathered per row (k_xdelta_rows) instead of per entry
    int32_t *uoldv = nullptr;  // image before the step, indexed by moved vertex
    uint64_t x_epoch = ~0ull;  // map_epoch at the end of the last run (X then matched the mapping)
    int inc_max = 0;
    int pinc_max = 0;
    long long inc_budget = 0;  // list work an incremental update may take (k_collect)
    RefineState *st = nullptr;
    RefineState *hst = nullptr;  // pinned copy
    bm_refine_step *hist = nullptr;
    MoveRec *log = nullptr;
    int log_cap = 0;
    long long M = 0;
    size_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool in_flight = false;
    int64_t run_steps = 0;
};

static CandView cand_view(bm_ctx *c) {
    bm_batch_state *b = c->batch;
    CandView cv;
    cv.rowkeys = c->keys;
    cv.tri = b->tri;
    cv.sets = b->sets;
    cv.tri_d = b->tri_d;
    cv.tri_p = b->tri_p;
    cv.set_d = b->set_d;
    cv.set_r = b->set_r;
    cv.N1 = c->N1;
    cv.n = c->n;
    cv.T = b->prm.triples;
    cv.S = b->prm.sets;
    cv.k = b->prm.set_k;
    cv.pairs = b->prm.pairs;
    return cv;
}

// the incremental update's byte flags share one allocation: inU, the row
// marks, the moved flags, (N1 + 3) / 4 words each
static uint8_t *rowmark(bm_batch_state *b, int N1) { return b->inU + (size_t)(N1 + 3) / 4 * 4; }
static uint8_t *moved_flags(bm_batch_state *b, int N1) { return b->inU + (size_t)(N1 + 3) / 4 * 8; }
// then the per-row counts of the row updates: rowq (entries), roww (words), N1 ints each
static int32_t *row_q(bm_batch_state *b, int N1) { return (int32_t *)(b->inU + (size_t)(N1 + 3) / 4 * 12); }

static void drop_graph(bm_batch_state *b) {
    std::lock_guard<std::mutex> g(sync_mutex());
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int bm::free_batch(bm_ctx *c) {
    bm_batch_state *b = c->batch;
    if (!b) return BM_OK;
    drop_graph(b);
    (void)hipStreamSynchronize(c->stream);  // no step of this state is still running
    auto fr = [&](void *p) { block_free(p); };
    fr(b->tri), fr(b->tri_d), fr(b->tri_p), fr(b->sets), fr(b->set_d), fr(b->set_r), fr(b->ovf);
    fr(b->ckey), fr(b->vmax), fr(b->taken), fr(b->win), fr(b->clist), fr(b->st), fr(b->hist), fr(b->log);
    fr(b->ulist), fr(b->uold), fr(b->ubase), fr(b->cbase), fr(b->xscratch), fr(b->inU), fr(b->uoldv);
    fr(b->ulists), fr(b->hrows), fr(b->drows);
    {
        std::lock_guard<std::mutex> g(sync_mutex());
        if (b->hst) (void)hipHostFree(b->hst);
        if (b->done) (void)hipEventDestroy(b->done);
    }
    c->dev_bytes -= (int64_t)b->bytes;
    delete b;
    c->batch = nullptr;
    return BM_OK;
}

static int balloc(bm_ctx *c, void **p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    const hipError_t e = block_alloc(c->device, p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(BM_ENOMEM, "device allocation of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
    }
    c->batch->bytes += bytes;
    c->dev_bytes += (int64_t)bytes;
    return BM_OK;
}

// Everything one step enqueues; capturable (kernels only).
static int enqueue_step(bm_ctx *c) {
    bm_batch_state *b = c->batch;
    const bm_refine_params &p = b->prm;
    const int N1 = c->N1;
    const CandView cv = cand_view(c);
    int rc;
    // X is updated in place between steps only with the sparse sweep and swap candidates
    const bool inc = p.pairs && sweep_is_sparse(c);
    b->incremental = inc;
    // every kernel of the step can read W2 through the mapping instead of P / P^T when X is kept
    // current and the triples and sets are scored from it; P / P^T are then rebuilt only for a
    // whole sweep, and otherwise on the next use outside the run (ensure_p)
    const bool pfree = inc && (p.triples == 0 || b->rot3_x) && (p.sets == 0 || b->rot3_x);
    b->pfree = pfree;
    const int16_t *Pq = pfree ? c->W2 : c->Q;
    const int32_t *pmap = pfree ? c->map : nullptr;
    {
        const long long vm_words = (long long)p.rounds * N1 * 2;
        // inU, the row marks, the moved flags, the row counts
        const long long mark_words = inc ? (long long)(N1 + 3) / 4 * 3 + 2LL * N1 : 0;
        const long long most = std::max(std::max(vm_words, 2LL * N1), mark_words);
        BM_LAUNCH(c, "refine_begin",
                  hipLaunchKernelGGL(k_step_prologue, dim3((unsigned)((most + 255) / 256)), dim3(256), 0, c->stream,
                                     b->st, pfree ? 1 : 0, (uint32_t *)b->vmax, vm_words, (uint32_t *)b->taken, 2LL * N1,
                                     (uint32_t *)b->inU, mark_words));
    }
    if (pfree && (rc = launch_rebuild_p_if(c, &b->st->prebuild))) return rc;
    if (p.pairs) {
        // the sweep's fill and combine skipped after a step that moved nothing (same mapping, same keys)
        if ((rc = sweep_enqueue(c, nullptr, inc ? &b->st->xstale : nullptr, &b->st->keep_keys))) return rc;
    }
    if (p.triples > 0) {
        BM_LAUNCH(c, "gen_triples",
                  hipLaunchKernelGGL(k_gen_triples, dim3((p.triples + 255) / 256), dim3(256), 0, c->stream, b->tri,
                                     p.triples, b->st, (unsigned long long)p.seed, c->cptr, c->cr, N1, c->n));
        rc = b->rot3_x ? launch_rot3_x(c, b->tri, p.triples, BM_ROT3_ALL, b->tri_d, b->tri_p, pfree)
                       : launch_rot3(c, b->tri, p.triples, BM_ROT3_ALL, b->tri_d, b->tri_p);
        if (rc) return rc;
    }
    if (p.sets > 0) {
        BM_LAUNCH(c, "gen_sets",
                  hipLaunchKernelGGL(k_gen_sets, dim3((p.sets + 255) / 256), dim3(256), 0, c->stream, b->sets, p.sets,
                                     p.set_k, p.set_gen, b->st, (unsigned long long)p.seed, c->cptr, c->cr, N1, c->n));
        if ((rc = launch_permset(c, p.set_k, b->sets, p.sets, b->set_d, b->set_r, b->ovf, 1, b->rot3_x, pfree)))
            return rc;
    }
    const unsigned mb = (unsigned)((b->M + 255) / 256);
    BM_LAUNCH(c, "cand_keys",
              hipLaunchKernelGGL(k_cand_keys, dim3(mb), dim3(256), 0, c->stream, b->ckey, b->M, cv, c->vsimple,
                                 b->st, inc ? 1 : 0, b->vmax, b->clist));
    for (int r = 0; r < p.rounds; r++) {
        unsigned long long *vm = b->vmax + (size_t)r * N1;
        unsigned long long *vmn = r + 1 < p.rounds ? vm + N1 : nullptr;
        DecideArgs da{b->ckey, cv, vm, vmn, b->taken, b->win, N1 / 2 + 1, r, b->st, c->cptr,
                      (const int2 *)c->crw, c->map, c->W2, c->ld, N1};
        BM_LAUNCH(c, "decide",
                  hipLaunchKernelGGL(k_decide, dim3((unsigned)std::min<long long>((b->M + 3) / 4, 2048)), dim3(256),
                                     0, c->stream, da, b->clist + (size_t)(r & 1) * b->M,
                                     b->clist + (size_t)((r + 1) & 1) * b->M));
    }
    const unsigned ublocks = 1024;  // grid-stride kernels of the incremental update
    if (inc) {
        BM_LAUNCH(c, "collect", hipLaunchKernelGGL(k_collect, dim3(1), dim3(1024), 0, c->stream, b->win, cv, c->map,
                                                   c->cptr, c->g2optr, c->g2iptr, b->ulist, b->uold, b->ubase, b->inU,
                                                   b->uoldv, b->st, b->inc_max, b->pinc_max, b->inc_budget,
                                                   (int)(c->ncomb / std::max(1, c->N1)), pfree ? 1 : 0, N1));
        BM_LAUNCH(c, "corr_inc", hipLaunchKernelGGL(k_corr_inc, dim3(ublocks), dim3(256), 0, c->stream, -1, b->st,
                                                    b->ulist, b->ubase, b->inU, c->cptr, (const int2 *)c->crw, c->cdw,
                                                    Pq, c->ld, N1, c->X, c->ldx, rowmark(b, N1), pmap));
    }
    const int max_win = N1 / 2 + 1;
    BM_LAUNCH(c, "apply", hipLaunchKernelGGL(k_apply, dim3((max_win + 255) / 256), dim3(256), 0, c->stream, b->win,
                                             max_win, cv, c->map, c->pinv, c->sparse_ok ? 1 : 0, b->st, b->log,
                                             b->log_cap, inc ? moved_flags(b, N1) : nullptr,
                                             inc ? b->ulists : nullptr, c->g2optr, c->g2iptr, (int)c->ng2));
    if (!pfree && (rc = inc ? launch_rebuild_p(c, &b->st->pinc, b->ulist, &b->st->nu) : launch_rebuild_p(c)))
        return rc;
    if (c->sparse_ok) {
        if (inc)  // only the words at the moved images change
            BM_LAUNCH(c, "g2y_patch", hipLaunchKernelGGL(k_g2y_patch, dim3(1024), dim3(256), 0, c->stream, b->st,
                                                         b->ulist, c->map, c->g2oy, c->g2iy, c->g2o, c->g2i,
                                                         c->g2optr, c->g2iptr, c->g2o_at_in, c->g2i_at_out,
                                                         moved_flags(b, N1), b->ulists, c->cptr,
                                                         (const int2 *)c->crw, N1,
                                                         b->xdelta_rows ? row_q(b, N1) : nullptr));
        else if ((rc = launch_refresh_g2y(c)))
            return rc;
    }
    if (inc) {
        if (b->xdelta_rows) {
            BM_LAUNCH(c, "xrow_class",
                      hipLaunchKernelGGL(k_xrow_class, dim3((N1 + 255) / 256), dim3(256), 0, c->stream, b->st,
                                         row_q(b, N1), row_q(b, N1) + N1, N1, b->hrows, b->drows));
            BM_LAUNCH(c, "xdelta_hash",
                      hipLaunchKernelGGL(k_xdelta_hash, dim3((unsigned)std::min(N1, 4096)), dim3(256), 0, c->stream,
                                         b->st, b->hrows, moved_flags(b, N1), b->ulists, c->cptr,
                                         (const int2 *)c->crw, c->g2oy, N1, c->X, c->ldx));
            BM_LAUNCH(c, "xdelta", hipLaunchKernelGGL(k_xdelta_rows, dim3((unsigned)std::min(N1, 2048)), dim3(1024),
                                                      (size_t)N1 * 4, c->stream, b->st, b->drows,
                                                      moved_flags(b, N1), b->ulists, c->cptr, (const int2 *)c->crw,
                                                      c->g2oy, N1, c->X, c->ldx));
        }
        else
            BM_LAUNCH(c, "xdelta", hipLaunchKernelGGL(k_xdelta, dim3(ublocks), dim3(256), 0, c->stream, b->st,
                                                      b->ulist, b->uold, b->ubase, b->inU, c->map, c->cptr,
                                                      (const int2 *)c->crw, c->g2optr, c->g2oy, c->g2iptr, c->g2iy,
                                                      N1, c->X, c->ldx));
        BM_LAUNCH(c, "xcol_plan", hipLaunchKernelGGL(k_xcol_plan, dim3(1), dim3(1024), 0, c->stream, b->st, b->ulist,
                                                     c->map, c->g2optr, c->g2iptr, b->cbase, b->xcap));
        BM_LAUNCH(c, "xcol", hipLaunchKernelGGL(k_xcol, dim3(kXcolBlocks), dim3(1024), (size_t)N1 * 4, c->stream,
                                                b->st, b->ulist, b->cbase, c->map, c->cptr, c->cmid,
                                                (const int2 *)c->crw, c->g2optr, c->g2oy, c->g2iptr, c->g2iy, N1,
                                                b->xscratch));
        BM_LAUNCH(c, "xcol_sum", hipLaunchKernelGGL(k_xcol_sum, dim3(2048), dim3(256), 0, c->stream, b->st, b->ulist,
                                                    b->cbase, c->cneg, b->xscratch, N1, c->X, c->ldx));
        BM_LAUNCH(c, "corr_inc", hipLaunchKernelGGL(k_corr_inc, dim3(ublocks), dim3(256), 0, c->stream, 1, b->st,
                                                    b->ulist, b->ubase, b->inU, c->cptr, (const int2 *)c->crw, c->cdw,
                                                    Pq, c->ld, N1, c->X, c->ldx, nullptr, pmap));
    }
    // the exact check: with U collected, the exact change over the lines at U (the whole score
    // is recomputed at the end of the run); else the whole score
    if (inc)
        BM_LAUNCH(c, "score_change", hipLaunchKernelGGL(k_score_change, dim3(ublocks), dim3(256), 0, c->stream, b->st,
                                                        b->ulist, b->ubase, b->inU, b->uoldv, c->map, c->cptr,
                                                        (const int2 *)c->crw, c->W2, c->ld, N1, c->X, c->ldx,
                                                        c->diag));
    else if ((rc = launch_score(c, &b->st->score_slot, pfree)))
        return rc;
    BM_LAUNCH(c, "refine_end", hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, c->stream, b->st, b->hist,
                                                  inc ? 1 : 0));
    return BM_OK;
}

static int capture(bm_ctx *c) {
    bm_batch_state *b = c->batch;
    drop_graph(b);
    const bool timing = c->timing, dirty = c->g2y_dirty;
    c->timing = false;     // no event records inside the graph
    c->g2y_dirty = false;  // each step ends with the refresh; refine_start covers the first
    hipGraph_t g = nullptr;
    hipError_t e;
    int rc;
    {
        std::lock_guard<std::mutex> lk(sync_mutex());
        e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            c->timing = timing;
            c->g2y_dirty = dirty;
            return hip_fail(e, "hipStreamBeginCapture");
        }
        rc = enqueue_step(c);
        e = hipStreamEndCapture(c->stream, &g);
        if (rc && g) (void)hipGraphDestroy(g);
        if (!rc && e == hipSuccess) {
            b->graph = g;
            e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                b->graph = nullptr;
                b->exec = nullptr;
                c->timing = timing;
                c->g2y_dirty = dirty;
                return hip_fail(e, "hipGraphInstantiate");
            }
        }
    }
    c->timing = timing;
    c->g2y_dirty = dirty;
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->captured_kind = (c->sparse_ok ? 1 : 0) | (c->all_mapped ? 2 : 0);
    return BM_OK;
}



// Lanes per triple and threads per set, timed on this instance's own
// generated candidates (the move-kernel autotuner; the choice never changes results).
static int autotune(bm_ctx *c) {
    bm_batch_state *b = c->batch;
    const bm_refine_params &p = b->prm;
    int rc;
    if (p.triples > 0 && !b->rot3_x) {
        BM_LAUNCH(c, "gen_triples",
                  hipLaunchKernelGGL(k_gen_triples, dim3((p.triples + 255) / 256), dim3(256), 0, c->stream, b->tri,
                                     p.triples, b->st, (unsigned long long)p.seed, c->cptr, c->cr, c->N1, c->n));
        if ((rc = tune_rot3(c, b->tri, p.triples, b->tri_d, b->tri_p))) return rc;
    }
    if (p.sets > 0) {
        BM_LAUNCH(c, "gen_sets",
                  hipLaunchKernelGGL(k_gen_sets, dim3((p.sets + 255) / 256), dim3(256), 0, c->stream, b->sets, p.sets,
                                     p.set_k, p.set_gen, b->st, (unsigned long long)p.seed, c->cptr, c->cr, c->N1,
                                     c->n));
        if ((rc = tune_permset(c, p.set_k, b->sets, p.sets, b->set_d, b->set_r, b->ovf, 1, b->rot3_x))) return rc;
    }
    BM_HIP(hipStreamSynchronize(c->stream));
    return BM_OK;
}

namespace bm {

int refine_configure(bm_ctx *c, const bm_refine_params *prm) {
    if (!c || !prm) return set_error(BM_EINVAL, "NULL argument");
    if (!c->mapped) return set_error(BM_ESTATE, "mapping not set");
    if (int prc = ensure_p(c)) return prc;
    bm_refine_params p = *prm;
    if (p.triples < 0 || p.sets < 0 || p.triples > (1 << 24) || p.sets > (1 << 20))
        return set_error(BM_EINVAL, "triples %d / sets %d outside [0, 2^24] / [0, 2^20]", p.triples, p.sets);
    if (p.sets > 0 && (p.set_k < 2 || p.set_k > 8)) return set_error(BM_EINVAL, "set_k=%d outside [2,8]", p.set_k);
    if (p.set_gen != 0 && p.set_gen != 1) return set_error(BM_EINVAL, "set_gen=%d not 0 or 1", p.set_gen);
    // rounds 0: by size (a round's launch outweighs its work on small pairs; on large ones the hub
    // moves settle one per round, and more rounds per step mean fewer steps)
    if (p.rounds == 0) p.rounds = c->n >= 8000 ? kMaxRounds : 32;
    if (p.rounds < 1 || p.rounds > kMaxRounds)
        return set_error(BM_EINVAL, "rounds=%d outside [0,%d]", p.rounds, kMaxRounds);
    if (p.steps_per_check < 1) return set_error(BM_EINVAL, "steps_per_check=%d < 1", p.steps_per_check);
    if (!p.pairs && !p.triples && !p.sets) return set_error(BM_EINVAL, "no move kind enabled");
    if (c->n < 3) return set_error(BM_ERANGE, "n=%d too small for batched refinement", c->n);
    if (p.sets == 0) p.set_k = 0;
    BM_HIP(hipSetDevice(c->device));
    BM_HIP(hipStreamSynchronize(c->stream));
    free_batch(c);
    c->batch = new bm_batch_state();
    bm_batch_state *b = c->batch;
    b->prm = p;
    const int N1 = c->N1;
    b->M = (long long)N1 + p.triples + p.sets;
    int rc;
    auto fail = [&](int r) {
        free_batch(c);
        return r;
    };
    if ((rc = balloc(c, (void **)&b->tri, (size_t)p.triples * 12))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->tri_d, (size_t)p.triples * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->tri_p, (size_t)p.triples * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->sets, (size_t)p.sets * 4 * std::max(p.set_k, 1)))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->set_d, (size_t)p.sets * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->set_r, (size_t)p.sets * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->ovf, 64))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->ckey, (size_t)b->M * 8))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->vmax, (size_t)p.rounds * N1 * 8))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->taken, (size_t)N1 * 8))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->win, (size_t)(N1 / 2 + 1) * 12))) return fail(rc);  // index, change, round
    if ((rc = balloc(c, (void **)&b->clist, (size_t)b->M * 8))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->ulist, (size_t)N1 * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->uold, (size_t)N1 * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->ubase, (size_t)(N1 + 1) * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->cbase, (size_t)(N1 + 1) * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->uoldv, (size_t)N1 * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->inU, (size_t)(N1 + 3) / 4 * 12 + (size_t)N1 * 8))) return fail(rc);  // + rowq, roww
    if ((rc = balloc(c, (void **)&b->ulists, (size_t)N1 * 2 * sizeof(int4)))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->hrows, (size_t)N1 * 4))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->drows, (size_t)N1 * 4))) return fail(rc);
    // steps moving at most inc_max vertices update X in place (about 3 |U| / n of a sweep)
    b->inc_max = std::max(64, c->N1 / 8);
    if (const char *e = getenv("BM_REFINE_INC_MAX")) b->inc_max = atoi(e);
    // chunk rows of the column recompute: one per moved vertex plus kXcolSpare
    b->xcap = std::min(std::max(b->inc_max, 0), N1) + kXcolSpare;
    if ((rc = balloc(c, (void **)&b->xscratch, (size_t)b->xcap * N1 * 4))) return fail(rc);
    // The sweep walks about sum_v deg1(v) deg2(p(v)) list words with LDS atomics; the
    // update, with global atomics, may take a quarter of that (measured on the served
    // workload: 1/2 to 1/4 are equally fast, 1/32 falls back to whole sweeps too often;
    // BM_REFINE_INC_RATIO overrides the 4).
    {
        long long wfull = 0;
        for (int v = 0; v < N1; v++) wfull += (long long)c->hcdeg[v] * c->hg2deg[c->hmap[v]];
        long long ratio = 4;
        if (const char *e = getenv("BM_REFINE_INC_RATIO")) ratio = std::max(1, atoi(e));
        b->inc_budget = wfull / ratio;
        if (getenv("BM_REFINE_INC_MAX")) b->inc_budget = (long long)1 << 62;  // count rule only
    }
    // and patch P / P^T at the moved rows and columns up to pinc_max vertices
    b->pinc_max = std::max(32, c->N1 / 64);
    if (const char *e = getenv("BM_REFINE_PATCH_MAX")) b->pinc_max = atoi(e);
    if ((rc = balloc(c, (void **)&b->st, sizeof(RefineState)))) return fail(rc);
    if ((rc = balloc(c, (void **)&b->hist, sizeof(bm_refine_step) * kHist))) return fail(rc);
    hipError_t he;
    {
        std::lock_guard<std::mutex> g(sync_mutex());
        he = hipHostMalloc((void **)&b->hst, sizeof(RefineState), hipHostMallocDefault);
    }
    if (he != hipSuccess) return fail(set_error(BM_ENOMEM, "pinned host allocation failed"));
    if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess)
        return fail(set_error(BM_EHIP, "hipEventCreate failed"));
    if (hipMemsetAsync(b->st, 0, sizeof(RefineState), c->stream) != hipSuccess ||
        hipMemsetAsync(b->ovf, 0, 64, c->stream) != hipSuccess)
        return fail(set_error(BM_EHIP, "hipMemsetAsync failed"));
    if (p.pairs && (rc = sweep_prepare(c))) return fail(rc);
    if (c->sparse_ok && (rc = launch_refresh_g2y(c))) return fail(rc);
    // With pair moves the step keeps X current before the triples are scored; when every G1 line
    // is its own dense cell, the triples read their deltas from it (BM_REFINE_ROT3=direct: lists)
    b->rot3_x = false;
    if ((p.triples > 0 || p.sets > 0) && p.pairs && c->dense_equal) {
        const char *e = getenv("BM_REFINE_ROT3");
        if (!e || strcmp(e, "direct") != 0) {
            if ((rc = ensure_df_cells(c))) return fail(rc);
            b->rot3_x = true;
        }
    }
    // the direct triple kernel reads the de-duplicated lists: on the device before the capture
    if (p.triples > 0 && !b->rot3_x && (rc = ensure_dd(c))) return fail(rc);
    if ((rc = autotune(c))) return fail(rc);
    if (p.pairs && c->sweep_kind == 2 && c->sparse_ok && !c->sparse_attr) {
        // first sparse launch outside the capture sets the kernel's LDS attribute
        if ((rc = sweep_enqueue(c, nullptr))) return fail(rc);
    }
    if (hipFuncSetAttribute((const void *)k_xcol, hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024) !=
            hipSuccess ||
        hipFuncSetAttribute((const void *)k_xdelta_rows, hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024) !=
            hipSuccess)
        return fail(set_error(BM_EHIP, "hipFuncSetAttribute failed"));
    // BM_REFINE_XDELTA=entries: the per-entry form with global atomics
    b->xdelta_rows = !(getenv("BM_REFINE_XDELTA") && strcmp(getenv("BM_REFINE_XDELTA"), "entries") == 0);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_error(BM_EHIP, "configure sync failed"));
    if ((rc = capture(c))) return fail(rc);
    b->configured = true;
    return BM_OK;
}

int refine_set_log(bm_ctx *c, int32_t cap) {
    bm_batch_state *b = c->batch;
    if (b->log) {
        BM_HIP(hipStreamSynchronize(c->stream));
        block_free(b->log);
        b->bytes -= (size_t)b->log_cap * sizeof(MoveRec);
        c->dev_bytes -= (int64_t)b->log_cap * (int64_t)sizeof(MoveRec);
        b->log = nullptr;
    }
    b->log_cap = 0;
    int rc;
    if (cap > 0) {
        if ((rc = balloc(c, (void **)&b->log, (size_t)cap * sizeof(MoveRec)))) return rc;
        b->log_cap = cap;
    }
    return capture(c);  // the log pointer is a kernel argument
}

// New run: exact score of the current mapping, counters zeroed.
int refine_start(bm_ctx *c) {
    bm_batch_state *b = c->batch;
    if (b->captured_kind != ((c->sparse_ok ? 1 : 0) | (c->all_mapped ? 2 : 0))) {  // the mapping changed kind
        int rc = capture(c);
        if (rc) return rc;
    }
    if (c->sparse_ok && c->g2y_dirty) {
        int rc = launch_refresh_g2y(c);
        if (rc) return rc;
    }
    if (!b->pfree) {  // this step keeps P / P^T current from where they are
        int rc = ensure_p(c);
        if (rc) return rc;
    }
    BM_HIP(hipMemsetAsync(&b->st->score_slot, 0, 8, c->stream));
    int rc = launch_score(c, &b->st->score_slot, b->pfree);
    if (rc) return rc;
    const int keep_x = b->x_epoch == c->map_epoch && b->incremental ? 1 : 0;
    BM_LAUNCH(c, "refine_init", hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, c->stream, b->st, keep_x,
                                                   c->p_stale ? 1 : 0));
    b->run_steps = 0;
    b->in_flight = false;
    return BM_OK;
}

// Enqueue `steps` replays and an async copy of the state; returns at once.
int refine_enqueue(bm_ctx *c, int steps) {
    bm_batch_state *b = c->batch;
    {
        // graph launches, like every other graph call (capture, instantiate,
        // destroy), are serialised across the threads of a process: the
        // serving loop launches from its scheduler thread while preparation
        // threads capture and instantiate other jobs' graphs and destroy
        // retired ones
        std::lock_guard<std::mutex> lk(sync_mutex());
        for (int s = 0; s < steps; s++) BM_HIP(hipGraphLaunch(b->exec, c->stream));
    }
    BM_HIP(hipMemcpyAsync(b->hst, b->st, sizeof(RefineState), hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipEventRecord(b->done, c->stream));
    b->in_flight = true;
    b->run_steps += steps;
    return BM_OK;
}

// 1 when the last enqueued chunk finished, 0 while it runs, < 0 on error
int refine_poll(bm_ctx *c, bool wait) {
    bm_batch_state *b = c->batch;
    if (!b->in_flight) return 1;
    hipError_t e = wait ? hipEventSynchronize(b->done) : hipEventQuery(b->done);
    if (e == hipErrorNotReady) return 0;
    if (e != hipSuccess) return hip_fail(e, "refine chunk");
    b->in_flight = false;
    return 1;
}

void refine_stats(bm_ctx *c, bm_refine_stats *out) {
    const RefineState &s = *c->batch->hst;
    memset(out, 0, sizeof *out);
    out->steps = (int64_t)s.step;
    out->moves = s.moves;
    out->pair_moves = s.kind_moves[0];
    out->triple_moves = s.kind_moves[1];
    out->set_moves = s.kind_moves[2];
    out->initial_score = s.initial_score;
    out->final_score = s.score;
    out->idle_steps = s.idle;
    out->verified = s.bad_step == 0 ? 1 : 0;
    out->bad_step = s.bad_step ? (int64_t)s.bad_step - 1 : -1;
}

// Copies the mapping back after a run (the device holds the truth).  When the
// steps checked their scores by the exact change over the lines at the moved
// vertices, the whole score of the final mapping is recomputed and must equal
// the run's running score.
int refine_finish(bm_ctx *c) {
    bm_batch_state *b = c->batch;
    const bool rescore = b->incremental && b->hst->moves > 0 && b->hst->bad_step == 0;
    if (rescore) {
        BM_HIP(hipMemsetAsync(&b->st->score_slot, 0, 8, c->stream));
        int rc = launch_score(c, &b->st->score_slot, b->pfree);
        if (rc) return rc;
        BM_HIP(hipMemcpyAsync(&b->hst->score_slot, &b->st->score_slot, 8, hipMemcpyDeviceToHost, c->stream));
    }
    BM_HIP(hipMemcpyAsync(c->hmap.data(), c->map, (size_t)c->N1 * 4, hipMemcpyDeviceToHost, c->stream));
    BM_HIP(hipStreamSynchronize(c->stream));
    if (c->sparse_ok) {
        std::fill(c->hpinv.begin(), c->hpinv.end(), -1);
        for (int v = 0; v < c->N1; v++)
            if (c->hmap[v]) c->hpinv[c->hmap[v]] = v;
    }
    c->diag_valid = c->dfdiag_valid = false;
    c->g2y_dirty = !c->sparse_ok;
    if (c->batch->pfree && c->batch->hst->moves > 0) c->p_stale = true;  // rebuilt on next use
    c->map_epoch++;  // the device moved the mapping; X matches it unless the last step left it stale
    c->batch->x_epoch = c->map_epoch;
    if (rescore && (long long)b->hst->score_slot != b->hst->score)
        return set_error(BM_ERANGE, "whole rescore %lld after the run differs from its running score %lld",
                         (long long)b->hst->score_slot, b->hst->score);
    return BM_OK;
}

}  // namespace bm

extern "C" int bm_refine_configure(bm_ctx *c, const bm_refine_params *p) { return refine_configure(c, p); }

extern "C" int bm_refine_run(bm_ctx *c, int64_t max_steps, int64_t move_budget, int32_t patience,
                             bm_refine_stats *stats) {
    if (!c || !stats) return set_error(BM_EINVAL, "NULL argument");
    if (!c->batch || !c->batch->configured) return set_error(BM_ESTATE, "bm_refine_configure not called");
    if (max_steps < 0 || patience < 1) return set_error(BM_EINVAL, "max_steps %lld / patience %d", (long long)max_steps,
                                                        patience);
    BM_HIP(hipSetDevice(c->device));
    bm_batch_state *b = c->batch;
    int rc = refine_start(c);
    if (rc) return rc;
    bm_refine_stats s;
    memset(&s, 0, sizeof s);
    int64_t done = 0;
    for (;;) {
        const int chunk = (int)std::min<int64_t>(b->prm.steps_per_check, max_steps - done);
        if ((rc = refine_enqueue(c, chunk))) return rc;
        if ((rc = refine_poll(c, true)) < 0) return rc;
        done += chunk;
        refine_stats(c, &s);
        if (!s.verified) break;
        if (done >= max_steps) break;
        if (move_budget > 0 && s.moves >= move_budget) break;
        if (s.idle_steps >= patience) {
            s.converged = 1;
            break;
        }
    }
    if ((rc = refine_finish(c))) return rc;
    *stats = s;
    if (!s.verified)
        return set_error(BM_ERANGE, "exact score check failed at step %lld", (long long)s.bad_step);
    return BM_OK;
}

// Several graph pairs at once from one host thread: every context's step
// graph is replayed on the context's own stream, steps_per_check steps per
// chunk followed by the copy of its counters, so the pairs' kernels overlap on
// the device (one parent graph holding every pair's steps as child nodes was
// slower than one pair after another: ROCm runs the branches of a graph in
// order).  When a pair's chunk has finished, its stop rules are applied
// exactly as bm_refine_run applies them and its next chunk is enqueued at
// once; the others keep running.
extern "C" int bm_refine_run_many(bm_ctx *const *ctxs, int32_t count, int64_t max_steps, int64_t move_budget,
                                  int32_t patience, bm_refine_stats *stats) {
    if (!ctxs || !stats || count < 1) return set_error(BM_EINVAL, "NULL argument or count < 1");
    if (max_steps < 0 || patience < 1)
        return set_error(BM_EINVAL, "max_steps %lld / patience %d", (long long)max_steps, patience);
    for (int i = 0; i < count; i++) {
        bm_ctx *c = ctxs[i];
        if (!c || !c->batch || !c->batch->configured)
            return set_error(BM_ESTATE, "context %d: bm_refine_configure not called", i);
        if (c->device != ctxs[0]->device) return set_error(BM_EINVAL, "contexts on different devices");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == c) return set_error(BM_EINVAL, "context %d given twice", i);
    }
    BM_HIP(hipSetDevice(ctxs[0]->device));
    int rc;
    std::vector<int64_t> done(count, 0);
    std::vector<int> chunk(count, 0);
    auto enqueue = [&](int i) {
        chunk[i] = (int)std::min<int64_t>(ctxs[i]->batch->prm.steps_per_check, max_steps - done[i]);
        return refine_enqueue(ctxs[i], chunk[i]);
    };
    for (int i = 0; i < count; i++) {
        memset(&stats[i], 0, sizeof stats[i]);
        if ((rc = refine_start(ctxs[i])) || (rc = enqueue(i))) return rc;
    }
    std::vector<int> live(count);
    for (int i = 0; i < count; i++) live[i] = i;
    int first_error = BM_OK;
    while (!live.empty()) {
        std::vector<int> next;
        bool progressed = false;
        for (int i : live) {
            bm_ctx *c = ctxs[i];
            const int pr = refine_poll(c, false);
            if (pr < 0) return pr;
            if (pr == 0) {  // still running
                next.push_back(i);
                continue;
            }
            progressed = true;
            done[i] += chunk[i];
            bm_refine_stats &st = stats[i];
            refine_stats(c, &st);
            bool fin = !st.verified || done[i] >= max_steps || (move_budget > 0 && st.moves >= move_budget);
            if (!fin && st.idle_steps >= patience) {
                st.converged = 1;
                fin = true;
            }
            if (!fin) {
                if ((rc = enqueue(i))) return rc;
                next.push_back(i);
                continue;
            }
            if ((rc = refine_finish(c))) return rc;
            if (!st.verified && !first_error)
                first_error = set_error(BM_ERANGE, "context %d: exact score check failed at step %lld", i,
                                        (long long)st.bad_step);
        }
        live.swap(next);
        if (!progressed && !live.empty()) std::this_thread::yield();
    }
    return first_error;
}

extern "C" int bm_refine_history(bm_ctx *c, int64_t first, int64_t count, bm_refine_step *out) {
    if (!c || (count > 0 && !out)) return set_error(BM_EINVAL, "NULL argument");
    if (!c->batch || !c->batch->configured) return set_error(BM_ESTATE, "bm_refine_configure not called");
    const int64_t steps = (int64_t)c->batch->hst->step;
    if (first < 0 || count < 0 || first + count > steps || (steps - first) > kHist)
        return set_error(BM_ERANGE, "steps [%lld, %lld) not held (last %d of %lld kept)", (long long)first,
                         (long long)(first + count), kHist, (long long)steps);
    BM_HIP(hipSetDevice(c->device));
    // the requested steps occupy at most two runs of the ring
    for (int64_t done = 0; done < count;) {
        const int64_t slot = (first + done) % kHist;
        const int64_t n = std::min<int64_t>(count - done, kHist - slot);
        if (int rc_ = copy_sync(out + done, c->batch->hist + slot, sizeof(bm_refine_step) * (size_t)n,
                                hipMemcpyDeviceToHost))
            return rc_;
        done += n;
    }
    return BM_OK;
}

extern "C" int bm_refine_log(bm_ctx *c, int32_t capacity) {
    if (!c) return set_error(BM_EINVAL, "ctx is NULL");
    if (!c->batch || !c->batch->configured) return set_error(BM_ESTATE, "bm_refine_configure not called");
    if (capacity < 0) return set_error(BM_EINVAL, "capacity %d", capacity);
    BM_HIP(hipSetDevice(c->device));
    BM_HIP(hipStreamSynchronize(c->stream));
    return refine_set_log(c, capacity);
}

extern "C" int bm_refine_moves(bm_ctx *c, int64_t *count, int32_t *recs, int64_t capacity) {
    if (!c || !count) return set_error(BM_EINVAL, "NULL argument");
    if (!c->batch || !c->batch->configured) return set_error(BM_ESTATE, "bm_refine_configure not called");
    bm_batch_state *b = c->batch;
    const int64_t held = std::min<int64_t>(b->hst->log_count, b->log_cap);
    *count = b->hst->log_count;
    if (recs && capacity > 0 && held > 0) {
        BM_HIP(hipSetDevice(c->device));
        if (int rc_ = copy_sync(recs, b->log, (size_t)std::min(held, capacity) * sizeof(MoveRec), hipMemcpyDeviceToHost))
            return rc_;
    }
    return BM_OK;
}

extern "C" int bm_refine_config(bm_ctx *c, int32_t *rot3_lanes, int32_t *permset_threads) {
    if (!c || !rot3_lanes || !permset_threads) return set_error(BM_EINVAL, "NULL argument");
    *rot3_lanes = c->rot3_lanes;
    *permset_threads = c->permset_threads;
    return BM_OK;
}
