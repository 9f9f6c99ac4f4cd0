// Internal declarations of the MI355X move-delta engine (gfx950, wave64).
//
// Device-resident state of one context (one graph pair):
//   W2   int16 [2(n+1) x ld]     rows 0..n: G2 dense adjacency (last line per
//                                pair wins); rows n+1..2n+1: its transpose
//   Q    int16 [2(n+1) x ld]     rows 0..n     : P[x][y]  = W2[p(x)][p(y)]
//                                rows n+1..2n+1: PT[x][y] = P[y][x]
//   comb lists, one per vertex v (out entries of v, then in entries of v):
//        r  = n+1+d for an out entry v->d  (row of PT: Q[r][b] = P[b][d])
//        r  = s     for an in entry  s->v  (row of P : Q[r][b] = P[s][b])
//        w  = the CSV weight, dw = (short)weight of the line that owns the
//        dense G1 cell for that ordered pair (0 for earlier duplicates)
//   diag int32 [n+1]             X(v,v), X(u,v) = sum_{e in comb(u)} min(w_e, Q[r_e][v])
// With these, Delta_B(a,b) = X(a,b) + X(b,a) - X(a,a) - X(b,b) + Corr(a,b),
// where Corr collects the skipped a<->b list entries and the two dense cells.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "bmatch.h"

namespace bm {

constexpr int kWave = 64;

struct TimerRec {
    int kernel;  // index into kernel name table
    hipEvent_t start, stop;
};

struct KernelStat {
    std::string name;
    int64_t launches = 0;
    double ms = 0.0;
};

}  // namespace bm

struct bm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;

    // sizes
    int32_t n = 0;    // largest vertex id
    int32_t N1 = 0;   // n + 1
    int32_t ld = 0;   // row stride (elements) of W2 and Q, multiple of 64
    int64_t e1 = 0;   // G1 CSV lines
    int64_t e1_v0 = 0; // of them, lines leaving vertex 0: the head of the flat out-line list, not scored
    int64_t ncomb = 0;
    int64_t ndd = 0;
    bool graphs = false;
    bool mapped = false;
    bool diag_valid = false;
    bool nonneg = false;  // all weights >= 0 (packed-accumulate fast paths)
    bool g2nonneg = false;  // every dense G2 cell >= 0 as int16 (so is every entry of P)
    bool g1nonneg = false;  // every G1 weight >= 0
    bool dense_equal = false;  // every G1 line is its own dense cell: comb lists = dense-form lists
    int32_t max_w1 = 0;

    // device buffers
    int16_t *W2 = nullptr;
    int16_t *Q = nullptr;
    int32_t *map = nullptr;
    int32_t *cptr = nullptr, *cr = nullptr, *cw = nullptr, *cdw = nullptr;
    int32_t *crw = nullptr;     // (r, w) pairs of the comb lists, 8 bytes per entry
    int32_t *cmid = nullptr;    // cmid[v]: first in entry of comb(v)
    int32_t *cneg = nullptr;    // cneg[v] = sum over comb(v) of min(w, 0)
    int32_t *rowpart = nullptr;  // part rows of the sparse row kernel: rows x parts x (n+1) ints (lazy, grows)
    size_t rowpart_bytes = 0;
    unsigned long long *grpres = nullptr;   // kRowGroupMax keys of a row group (lazy)
    unsigned long long *hgrpres = nullptr;  // pinned host copy of those keys (lazy)
    int32_t *ddptr = nullptr, *ddmid = nullptr, *ddr = nullptr, *ddw = nullptr;
    // dense-form lists (lazy): the de-duplicated entries whose int16 G1 cell is
    // positive, without self-loops and vertex 0; weights are the int16 cells
    int32_t *dfptr = nullptr, *dfmid = nullptr, *dfr = nullptr, *dfw = nullptr;
    int32_t *dfrw = nullptr;    // (r, w) pairs of the same lists
    int32_t *dfneg = nullptr;   // zeros: the sum of min(w, 0) over lists of positive weights
    int32_t *dfdiag = nullptr;  // X(v,v) over the dense-form lists
    int32_t *dfes = nullptr, *dfed = nullptr, *dfew = nullptr;  // their out entries as flat lines (src, dst, cell)
    int64_t ndfout = 0;
    bool dfdiag_valid = false;
    int64_t ndf = 0;
    int32_t dfmaxw = 0;         // largest dense-form weight
    // 3-vertex table form (lazy): X over the dense-form lists for all pairs
    // (tile-major like X) and the dense-form G1 cells as an int16 matrix [n+1 x ld]
    int32_t *dfX = nullptr;
    int16_t *dfW1 = nullptr;
    int32_t *dfXP = nullptr;    // (Xd, cell << 16 | P) per pair, row-major [n+1 x ld] int2
    int32_t *esrc = nullptr, *edst = nullptr, *ew = nullptr, *edw = nullptr;
    // G2 de-duplicated (last line wins) out- and in-lists, packed (vertex << 16 | (uint16)weight)
    int32_t *g2optr = nullptr, *g2iptr = nullptr;
    uint32_t *g2o = nullptr, *g2i = nullptr;
    int64_t ng2 = 0;
    bool g2_vertex0 = false;  // a G2 line names vertex 0
    int32_t *pinv = nullptr;  // pinv[g] = the vertex mapped to g (-1 if none); valid when sparse_ok
    // the G2 lists with targets as P coordinates (p^-1), 0xFFFF = none; one allocation, g2iy = g2oy + ng2
    uint32_t *g2oy = nullptr, *g2iy = nullptr;
    // the sparse sweep's copy of the translated lists (lazy, rewritten before each sparse
    // sweep): every list padded to whole 16-byte quads of 4 words, out-lists then
    // in-lists, an untranslated or padding word replaced by N1 << 16 (a zero term
    // aimed at the row accumulator's spare slot N1)
    uint32_t *g2q = nullptr;
    int32_t *qoptr = nullptr, *qiptr = nullptr;  // quad index of each out- and in-list in g2q
    int32_t *g2qpos = nullptr;  // per word of g2oy / g2iy (2 ng2): its place in g2q
    int64_t nquads = 0;
    // crw and cdw with each row's entries in one global order (the sparse sweep's walk)
    int32_t *crws = nullptr, *cdws = nullptr;
    int32_t *pdg = nullptr;  // P[x][x] (the sparse sweep, written with g2q)
    bool g2q_small = false;  // every list weight in [0, 16383]: word >> 14 = 4 * target
    int32_t *g2o_at_in = nullptr;   // per in-list entry (line u->g): its index in the out-lists
    int32_t *g2i_at_out = nullptr;  // per out-list entry (line g->v): its index in the in-lists
    bool g2y_dirty = true;
    bool sparse_ok = false;   // p injective on mapped vertices and !g2_vertex0
    bool all_mapped = false;  // sparse_ok and p a bijection of 1..n (moves keep it): every G2 target translates
    int sweep_kind = 0;       // all-pairs gain kernel: 0 = not tuned, 1 = dense, 2 = sparse
    float sparse_ms = 0.f;
    int32_t *diag = nullptr;
    int32_t *X = nullptr;  // all-pairs gain matrix (lazy)
    int64_t ldx = 0;
    int gain_cols = 0;        // all-pairs slab width per lane (1, 2, 4), 0 = not tuned yet
    float gain_ms[3] = {0, 0, 0};  // autotuner timings of widths 1, 2, 4
    bool sparse_attr = false;      // dynamic-LDS attribute of the sparse gain kernel set
    bool gather_attr = false;      // dynamic-LDS attribute of the P / P^T rebuild kernel set
    int4 *plan = nullptr;          // sparse-kernel row plan: (row, entry lo, entry hi, flags)
    int64_t nplan = 0;
    int32_t *split_rows = nullptr; // rows the plan splits over several blocks
    int nsplit = 0;
    int dense_len = 1 << 30;       // sparse kernel: G2 lists longer than this stream their Q row
    int g2_maxlen = 0;             // longest de-duplicated G2 out- or in-list
    int rot3_lanes = 16;           // lanes per triple of the 3-vertex kernel (8, 16, 32)
    int permset_threads = 256;     // workgroup size of the k-set kernel (64, 256)
    bool rot3_tuned = false;       // rot3_lanes chosen by the autotuner for this graph pair
    int permset_threads_k = 0;     // the k permset_threads was tuned for (0: none)
    uint8_t *vsimple = nullptr;    // 1: no self-loop and no repeated line touches the vertex
    unsigned long long *keys = nullptr;  // N1 + 8 reduction slots
    void *stage = nullptr;               // move-list staging (lazy, grows)
    unsigned long long *hres = nullptr;  // pinned host slots for small results (8 x 8 bytes)
    void *hrows = nullptr;               // pinned host copy of bm_eval_rows' output (grown on demand)
    size_t hrows_bytes = 0;
    size_t stage_bytes = 0;
    int64_t dev_bytes = 0;

    // batched refinement (bm_batch_*)
    struct bm_batch_state *batch = nullptr;

    // device-resident candidate list (bm_moves_*)
    int32_t mv_k = 0;
    int64_t mv_count = 0, mv_cap = 0;
    int32_t *mv_verts = nullptr, *mv_delta = nullptr, *mv_aux = nullptr;
    int *mv_flag = nullptr;

    // counts host-side mapping changes (set_mapping, apply_*, the end of a batched run)
    uint64_t map_epoch = 0;
    // P / P^T lag the mapping: a batched run whose step reads W2 through the
    // mapping does not keep them current; ensure_p rebuilds them on next use
    bool p_stale = false;

    // host mirror of the mapping
    std::vector<int32_t> hmap;
    std::vector<int32_t> hpinv;
    std::vector<int32_t> hcdeg;   // comb-list length per vertex
    std::vector<int32_t> hdfdeg;  // dense-form list length per vertex (with the lists)
    // host copies behind the de-duplicated view: comb offsets, out/in split, neighbour + weight per entry,
    // whether the entry is the last of its (direction, neighbour) pair, and the view's own offsets; its
    // device copy (ddptr ...) is made on first use (ensure_dd)
    std::vector<int32_t> hcptr, hcmid, hcr, hcw, hddptr, hddmid;
    std::vector<uint8_t> hlast;
    std::vector<int32_t> hg2deg;  // G2 in- plus out-degree per G2 vertex (de-duplicated)
    std::vector<float> hvscore;   // perturbation weights per G1 vertex (computed on first use)
    std::vector<int32_t> hg2optr, hg2iptr;  // host copies of the G2 list offsets

    // kernel timing
    bool timing = false;
    std::vector<bm::KernelStat> kstats;
    std::vector<bm::TimerRec> pending;
    std::vector<hipEvent_t> ev_pool;
};

namespace bm {

int set_error(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

#define BM_HIP(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return bm::hip_fail(e_, #call); \
    } while (0)

// Held across a stream capture and around the engine's allocation, free and
// synchronous-copy calls (and stream / graph destruction): made from another
// thread in HIP's default capture mode, such a call would invalidate a
// capture in progress.
std::mutex &sync_mutex();

// device blocks through the process-wide cache (free only blocks no queued work still uses)
hipError_t block_alloc(int device, void **p, size_t bytes);
void block_free(void *p);
int dev_alloc(bm_ctx *c, void **p, size_t bytes);
void dev_free(bm_ctx *c, void **p, size_t bytes);
// hipMemcpy under sync_mutex: a blocking copy runs on the null stream, which
// waits for every blocking stream, a capturing one included
int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
int ensure_stage(bm_ctx *c, size_t bytes);

// kernel timing helpers (no-ops unless c->timing)
int timer_begin(bm_ctx *c, const char *name);
void timer_end(bm_ctx *c, int token);

int ensure_diag(bm_ctx *c);
int ensure_dense_lists(bm_ctx *c);
// the de-duplicated view (last line of each (direction, neighbour) pair per vertex) on the host, and on the device
void host_dd(const bm_ctx *c, std::vector<int32_t> &r, std::vector<int32_t> &w);
int ensure_dd(bm_ctx *c);
// out[v] = sum over the list of v of min(w, Q[r][v]), for every v
int launch_list_diag(bm_ctx *c, int32_t *out, const int32_t *ptr, const int32_t *r, const int32_t *w);
// X[u][v] = sum over the list of u of min(w, Q[r][v]) for all u, v (tile-major,
// stride ld), lists of non-negative weights <= max_w given as (r, w) pairs
// (the packed partial sums also need P >= 0: c->g2nonneg)
int launch_list_gain(bm_ctx *c, const int32_t *ptr, const int32_t *rw, int32_t *X, int max_w);
int launch_check(const char *what);

// launchers shared across translation units (device pointers, context stream)
int launch_eval_pairs(bm_ctx *c, const int32_t *dA, const int32_t *dB, int64_t m, int32_t *dD);
int launch_rot3(bm_ctx *c, const int32_t *dT, int64_t m, int mode, int32_t *dD, int32_t *dP);
// mult = 2: Delta_P as the reference (intra-set lines twice); 1: the exact score change
int launch_permset(bm_ctx *c, int k, const int32_t *dS, int64_t m, int32_t *dD, int32_t *dR, int *dOv, int mult = 2,
                   bool from_x = false, bool via_map = false);
int free_moves(bm_ctx *c);
// autotuner of the move kernels: times the variants on the given device list
int tune_rot3(bm_ctx *c, const int32_t *dT, int64_t m, int32_t *dD, int32_t *dP);
// Delta_G3 of a move list outside the batched step: the direct kernel or the
// table form (choose_rot3_table times both once per instance and size class);
// prepare = build the table form's all-pairs pass for the current mapping.
int choose_rot3_table(bm_ctx *c, const int32_t *dT, int64_t m, int32_t *dD, int32_t *dP, int *table);
int launch_rot3_moves(bm_ctx *c, const int32_t *dT, int64_t m, int mode, int32_t *dD, int32_t *dP, int table,
                      bool prepare);
// the dense-form lists and G1 cell matrix; Delta_G3 from the current X (needs
// c->dense_equal, the cell matrix and X matching the mapping)
int ensure_df_cells(bm_ctx *c);
// via_map: P read as W2[p(x)][p(y)] (P / P^T may lag the mapping)
int launch_rot3_x(bm_ctx *c, const int32_t *dT, int64_t m, int mode, int32_t *dD, int32_t *dP, bool via_map = false);
int tune_permset(bm_ctx *c, int k, const int32_t *dS, int64_t m, int32_t *dD, int32_t *dR, int *dOv, int mult,
                 bool from_x = false);
int free_batch(bm_ctx *c);
// P and P^T from W2 and the device mapping; with patch_flag, when *patch_flag
// is 1 only the rows and columns of verts[0 .. *nverts) are rebuilt
int launch_rebuild_p(bm_ctx *c, const int *patch_flag = nullptr, const int32_t *verts = nullptr,
                     const int *nverts = nullptr);
// P and P^T whole when *flag is 1 (capturable)
int launch_rebuild_p_if(bm_ctx *c, const int *flag);
int launch_diag(bm_ctx *c);             // X(v,v) for every v
// S_opt into *out; via_map reads the cells as W2[p(x)][p(y)] instead of P[x][y]
int launch_score(bm_ctx *c, unsigned long long *out, bool via_map = false);
int ensure_p(bm_ctx *c);                // rebuilds P / P^T when they lag the mapping
int launch_refresh_g2y(bm_ctx *c);      // G2 lists through p^-1
int sweep_prepare(bm_ctx *c);           // allocations and kernel choice of the all-pairs sweep
// row-best keys into keys[0..n], global best into keys[n+1]; with run_flag the
// sparse gain kernels run only when *run_flag != 0 (X stale), the combine always
// skip_flag (device): when it reads nonzero, the row-best keys are left as they are
int sweep_enqueue(bm_ctx *c, int32_t *dmat, const int *run_flag = nullptr, const int *skip_flag = nullptr);
bool sweep_is_sparse(const bm_ctx *c);  // the sweep runs the sparse gain kernel

// batched refinement, split for the serving loop: start a run, enqueue a chunk
// of steps (returns at once), poll / wait for it, read the counters, finish
int refine_configure(bm_ctx *c, const bm_refine_params *p);
int refine_start(bm_ctx *c);
int refine_enqueue(bm_ctx *c, int steps);
int refine_poll(bm_ctx *c, bool wait);
void refine_stats(bm_ctx *c, bm_refine_stats *out);
int refine_finish(bm_ctx *c);

// device-side row/column permutation of P and PT for a vertex set
int permute_rows_cols(bm_ctx *c, int k, const int32_t *set, const int32_t *src_of);

}  // namespace bm

// Launch helper: enqueues `launch` on c->stream, timed under `name`.
#define BM_LAUNCH(c, name, ...)                        \
    do {                                               \
        int tok_ = bm::timer_begin((c), (name));       \
        __VA_ARGS__;                                   \
        bm::timer_end((c), tok_);                      \
        int rc_ = bm::launch_check(name);              \
        if (rc_) return rc_;                           \
    } while (0)
