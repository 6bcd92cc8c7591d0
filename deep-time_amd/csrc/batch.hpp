// The batched solve (dtd_solve_batch): many independent grids of different shapes, each solved as
// dtd_solve_rel_device solves one (the operator and its levels, the splitting's first guess, f64
// conjugate gradients preconditioned by the f32 V-cycle, the put-back, the true residual), with one
// launch per pass and level over all of them.  Included once, by dtdiff.hip.
//
// Reference files (relative to the deep-time repository root):
//   D  crates/deeptime-core/src/surface/diffusion.rs
//   MG crates/deeptime-core/src/surface/diffusion/multigrid.rs
//
// The reference solves one grid a call (D:77-116); a batch here is that call on every grid.
// Layout: the grids' f64 vectors lie back to back, each starting on a multiple of kCgCells cells;
// each level of the cycle has one pool holding every grid's copy of that level.  A launch takes a
// table of (grid, block of that grid) pairs, so its blocks never straddle two grids, and writes one
// partial sum per block; a step kernel with one block per grid adds each grid's partials in block
// order into that grid's CgState.  Band heights depend on the grid's shape alone, so a grid's
// answer is the same bits whatever else is in the batch.  An iteration is the single grid's: the
// direction pass stores no A p, the finest level's fused pre-smoothing updates the residual
// (r -= alpha A p, A p and the level's records formed from p and kd) and leaves r.r's partials,
// and the residual and direction buffers alternate with the parity; a grid of one level (its
// finest level is the coarsest) keeps the stored A p and the separate update.  A grid whose loop has stopped skips
// every later pass (its state says done) while the others go on; the host loop ends when all
// have stopped.
#pragma once

#include "context.hpp"
#include "kernels.hpp"

namespace dtd {

// ---- what one launch knows of each grid ------------------------------------------------------

struct BGrid {  // one grid's f64 side
    Grid g;
    int wrap_x;
    int th;  // band height of the streamed f64 passes
    double dt, s2;
    const double *kd, *b;  // the caller's arrays (b: h on entry, read until the end)
    double *h;             // the caller's array, written last
    double *x, *r, *p0, *p1, *ap;
    double *ap_dir;             // where the direction pass stores A p (null: not stored)
    float *r32, *z32, *coef32;  // the finest f32 level (r32 null when the cycle reads r itself)
    CgState *st;
};

struct BLevel {      // one grid's level l of the f32 cycle
    Lvl<float> L, C;  // C: level l + 1 (unused on the grid's coarsest level)
    int th, th_post;  // band heights of the fused pre- and post-smoothing on this level
    const double *dot;  // the f64 residual on the finest level (the r.z sum), else null
    const CgState *st;
    // The finest level of a grid with more levels: the two residual buffers (r, r2) and the two
    // direction buffers (p0, p1); iteration parity q reads r64[q], writes r64[q ^ 1], and forms
    // A p from pd[q].
    double *r64[2];
    const double *pd[2];
};

struct BInit {  // a solve's start for one grid
    double tol2, rel2;
    int active, pad;
};

// Band height of a streamed kernel on an h-row level of a batch: a function of the shape alone,
// by the policy the batch was planned with (the built-in one: 16, 8 or 4 rows).
inline int batch_band(int h, const BatchPolicy &p = BatchPolicy{}) { return p.level(h); }

// ---- kernels: a table maps each block to (grid, block of the grid) ---------------------------

__global__ __launch_bounds__(kBlock) void kb_operator(const BGrid *__restrict__ t,
                                                      const int2 *__restrict__ blk) {
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    const int c = m.y * kBlock + threadIdx.x;
    if (c < q.g.n) operator_cell(q.g, q.kd, q.dt, q.s2, nullptr, q.coef32, nullptr, c);
}

// Level l + 1 from level l, one coarse cell a thread (t: the grids' level l).
__global__ __launch_bounds__(kBlock) void kb_coarsen(const BLevel *__restrict__ t,
                                                     const int2 *__restrict__ blk, float scale) {
    const int2 m = blk[blockIdx.x];
    const BLevel &q = t[m.x];
    const int k = m.y * kBlock + threadIdx.x;
    if (k < q.C.n) coarsen_cell(q.L, q.C, scale, k);
}

// Level 1 of a grid's cycle straight from kd (coarsen_kd_walk, as k_coarsen_kd); the level-0 task
// of grid g is task g; its blocks as k_coarsen_kd's grid, row-major.
__global__ __launch_bounds__(kBlock) void kb_coarsen_kd(const BGrid *__restrict__ t,
                                                        const BLevel *__restrict__ lv0,
                                                        const int2 *__restrict__ blk,
                                                        float scale) {
    const int2 m = blk[blockIdx.x];
    const BLevel &q = lv0[m.x];
    const BGrid &G = t[m.x];
    const int ncb = (q.C.w + 63) / 64;
    const int ic = (m.y % ncb) * 64 + (threadIdx.x & 63);
    const int jc0 = ((m.y / ncb) * (kBlock / 64) + (int)(threadIdx.x >> 6)) * kCoarsenRows;
    if (jc0 >= q.C.h) return;  // the whole wave
    coarsen_kd_walk(q.L, q.C, scale, G.kd, G.dt, G.s2, ic, jc0);
}

// The caller's h = x.
__global__ __launch_bounds__(kBlock) void kb_copy_out(const BGrid *__restrict__ t,
                                                      const int2 *__restrict__ blk) {
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    const int c = m.y * kBlock + threadIdx.x;
    if (c < q.g.n) q.h[c] = q.x[c];
}

// The first guess (`lod`, D:370-489) with the reference's recurrences, one thread a line: the
// rows on the transposed grid, where a wave's 64 rows lie side by side in memory, then the columns
// in place.  Transposes through 32 x 32 tiles: kd into r (mode 0) and b into ap (mode 1), w x h;
// then ap back into x (mode 2).
__global__ __launch_bounds__(kBlock) void kb_transpose(const BGrid *__restrict__ t,
                                                       const int2 *__restrict__ blk, int mode) {
    __shared__ double tile[32][33];
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    const double *src = mode == 0 ? q.kd : (mode == 1 ? q.b : q.ap);
    double *dst = mode == 0 ? q.r : (mode == 1 ? q.ap : q.x);
    const int rows = mode == 2 ? q.g.w : q.g.h, cols = mode == 2 ? q.g.h : q.g.w;
    const int ntc = (cols + 31) / 32;
    transpose_tile(src, dst, rows, cols, (m.y / ntc) * 32, (m.y % ntc) * 32, tile);
}
__global__ __launch_bounds__(kLineBlock) void kb_lod_rows(const BGrid *__restrict__ t,
                                                          const int2 *__restrict__ blk) {
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    const int j = m.y * kLineBlock + threadIdx.x;
    const int h = q.g.h;
    if (j < h)
        lod_line_x(Line<const double>{q.r + j, h}, Line<double>{q.ap + j, h},
                   Line<double>{q.p0 + j, h}, Line<double>{q.p1 + j, h}, q.g.w, q.wrap_x, q.dt,
                   q.s2);
}
// Whether a grid's columns go by the partition method in place (the others' by one thread's
// recurrences a column): a function of the height alone.
__host__ __device__ inline bool cols_by_segments(size_t h) { return h % kSeg == 0 && h >= 2 * kSeg; }

// One thread a column: the recurrences down the column, or (cols_by_segments) the column's reduced
// system, whose equations kb_col_reduce left in p0.
__global__ __launch_bounds__(kLineBlock) void kb_lod_cols(const BGrid *__restrict__ t,
                                                          const int2 *__restrict__ blk) {
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    const int i = m.y * kLineBlock + threadIdx.x;
    if (i >= q.g.w) return;
    if (cols_by_segments(q.g.h))
        col_thomas(q.g.w, q.g.h, q.p0, i);
    else
        lod_col(q.g, q.kd, q.dt, q.s2, q.x, q.p0, i);
}

// The column segments of the grids whose columns go by the partition method: a block takes 64
// columns x 4 segments (mode 0: the reduced equations into p0; 1: the segments' values into x).
__global__ __launch_bounds__(kBlock) void kb_col_seg(const BGrid *__restrict__ t,
                                                     const int2 *__restrict__ blk, int mode) {
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    const int w = q.g.w, h = q.g.h, ncb = (w + 63) / 64;
    const int i = (m.y % ncb) * 64 + (threadIdx.x & 63), s = (m.y / ncb) * 4 + (threadIdx.x >> 6);
    if (i >= w || s >= h / kSeg) return;
    if (mode == 0)
        col_reduce(w, h, q.kd, q.x, q.dt / q.s2, q.p0, i, s);
    else
        col_back(w, h, q.kd, q.x, q.dt / q.s2, q.p0, i, s);
}

// The row pass of the first guess by the partition method, one wave a row in place of the rows'
// recurrences and their transposes: b into x (widths a multiple of 16 from 32 to 1024); a
// block takes 4 rows of one grid.
__global__ __launch_bounds__(kBlock) void kb_lod_rows_wave(const BGrid *__restrict__ t,
                                                           const int2 *__restrict__ blk) {
    __shared__ double sm[kBlock / 64][6 * 128];
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    const int wv = threadIdx.x >> 6, w = q.g.w;
    const int j = m.y * (kBlock / 64) + wv;
    const bool on = j < q.g.h;
    const size_t o = (size_t)(on ? j : 0) * w;
    lod_line_wave(w, q.wrap_x && w > 2, on, q.kd + o, q.b + o, q.x + o, q.dt / q.s2, sm[wv]);
}

// Whether a grid's rows go by the one-wave partition method (the others' by the recurrences on
// the transposed grid): a function of the width alone, so a grid's answer does not depend on the
// batch it is in.
inline bool rows_by_wave(size_t w) { return w % kSeg == 0 && w >= 32 && w <= 64 * kSeg; }

// A fresh state for each grid (inactive grids are marked done and skipped by every pass).
__global__ void kb_init(CgState *st, const BInit *__restrict__ in, const BGrid *__restrict__ t,
                        int n, int max_iter) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    CgState s{};
    s.tol2 = in[g].tol2;
    s.rel2 = in[g].rel2;
    s.max_iter = max_iter;
    s.skip = !in[g].active;
    s.done = s.skip ? 5 : 0;
    s.kd = t[g].kd, s.dt = t[g].dt, s.s2 = t[g].s2;
    st[g] = s;
}

// The first residual r = b - A x (x = b first when from_b); (r.r, b.b) per block.
__global__ __launch_bounds__(kBlock) void kb_sresid(const BGrid *__restrict__ t,
                                                    const int2 *__restrict__ blk, int from_b,
                                                    double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    if (q.st->skip) return;
    block_pair(sresid_band<float>(q.g, q.kd, q.dt, q.s2, q.b, from_b ? q.b : q.x, q.x, q.r, q.r32,
                                  q.th, block_wave(m.y)),
               parts, blockIdx.x, lds);
}

// rng[g] = (first block, block count) of grid g in the launch that left the partials.
template <int STEP>
__global__ void kb_step(const int2 *__restrict__ rng, CgState *st,
                        const double *__restrict__ parts) {
    __shared__ double lds[kReduceBlock / 64];
    CgState *s = st + blockIdx.x;
    if (s->skip) return;
    const int2 r = rng[blockIdx.x];
    cg_step_block<STEP>(parts + (size_t)(STEP == kStepRR0 ? 2 : 1) * r.x, r.y, s, lds);
}

__global__ __launch_bounds__(kBlock) void kb_sdir(const BGrid *__restrict__ t,
                                                  const int2 *__restrict__ blk, int parity,
                                                  double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    if (q.st->done) return;
    const double *pold = parity == 0 ? q.p1 : q.p0;
    double *pnew = parity == 0 ? q.p0 : q.p1;
    const double acc =
        sdir_band<float>(q.g, q.kd, q.dt, q.s2, q.z32, pold, pnew, q.ap_dir, q.x, q.th, q.st,
                         block_wave(m.y));
    const double sum = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = sum;
}

__global__ __launch_bounds__(kBlock) void kb_update(const BGrid *__restrict__ t,
                                                    const int2 *__restrict__ blk,
                                                    double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    if (q.st->done) return;
    const double acc = update_block<float>(q.g.n, q.ap_dir, q.r, q.r32, q.st->alpha, m.y);
    const double sum = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = sum;
}

template <int V>
__global__ __launch_bounds__(kBlock) void kb_spre(const BLevel *__restrict__ t,
                                                  const int2 *__restrict__ blk) {
    const int2 m = blk[blockIdx.x];
    const BLevel &q = t[m.x];
    if (q.st->done) return;
    spre_band<float, V>(q.L, q.L.z[0], q.C, q.th, block_wave(m.y));
}

template <int V>
__global__ __launch_bounds__(kBlock) void kb_spost(const BLevel *__restrict__ t,
                                                   const int2 *__restrict__ blk,
                                                   double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BLevel &q = t[m.x];
    if (q.st->done) return;
    const double acc = spost_band<float, V>(q.L, q.L.z[0], q.C.z[1], q.C.w, q.L.z[1], q.th_post,
                                            q.dot, block_wave(m.y));
    if (q.dot) {
        const double sum = block_sum<kBlock>(acc, lds);
        if (threadIdx.x == 0) parts[blockIdx.x] = sum;
    }
}

// The finest level of the grids with more levels, as the single grid's fused smoothing runs it
// (records formed from kd; RM = kR64K: the first cycle, r the f64 residual r64[0]; RM =
// kRUpdateK: iteration `parity`, r64[parity] - alpha A p into r64[parity ^ 1] with r.r's
// partials, A p from pd[parity]).
template <int V, int RM>
__global__ __launch_bounds__(kBlock) void kb_spre0(const BLevel *__restrict__ t,
                                                   const int2 *__restrict__ blk, int parity,
                                                   double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BLevel &q = t[m.x];
    const CgState *st = q.st;
    if (st->done) return;
    R64 rio{};
    rio.in = q.r64[RM == kRUpdateK ? parity : 0];
    rio.out = q.r64[RM == kRUpdateK ? parity ^ 1 : 0];
    rio.p = q.pd[parity];
    if (rm_update(RM)) rio.alpha = st->alpha;
    rio.kd = st->kd, rio.dt = st->dt, rio.s2 = st->s2, rio.rs2 = 1.0 / st->s2;
    const double acc = spre_any<float, V, RM, 0>(q.L, q.L.z[0], q.C, q.th, block_wave(m.y), rio);
    if (rm_update(RM)) {
        const double sum = block_sum<kBlock>(acc, lds);
        if (threadIdx.x == 0) parts[blockIdx.x] = sum;
    }
}

// Its post-smoothing: the records from kd, r.z against the current residual r64[cur].
template <int V>
__global__ __launch_bounds__(kBlock) void kb_spost0(const BLevel *__restrict__ t,
                                                    const int2 *__restrict__ blk, int cur,
                                                    double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BLevel &q = t[m.x];
    const CgState *st = q.st;
    if (st->done) return;
    R64 rio{};
    rio.kd = st->kd, rio.dt = st->dt, rio.s2 = st->s2, rio.rs2 = 1.0 / st->s2;
    const double acc = spost_any<float, V, kR64K, 0>(q.L, q.L.z[0], q.C.z[1], q.C.w, q.L.z[1],
                                                      q.th_post, q.r64[cur], block_wave(m.y), rio);
    const double sum = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = sum;
}

// The bottom of each grid's cycle: its levels from the first of at most the bottom's size down,
// one block a grid in LDS (bottom_cycle; tab: the grid's first level in lvs, its level count, the
// grid).
__global__ __launch_bounds__(kBottomBlock) void kb_bottom(const Lvl<float> *__restrict__ lvs,
                                                          const int4 *__restrict__ tab,
                                                          const CgState *__restrict__ st) {
    const int4 m = tab[blockIdx.x];
    if (st[m.z].done) return;
    bottom_cycle<float>(lvs + m.x, 0, m.y);
}

// The coarsest level of each grid whose cycle ends here: one wave a grid.
__global__ __launch_bounds__(64) void kb_coarsest(const BLevel *__restrict__ t,
                                                  const int2 *__restrict__ blk,
                                                  double *__restrict__ parts) {
    __shared__ float zs[2][64];
    const BLevel &q = t[blk[blockIdx.x].x];
    if (q.st->done) return;
    coarsest_wave<float>(q.L, q.L.z[1], 16, q.dot, parts + blockIdx.x, zs);
}

__global__ __launch_bounds__(kBlock) void kb_finish1(const BGrid *__restrict__ t,
                                                     const int2 *__restrict__ blk,
                                                     double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    if (q.st->skip) return;
    block_pair(finish1_band(q.g, q.kd, q.dt, q.s2, q.b, q.x, q.p0, q.p1, q.ap, q.th, q.st,
                            block_wave(m.y)),
               parts, blockIdx.x, lds);
}

__global__ __launch_bounds__(kBlock) void kb_finish2(const BGrid *__restrict__ t,
                                                     const int2 *__restrict__ blk,
                                                     double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int2 m = blk[blockIdx.x];
    const BGrid &q = t[m.x];
    if (q.st->skip) return;
    block_pair(finish2_band(q.g, q.kd, q.dt, q.s2, q.b, q.ap, q.x, q.th, q.st, block_wave(m.y)),
               parts, blockIdx.x, lds);
}

// The put-back's sums into each grid's state; then each grid's (r.r, b.b) of the true residual.
__global__ void kb_putback(const int2 *__restrict__ rng, CgState *st,
                           const double *__restrict__ parts) {
    __shared__ double lds[kReduceBlock / 64];
    CgState *s = st + blockIdx.x;
    if (s->skip) return;
    const int2 r = rng[blockIdx.x];
    const double l = reduce_parts(parts + 2 * (size_t)r.x, r.y, 2, 0, lds);
    const double wg = reduce_parts(parts + 2 * (size_t)r.x, r.y, 2, 1, lds);
    if (threadIdx.x == 0) s->lost = l, s->weight = wg;
}
__global__ void kb_pair(const int2 *__restrict__ rng, const CgState *__restrict__ st,
                        const double *__restrict__ parts, double *__restrict__ out) {
    __shared__ double lds[kReduceBlock / 64];
    if (st[blockIdx.x].skip) return;
    const int2 r = rng[blockIdx.x];
    const double a = reduce_parts(parts + 2 * (size_t)r.x, r.y, 2, 0, lds);
    const double b = reduce_parts(parts + 2 * (size_t)r.x, r.y, 2, 1, lds);
    if (threadIdx.x == 0) out[2 * blockIdx.x] = a, out[2 * blockIdx.x + 1] = b;
}

// flag = 1 when every grid's loop has stopped.
// flag[0] = every grid has stopped; also into host (host memory the device can write) if given.
// With mask given (host memory the device can write), also mask[g] = grid g still runs and
// mask[n] = how many do.
__global__ __launch_bounds__(kBlock) void kb_all_done(const CgState *__restrict__ st, int n,
                                                      int *flag, int *host = nullptr,
                                                      int *mask = nullptr) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int open = 0;
    for (int g = threadIdx.x; g < n; g += kBlock) {
        const int o = st[g].done == 0;
        open += o;
        if (mask) reinterpret_cast<volatile int *>(mask)[g] = o;
    }
    if (open) atomicAdd(&cnt, open);
    __syncthreads();
    if (threadIdx.x == 0) {
        flag[0] = cnt == 0;
        if (mask) reinterpret_cast<volatile int *>(mask)[n] = cnt;
        __threadfence_system();
        if (host) *reinterpret_cast<volatile int *>(host) = cnt == 0;
    }
}

// ---- the plan of one batch shape -------------------------------------------------------------

struct BLaunch {  // one launch: its block table (device) and block count
    const int2 *blk = nullptr;
    int nblk = 0;
};

// The band height of every streamed launch in a pass that solves a few grids again.
constexpr int kRestartTh = 4;

// The launch tables of the iterations (and of a pass's first residual, first cycle and finish)
// over one set of grids: every grid of the batch, or the ones still running (the others' blocks
// left out, so that a launch dispatches only blocks with work).
struct BTables {
    BLaunch update_l, band_l;
    const int2 *rng_band = nullptr, *rng_rr = nullptr, *rng_rz = nullptr;
    int rr_base_pre1 = 0, rr_base_update = 0;  // r.r partials: pre2[0]'s, pre1[0]'s, the update's
    int rz_base_post1 = 0, rz_base_coarsest = 0;
    std::vector<BLaunch> pre2, pre1, post2, post1, coarsest;  // per level
    const int4 *bottom_tab = nullptr;
    int nbottom = 0;
    int ngrids = 0;  // grids with blocks in the tables
    // the grid and level tables the launches read (their band heights made the block tables)
    const BGrid *grids = nullptr;
    std::vector<const BLevel *> lv;
};

struct BatchWs {
    std::vector<long> key;  // (w, h, wrap) of each grid, omega and scale bits
    BatchPolicy policy;     // the band heights the launch tables were built with
    int ngrids = 0;
    int nlevels = 0;  // the deepest grid's level count
    size_t cells = 0;
    std::vector<int> nlev;       // levels of each grid
    std::vector<BGrid> grids;    // host copy of the f64-side table
    BGrid *d_grids = nullptr;
    std::vector<BLevel *> d_lv;  // per level: the table of every grid that has the level
    CgState *st = nullptr;
    BInit *init = nullptr;
    double *parts = nullptr, *pairs = nullptr;
    int *flag = nullptr;
    int2 *tables = nullptr;  // the set-up's block tables, one allocation
    // the set-up's launches: every cell (the answer's copy), the finest records of the grids of
    // one level (the others' level 1 comes from kd), the first guess, the coarsening per level
    BLaunch cells_l, oper_l, rows_l, rows4_l, cols_l, colseg_l, tiles_l;
    std::vector<BLaunch> coarsen_l;
    // the iterations' launches over every grid, and over the grids still running
    BTables all, part;
    int2 *part_tab = nullptr;  // room for part's tables (as many entries as all's)
    int4 *part_bottom = nullptr;
    size_t tab_entries = 0;
    // what the iterations' tables are built from: each level's tasks (grid and record), each
    // grid's first level its bottom cycles (its level count: none), the bottoms' table
    std::vector<std::vector<int>> lv_grid;
    std::vector<std::vector<BLevel>> lv;
    std::vector<int> b0;
    std::vector<int4> btab;
    // the bottoms: every grid's levels from its bottom on, the LDS bytes
    const Lvl<float> *bottom_lv = nullptr;
    size_t bottom_lds = 0;
    // The same grid and level tables with every band height kRestartTh, for the pass that solves
    // a few grids again: their launches then have kRestartTh-row bands, many more blocks a grid,
    // which shortens each launch's longest walk (such a pass is latency-bound).
    BGrid *d_grids_r = nullptr;
    std::vector<BLevel *> d_lv_r;
    std::vector<BGrid> grids_r;  // host copy, refreshed before each such pass
    std::vector<void *> owned;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // pinned host
    CgState *h_st = nullptr;
    double *h_pairs = nullptr;
    BInit *h_init = nullptr;
    int *h_mask = nullptr, *d_mask = nullptr;  // two slots of n + 1 (kb_all_done's mask)
    std::vector<int2> h_part_tab;               // part's tables as last uploaded
    std::vector<int4> h_part_bot;
};

void free_batch(dtd_ctx *ctx) {
    BatchWs *bw = ctx ? ctx->batch : nullptr;
    if (!bw) return;
    for (auto &g : bw->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (void *p : bw->owned) (void)hipFree(p);
    if (bw->h_st) (void)hipHostFree(bw->h_st);
    if (bw->h_pairs) (void)hipHostFree(bw->h_pairs);
    if (bw->h_init) (void)hipHostFree(bw->h_init);
    if (bw->h_mask) (void)hipHostFree(bw->h_mask);
    delete bw;
    ctx->batch = nullptr;
}

static std::vector<long> batch_key(const dtd_ctx *ctx, size_t n, const dtd_grid *gs) {
    std::vector<long> key;
    key.reserve(3 * n + 2);
    for (size_t k = 0; k < n; k++) {
        key.push_back((long)gs[k].w);
        key.push_back((long)gs[k].h);
        key.push_back(gs[k].wrap_x && gs[k].w > 1);
    }
    long om = 0, sc = 0;
    std::memcpy(&om, &ctx->opts.omega, sizeof(float));
    std::memcpy(&sc, &ctx->opts.scale, sizeof(float));
    key.push_back(om);
    key.push_back(sc);
    return key;
}

static inline size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// The iterations' launch tables of the grids with active[g] (every grid when active is null):
// block tables and each grid's ranges of partials into tab, whose device copy will sit at dtab,
// and the bottoms' table into bot (at dbot).  A grid left out has no blocks and empty ranges.
static void iteration_tables(const BatchWs &bw, const int *active, BTables &T,
                             std::vector<int2> &tab, int2 *dtab, std::vector<int4> &bot,
                             int4 *dbot, bool restart_th = false) {
    const int n = bw.ngrids, L = bw.nlevels;
    T.grids = restart_th ? bw.d_grids_r : bw.d_grids;
    T.lv.assign(bw.d_lv.begin(), bw.d_lv.end());
    if (restart_th) T.lv.assign(bw.d_lv_r.begin(), bw.d_lv_r.end());
    auto gth = [&](int g) { return restart_th ? kRestartTh : bw.grids[g].th; };
    tab.clear();
    bot.clear();
    auto on = [&](int g) { return !active || active[g] != 0; };
    auto add_launch = [&](BLaunch &bl, const std::vector<int> &blocks_of,
                          std::vector<int2> *rng) {
        // blocks_of[task] blocks for each task, in task order.
        const size_t at = tab.size();
        for (size_t t = 0; t < blocks_of.size(); t++) {
            if (rng) rng->push_back(make_int2((int)(tab.size() - at), blocks_of[t]));
            for (int b = 0; b < blocks_of[t]; b++) tab.push_back(make_int2((int)t, b));
        }
        bl.nblk = (int)(tab.size() - at);
        bl.blk = dtab ? dtab + at : nullptr;
    };
    T.ngrids = 0;
    for (int g = 0; g < n; g++) T.ngrids += on(g);
    std::vector<int> cnt(n);
    for (int g = 0; g < n; g++)
        cnt[g] = on(g) && bw.nlev[g] <= 1 ? nblocks((long)bw.grids[g].g.n, kCgCells) : 0;
    std::vector<int2> rng_update, rng_band, rng_rz(n, make_int2(0, 0)), rng_rr(n, make_int2(0, 0));
    add_launch(T.update_l, cnt, &rng_update);
    for (int g = 0; g < n; g++)
        cnt[g] = on(g) ? stream_grid<1, 1>(bw.grids[g].g.w, bw.grids[g].g.h, gth(g)) : 0;
    add_launch(T.band_l, cnt, &rng_band);
    // The smoothing launches of the grids for which level l is neither the coarsest nor cycled
    // by the grid's bottom, split by lane width; the coarsest launches.
    T.pre2.assign(L, BLaunch{}), T.pre1.assign(L, BLaunch{});
    T.post2.assign(L, BLaunch{}), T.post1.assign(L, BLaunch{});
    T.coarsest.assign(L, BLaunch{});
    for (int l = 0; l < L; l++) {
        const size_t nt = bw.lv[l].size();
        std::vector<int> c2(nt, 0), c1(nt, 0), cc(nt, 0), q2(nt, 0), q1(nt, 0);
        for (size_t t = 0; t < nt; t++) {
            const BLevel &q = bw.lv[l][t];
            const int g = bw.lv_grid[l][t];
            if (!on(g)) continue;
            const int th = restart_th ? kRestartTh : q.th;
            const int th_post = restart_th ? kRestartTh : q.th_post;
            const bool last = l + 1 == bw.nlev[g];
            if (!last) {
                if (l >= bw.b0[g]) continue;  // the grid's bottom cycles this level
                // c: the pre-smoothing's blocks, q: the post-smoothing's.
                if (q.L.w % 2 == 0)
                    c2[t] = stream_grid<2, 1>(q.L.w, q.L.h, th),
                    q2[t] = stream_grid<2, 1>(q.L.w, q.L.h, th_post);
                else
                    c1[t] = stream_grid<1, 2>(q.L.w, q.L.h, th),
                    q1[t] = stream_grid<1, 2>(q.L.w, q.L.h, th_post);
            } else if (l < bw.b0[g]) {
                if (q.L.n > 64) throw Error{DTD_ESTATE, "the coarsest level exceeds one wave"};
                cc[t] = 1;
            }
        }
        std::vector<int2> r2, r1, rc, s2, s1;
        add_launch(T.pre2[l], c2, l == 0 ? &s2 : nullptr);
        add_launch(T.pre1[l], c1, l == 0 ? &s1 : nullptr);
        add_launch(T.post2[l], q2, l == 0 ? &r2 : nullptr);
        add_launch(T.post1[l], q1, l == 0 ? &r1 : nullptr);
        add_launch(T.coarsest[l], cc, l == 0 ? &rc : nullptr);
        if (l == 0) {
            // The r.r partials of an iteration: pre2's blocks, then pre1's, then the separate
            // update's (grids of one level).
            T.rr_base_pre1 = T.pre2[0].nblk;
            T.rr_base_update = T.pre2[0].nblk + T.pre1[0].nblk;
            // The r.z partials: post2's blocks, then post1's, then the coarsest launch's.
            T.rz_base_post1 = T.post2[0].nblk;
            T.rz_base_coarsest = T.post2[0].nblk + T.post1[0].nblk;
            for (size_t t = 0; t < nt; t++) {
                const int g = bw.lv_grid[0][t];
                if (c2[t])
                    rng_rr[g] = s2[t], rng_rz[g] = r2[t];
                else if (c1[t])
                    rng_rr[g] = make_int2(T.rr_base_pre1 + s1[t].x, s1[t].y),
                    rng_rz[g] = make_int2(T.rz_base_post1 + r1[t].x, r1[t].y);
                else
                    rng_rr[g] = make_int2(T.rr_base_update + rng_update[g].x, rng_update[g].y),
                    rng_rz[g] = make_int2(T.rz_base_coarsest + rc[t].x, cc[t]);
            }
        }
    }
    const size_t at_rng = tab.size();
    tab.insert(tab.end(), rng_band.begin(), rng_band.end());
    tab.insert(tab.end(), rng_rr.begin(), rng_rr.end());
    tab.insert(tab.end(), rng_rz.begin(), rng_rz.end());
    T.rng_band = dtab ? dtab + at_rng : nullptr;
    T.rng_rr = dtab ? T.rng_band + n : nullptr;
    T.rng_rz = dtab ? T.rng_rr + n : nullptr;
    for (const int4 &b : bw.btab)
        if (on(b.z)) bot.push_back(b);
    T.nbottom = (int)bot.size();
    T.bottom_tab = dbot;
}

// The entries of the iterations' tables over every grid.
static size_t iteration_entries(const BatchWs &bw) {
    BTables T;
    std::vector<int2> tab;
    std::vector<int4> bot;
    iteration_tables(bw, nullptr, T, tab, nullptr, bot, nullptr);
    return tab.size();
}

// Builds the device buffers and launch tables for grids gs (all non-empty).
static void plan_batch(dtd_ctx *ctx, size_t n, const dtd_grid *gs) {
    auto key = batch_key(ctx, n, gs);
    if (ctx->batch && ctx->batch->key == key && ctx->batch->policy == ctx->batch_policy) return;
    free_batch(ctx);
    BatchWs *bw = new BatchWs();
    ctx->batch = bw;
    bw->key = key;
    bw->policy = ctx->batch_policy;
    bw->ngrids = (int)n;
    auto own = [&](size_t bytes) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, bytes ? bytes : 1));
        bw->owned.push_back(p);
        return p;
    };
    // Shapes, and the offsets of each grid's f64 cells and of its levels in the level pools.
    std::vector<std::vector<Shape>> sh(n);
    std::vector<size_t> off(n);
    size_t total = 0;
    for (size_t g = 0; g < n; g++) {
        sh[g] = shapes((int)gs[g].w, (int)gs[g].h, gs[g].wrap_x && gs[g].w > 1);
        bw->nlev.push_back((int)sh[g].size());
        bw->nlevels = std::max(bw->nlevels, (int)sh[g].size());
        off[g] = total;
        total += round_up(gs[g].w * gs[g].h, kCgCells);
        bw->cells += gs[g].w * gs[g].h;
    }
    const int L = bw->nlevels;
    std::vector<std::vector<size_t>> loff(L, std::vector<size_t>(n, 0));
    std::vector<size_t> ltotal(L, 0);
    for (int l = 0; l < L; l++)
        for (size_t g = 0; g < n; g++)
            if (l < bw->nlev[g]) {
                loff[l][g] = ltotal[l];
                ltotal[l] += round_up((size_t)sh[g][l].w * sh[g][l].h, 64);
            }
    double *x = (double *)own(total * 8), *r = (double *)own(total * 8);
    double *rnext = (double *)own(total * 8);  // the residual's second buffer
    double *p0 = (double *)own(total * 8), *p1 = (double *)own(total * 8);
    double *ap = (double *)own(total * 8);
    bw->st = (CgState *)own(n * sizeof(CgState));
    bw->init = (BInit *)own(n * sizeof(BInit));
    bw->pairs = (double *)own(2 * n * sizeof(double));
    bw->flag = (int *)own(sizeof(int));
    struct Pool {
        float *coef, *mass, *r, *z0, *z1;
    };
    std::vector<Pool> pool(L);
    for (int l = 0; l < L; l++) {
        pool[l].coef = (float *)own(kRec * ltotal[l] * 4);
        pool[l].mass = l == 0 ? nullptr : (float *)own(ltotal[l] * 4);
        pool[l].r = (float *)own(ltotal[l] * 4);
        pool[l].z0 = (float *)own(ltotal[l] * 4);
        pool[l].z1 = (float *)own(ltotal[l] * 4);
    }
    auto level = [&](size_t g, int l) {
        Lvl<float> v{};
        if (l >= bw->nlev[g]) return v;
        const size_t o = loff[l][g];
        v.w = sh[g][l].w, v.h = sh[g][l].h, v.wrap = sh[g][l].wrap, v.n = v.w * v.h;
        v.omega = ctx->opts.omega;
        v.coef = pool[l].coef + kRec * o;
        v.mass = l == 0 ? nullptr : pool[l].mass + o;
        v.r = pool[l].r + o;
        v.z[0] = pool[l].z0 + o;
        v.z[1] = pool[l].z1 + o;
        return v;
    };
    // The f64 side's table (the caller's arrays and dt are filled per call).
    bw->grids.resize(n);
    for (size_t g = 0; g < n; g++) {
        BGrid &q = bw->grids[g];
        q = BGrid{};
        const int w = (int)gs[g].w, h = (int)gs[g].h;
        q.g = Grid{w, h, (int)(gs[g].wrap_x && w > 1), w * h};
        q.th = bw->policy.pass(h);
        const size_t o = off[g];
        q.x = x + o, q.r = r + o, q.p0 = p0 + o, q.p1 = p1 + o, q.ap = ap + o;
        const Lvl<float> L0 = level(g, 0);
        // A grid of one level keeps the stored A p, the separate update and the f32 residual its
        // coarsest solve reads; the others' finest level reads the f64 residual itself.
        const bool fused = bw->nlev[g] > 1;
        q.ap_dir = fused ? nullptr : q.ap;
        q.r32 = fused ? nullptr : L0.r, q.z32 = L0.z[1], q.coef32 = L0.coef;
        q.st = bw->st + g;
    }
    bw->d_grids = (BGrid *)own(n * sizeof(BGrid));
    bw->d_grids_r = (BGrid *)own(n * sizeof(BGrid));
    // Level tables: every grid that has level l, in grid order (task index = position).
    std::vector<std::vector<int>> lv_grid(L);  // grid of each task of level l
    std::vector<std::vector<BLevel>> lv(L);
    for (int l = 0; l < L; l++) {
        for (size_t g = 0; g < n; g++) {
            if (l >= bw->nlev[g]) continue;
            BLevel t{};
            t.L = level(g, l);
            t.C = level(g, l + 1);
            t.th = l == 0 ? bw->policy.finest(true, t.L.h) : bw->policy.level(t.L.h);
            t.th_post = l == 0 ? bw->policy.finest(false, t.L.h) : t.th;
            t.dot = l == 0 ? bw->grids[g].r : nullptr;
            t.st = bw->st + g;
            if (l == 0) {
                const size_t o = off[g];
                t.r64[0] = r + o, t.r64[1] = rnext + o;
                t.pd[0] = p0 + o, t.pd[1] = p1 + o;
            }
            lv[l].push_back(t);
            lv_grid[l].push_back((int)g);
        }
        BLevel *d = (BLevel *)own(lv[l].size() * sizeof(BLevel));
        HIPCHK(hipMemcpy(d, lv[l].data(), lv[l].size() * sizeof(BLevel), hipMemcpyHostToDevice));
        bw->d_lv.push_back(d);
        std::vector<BLevel> lr = lv[l];
        for (BLevel &q : lr) q.th = q.th_post = kRestartTh;
        BLevel *dr = (BLevel *)own(lr.size() * sizeof(BLevel));
        HIPCHK(hipMemcpy(dr, lr.data(), lr.size() * sizeof(BLevel), hipMemcpyHostToDevice));
        bw->d_lv_r.push_back(dr);
    }
    // Each grid's bottom: the first level from 1 on that one block cycles in LDS, as the single
    // grid's cycle chooses it (at most bottom_cells() cells; that level and the coarser ones'
    // records, r and z within kBottomLds; at most kBottomLevels levels).  A grid of one level
    // has none (b0 = its level count).
    bw->b0.assign(n, 0);
    {
        std::vector<Lvl<float>> blv;
        for (size_t g = 0; g < n; g++) {
            const int nl = bw->nlev[g];
            bw->b0[g] = nl;
            if (nl < 2 || !fused_enabled()) continue;
            for (int l = 1; l < nl; l++) {
                size_t lds = 0;
                for (int k = l; k < nl; k++) lds += (kRec + 3) * (size_t)sh[g][k].w * sh[g][k].h * 4;
                if ((long)sh[g][l].w * sh[g][l].h <= bottom_cells() && lds <= kBottomLds &&
                    nl - l <= kBottomLevels) {
                    bw->b0[g] = l;
                    bw->btab.push_back(make_int4((int)blv.size(), nl - l, (int)g, 0));
                    for (int k = l; k < nl; k++) blv.push_back(level(g, k));
                    bw->bottom_lds = std::max(bw->bottom_lds, lds);
                    break;
                }
            }
        }
        if (!blv.empty()) {
            Lvl<float> *dl = (Lvl<float> *)own(blv.size() * sizeof(Lvl<float>));
            HIPCHK(hipMemcpy(dl, blv.data(), blv.size() * sizeof(Lvl<float>), hipMemcpyHostToDevice));
            bw->bottom_lv = dl;
        }
    }
    // The set-up's block tables.
    std::vector<int2> tab;
    struct Pending {
        BLaunch *l;
        size_t at;
    };
    std::vector<Pending> pend;
    auto add_launch = [&](BLaunch &bl, const std::vector<int> &blocks_of) {
        // blocks_of[task] blocks for each task, in task order.
        const size_t at = tab.size();
        for (size_t t = 0; t < blocks_of.size(); t++)
            for (int b = 0; b < blocks_of[t]; b++) tab.push_back(make_int2((int)t, b));
        bl.nblk = (int)(tab.size() - at);
        pend.push_back({&bl, at});
    };
    std::vector<int> cnt(n);
    for (size_t g = 0; g < n; g++) cnt[g] = nblocks((long)gs[g].w * gs[g].h);
    add_launch(bw->cells_l, cnt);
    for (size_t g = 0; g < n; g++) cnt[g] = bw->nlev[g] > 1 ? 0 : nblocks((long)gs[g].w * gs[g].h);
    add_launch(bw->oper_l, cnt);
    // The first guess's rows: by the one-wave partition method (rows4_l, four rows a block), or
    // by the recurrences on the transposed grid (rows_l, tiles_l); then every grid's columns.
    for (size_t g = 0; g < n; g++)
        cnt[g] = rows_by_wave(gs[g].w) ? 0 : nblocks((long)gs[g].h, kLineBlock);
    add_launch(bw->rows_l, cnt);
    for (size_t g = 0; g < n; g++)
        cnt[g] = rows_by_wave(gs[g].w) ? nblocks((long)gs[g].h, kBlock / 64) : 0;
    add_launch(bw->rows4_l, cnt);
    for (size_t g = 0; g < n; g++) cnt[g] = nblocks((long)gs[g].w, kLineBlock);
    add_launch(bw->cols_l, cnt);
    for (size_t g = 0; g < n; g++)
        cnt[g] = cols_by_segments(gs[g].h)
                     ? (int)(((gs[g].w + 63) / 64) * ((gs[g].h / kSeg + 3) / 4))
                     : 0;
    add_launch(bw->colseg_l, cnt);
    for (size_t g = 0; g < n; g++)
        cnt[g] = rows_by_wave(gs[g].w) ? 0
                                       : (int)(((gs[g].w + 31) / 32) * ((gs[g].h + 31) / 32));
    add_launch(bw->tiles_l, cnt);
    // Coarsening of level l + 1 from level l's tasks.
    bw->coarsen_l.resize(L);
    for (int l = 0; l < L; l++) {
        std::vector<int> cz(lv[l].size(), 0);
        for (size_t t = 0; t < lv[l].size(); t++) {
            if (!(l + 1 < bw->nlev[lv_grid[l][t]])) continue;
            const Lvl<float> &C = lv[l][t].C;
            // Level 1 from kd by kb_coarsen_kd's blocks; the others a thread a coarse cell.
            cz[t] = l == 0 ? ((C.w + 63) / 64) * ((C.h + (kBlock / 64) * kCoarsenRows - 1) /
                                                  ((kBlock / 64) * kCoarsenRows))
                           : nblocks(C.n);
        }
        add_launch(bw->coarsen_l[l], cz);
    }
    bw->tables = (int2 *)own(tab.size() * sizeof(int2));
    HIPCHK(hipMemcpy(bw->tables, tab.data(), tab.size() * sizeof(int2), hipMemcpyHostToDevice));
    for (auto &q : pend) q.l->blk = bw->tables + q.at;
    bw->lv = std::move(lv);
    bw->lv_grid = std::move(lv_grid);
    // The iterations' tables over every grid, and room for those over the running grids.
    {
        std::vector<int2> itab;
        std::vector<int4> ibot;
        const size_t ne = iteration_entries(*bw);
        int2 *d_all = (int2 *)own(ne * sizeof(int2));
        bw->part_tab = (int2 *)own(ne * sizeof(int2));
        bw->tab_entries = ne;
        int4 *d_bot = (int4 *)own((bw->btab.size() + 1) * sizeof(int4));
        bw->part_bottom = (int4 *)own((bw->btab.size() + 1) * sizeof(int4));
        iteration_tables(*bw, nullptr, bw->all, itab, d_all, ibot, d_bot);
        HIPCHK(hipMemcpy(d_all, itab.data(), itab.size() * sizeof(int2), hipMemcpyHostToDevice));
        if (!ibot.empty())
            HIPCHK(hipMemcpy(d_bot, ibot.data(), ibot.size() * sizeof(int4),
                             hipMemcpyHostToDevice));
    }
    // Partial sums: pairs for the widest launch (and the r.z launches side by side).
    const BTables &T = bw->all;
    size_t most = (size_t)std::max({T.band_l.nblk, T.rr_base_update + T.update_l.nblk,
                                    T.post2[0].nblk + T.post1[0].nblk + T.coarsest[0].nblk});
    bw->parts = (double *)own(2 * most * sizeof(double) + 16);
    HIPCHK(hipHostMalloc((void **)&bw->h_st, n * sizeof(CgState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&bw->h_pairs, 2 * n * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&bw->h_init, n * sizeof(BInit), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&bw->h_mask, 2 * (n + 1) * sizeof(int),
                         hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void **)&bw->d_mask, bw->h_mask, 0));
}

// ---- enqueueing ------------------------------------------------------------------------------

// The V-cycle of every grid (`vcycle`, MG:500-523): down the levels (pre-smoothing and
// restriction, or the coarsest solve of the grids that end there), then up (post-smoothing; the
// finest level's leaves the r.z partials).  parity < 0: the first cycle (the finest level reads
// the first residual); else iteration `parity`, whose finest pre-smoothing updates the residual
// first, after which the r.r step runs (before the finest post-smoothing reuses the partials).
static void batch_vcycle(const BatchWs &bw, const BTables &T, hipStream_t s, int parity) {
    const int L = bw.nlevels;
    const int par = parity < 0 ? 0 : parity;
    for (int l = 0; l < L; l++) {
        if (l == 0) {
            if (T.pre2[0].nblk) {
                if (parity < 0)
                    kb_spre0<2, kR64K><<<T.pre2[0].nblk, kBlock, 0, s>>>(T.lv[0], T.pre2[0].blk,
                                                                       par, bw.parts);
                else
                    kb_spre0<2, kRUpdateK><<<T.pre2[0].nblk, kBlock, 0, s>>>(
                        T.lv[0], T.pre2[0].blk, par, bw.parts);
                LAUNCH_CHECK();
            }
            if (T.pre1[0].nblk) {
                if (parity < 0)
                    kb_spre0<1, kR64K><<<T.pre1[0].nblk, kBlock, 0, s>>>(
                        T.lv[0], T.pre1[0].blk, par, bw.parts + T.rr_base_pre1);
                else
                    kb_spre0<1, kRUpdateK><<<T.pre1[0].nblk, kBlock, 0, s>>>(
                        T.lv[0], T.pre1[0].blk, par, bw.parts + T.rr_base_pre1);
                LAUNCH_CHECK();
            }
            if (parity >= 0) {
                kb_step<kStepRR><<<bw.ngrids, kReduceBlock, 0, s>>>(T.rng_rr, bw.st, bw.parts);
                LAUNCH_CHECK();
            }
        } else {
            if (T.pre2[l].nblk) {
                kb_spre<2><<<T.pre2[l].nblk, kBlock, 0, s>>>(T.lv[l], T.pre2[l].blk);
                LAUNCH_CHECK();
            }
            if (T.pre1[l].nblk) {
                kb_spre<1><<<T.pre1[l].nblk, kBlock, 0, s>>>(T.lv[l], T.pre1[l].blk);
                LAUNCH_CHECK();
            }
        }
        if (T.coarsest[l].nblk) {
            kb_coarsest<<<T.coarsest[l].nblk, 64, 0, s>>>(
                T.lv[l], T.coarsest[l].blk, bw.parts + (l == 0 ? T.rz_base_coarsest : 0));
            LAUNCH_CHECK();
        }
    }
    // Every grid's bottom, once its levels above have restricted into it.
    if (T.nbottom) {
        kb_bottom<<<T.nbottom, kBottomBlock, bw.bottom_lds, s>>>(bw.bottom_lv, T.bottom_tab,
                                                                  bw.st);
        LAUNCH_CHECK();
    }
    // The current residual after iteration `parity`: r64[parity ^ 1]; in the first cycle r64[0].
    const int cur = parity < 0 ? 0 : parity ^ 1;
    for (int l = L - 2; l >= 0; l--) {
        if (T.post2[l].nblk) {
            if (l == 0)
                kb_spost0<2><<<T.post2[0].nblk, kBlock, 0, s>>>(T.lv[0], T.post2[0].blk, cur,
                                                                 bw.parts);
            else
                kb_spost<2><<<T.post2[l].nblk, kBlock, 0, s>>>(T.lv[l], T.post2[l].blk,
                                                                bw.parts);
            LAUNCH_CHECK();
        }
        if (T.post1[l].nblk) {
            if (l == 0)
                kb_spost0<1><<<T.post1[0].nblk, kBlock, 0, s>>>(T.lv[0], T.post1[0].blk, cur,
                                                                 bw.parts + T.rz_base_post1);
            else
                kb_spost<1><<<T.post1[l].nblk, kBlock, 0, s>>>(T.lv[l], T.post1[l].blk,
                                                                bw.parts + T.rz_base_post1);
            LAUNCH_CHECK();
        }
    }
}

// One iteration of every grid still running (MG:630-733), then the all-stopped flag.  The grids
// of one level take the separate update (their r.r partials after the finest pre-smoothings').
static void batch_iteration(const BatchWs &bw, const BTables &T, int parity, hipStream_t s,
                            int *host_flag, int *mask) {
    const int n = bw.ngrids;
    kb_sdir<<<T.band_l.nblk, kBlock, 0, s>>>(T.grids, T.band_l.blk, parity, bw.parts);
    LAUNCH_CHECK();
    kb_step<kStepPAP><<<n, kReduceBlock, 0, s>>>(T.rng_band, bw.st, bw.parts);
    LAUNCH_CHECK();
    if (T.update_l.nblk) {
        kb_update<<<T.update_l.nblk, kBlock, 0, s>>>(T.grids, T.update_l.blk,
                                                      bw.parts + T.rr_base_update);
        LAUNCH_CHECK();
    }
    batch_vcycle(bw, T, s, parity);
    kb_step<kStepRZ><<<n, kReduceBlock, 0, s>>>(T.rng_rz, bw.st, bw.parts);
    LAUNCH_CHECK();
    kb_all_done<<<1, kBlock, 0, s>>>(bw.st, n, bw.flag, host_flag, mask);
    LAUNCH_CHECK();
}

static void batch_graphs(dtd_ctx *ctx, BatchWs &bw) {
    for (int k = 0; k < 2; k++) {
        if (bw.graph[k]) continue;
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            batch_iteration(bw, bw.all, k, ctx->stream, ctx->dev_flag + k,
                            bw.d_mask + k * (bw.ngrids + 1));
        } catch (...) {
            (void)hipStreamEndCapture(ctx->stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        HIPCHK(hipStreamEndCapture(ctx->stream, &gr));
        hipError_t e = hipGraphInstantiate(&bw.graph[k], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHK(e);
    }
}

// The iterations' tables over the grids with active[g] into bw.part, its device copy queued on
// the stream.  The part tables must not be in use by queued work (they are rebuilt at most once
// a pass, before the pass launches anything with them).
// restart_th: with every band kRestartTh rows (the grid and level tables of that height), unless
// the tables would not fit the room; then with the batch's own heights.
static void part_tables(dtd_ctx *ctx, BatchWs &bw, const int *active, bool restart_th = false) {
    iteration_tables(bw, active, bw.part, bw.h_part_tab, bw.part_tab, bw.h_part_bot,
                     bw.part_bottom, restart_th);
    if (restart_th && bw.h_part_tab.size() > bw.tab_entries)
        iteration_tables(bw, active, bw.part, bw.h_part_tab, bw.part_tab, bw.h_part_bot,
                         bw.part_bottom, false);
    if (bw.h_part_tab.size() > bw.tab_entries)
        throw Error{DTD_ESTATE, "the running grids' tables exceed the batch's"};
    HIPCHK(hipMemcpyAsync(bw.part_tab, bw.h_part_tab.data(), bw.h_part_tab.size() * sizeof(int2),
                          hipMemcpyHostToDevice, ctx->stream));
    if (!bw.h_part_bot.empty())
        HIPCHK(hipMemcpyAsync(bw.part_bottom, bw.h_part_bot.data(),
                              bw.h_part_bot.size() * sizeof(int4), hipMemcpyHostToDevice,
                              ctx->stream));
}

// One pass of the loop over the active grids (h_init says which, and their tolerances): the
// first residual (from b when from_b, else from x), the cycle, the iterations, the last update,
// the put-back and the true residual.  Leaves h_st and h_pairs filled.  part: the pass runs on
// bw.part (built for its active grids) with the iterations launched one by one; otherwise on
// every grid's tables and the captured iterations, until at most a quarter of the grids still
// run: then the remaining iterations launch only their blocks.
static void batch_pass(dtd_ctx *ctx, BatchWs &bw, int from_b, bool part) {
    hipStream_t s = ctx->stream;
    const int n = bw.ngrids;
    const BTables *T = part ? &bw.part : &bw.all;
    HIPCHK(hipMemcpyAsync(bw.init, bw.h_init, n * sizeof(BInit), hipMemcpyHostToDevice, s));
    kb_init<<<nblocks(n), kBlock, 0, s>>>(bw.st, bw.init, bw.d_grids, n, (int)ctx->opts.max_iter);
    LAUNCH_CHECK();
    kb_sresid<<<T->band_l.nblk, kBlock, 0, s>>>(T->grids, T->band_l.blk, from_b, bw.parts);
    LAUNCH_CHECK();
    kb_step<kStepRR0><<<n, kReduceBlock, 0, s>>>(T->rng_band, bw.st, bw.parts);
    LAUNCH_CHECK();
    batch_vcycle(bw, *T, s, -1);
    kb_step<kStepRZ0><<<n, kReduceBlock, 0, s>>>(T->rng_rz, bw.st, bw.parts);
    LAUNCH_CHECK();
    kb_all_done<<<1, kBlock, 0, s>>>(bw.st, n, bw.flag);
    LAUNCH_CHECK();
    bool eager = part || !ctx->use_graph;
    if (!eager) batch_graphs(ctx, bw);
    // The loop: a pass ends when every grid has stopped.  Iteration k + 1 is launched before the
    // flag after iteration k is read (the flag before the loop counts as iteration 0's), so the
    // device runs at most one iteration that every grid skips; the state's cap stops every grid by
    // max_iter.  The running grids after iteration k - 1 include every grid that iteration k + 1
    // finds running.
    HIPCHK(hipMemcpyAsync(&ctx->host_flag[1], bw.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(ctx->flag_ev[1], s));
    const int max_iter = (int)ctx->opts.max_iter;
    for (int k = 0; k <= max_iter + 1; k++) {
        const int par = k & 1;
        if (!eager)
            HIPCHK(hipGraphLaunch(bw.graph[par], s));
        else
            batch_iteration(bw, *T, par, s, ctx->dev_flag + par, bw.d_mask + par * (n + 1));
        HIPCHK(hipEventRecord(ctx->flag_ev[par], s));
        HIPCHK(hipEventSynchronize(ctx->flag_ev[par ^ 1]));
        if (ctx->host_flag[par ^ 1]) break;
        const int *mask = bw.h_mask + (par ^ 1) * (n + 1);
        if (T == &bw.all && k >= 1 && 4 * mask[n] <= n) {
            part_tables(ctx, bw, mask);
            T = &bw.part;
            eager = true;
        }
    }
    // The last update, the put-back and the true residual of every grid the pass ran.
    const BTables &F = part ? bw.part : bw.all;
    kb_finish1<<<F.band_l.nblk, kBlock, 0, s>>>(F.grids, F.band_l.blk, bw.parts);
    LAUNCH_CHECK();
    kb_putback<<<n, kReduceBlock, 0, s>>>(F.rng_band, bw.st, bw.parts);
    LAUNCH_CHECK();
    kb_finish2<<<F.band_l.nblk, kBlock, 0, s>>>(F.grids, F.band_l.blk, bw.parts);
    LAUNCH_CHECK();
    kb_pair<<<n, kReduceBlock, 0, s>>>(F.rng_band, bw.st, bw.parts, bw.pairs);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(bw.h_st, bw.st, n * sizeof(CgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(bw.h_pairs, bw.pairs, 2 * n * sizeof(double), hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
}

// dtd_solve_batch.
static void solve_batch(dtd_ctx *ctx, size_t count, const dtd_grid *gs_in, double rel_tol,
                        size_t *iterations, double *rels, dtd_batch_stats *stats) {
    if (ctx->opts.cycle_precision != DTD_CYCLE_F32 || ctx->opts.nu != 2 ||
        ctx->opts.transfer != DTD_TRANSFER_AGGREGATE || ctx->opts.coarsest != DTD_COARSEST_SWEEPS ||
        ctx->opts.sums != DTD_SUMS_BLOCKS)
        throw Error{DTD_EINVAL, "the batched solve runs the reference's f32 cycle with nu = 2, "
                                "its sums per block"};
    const bool relative = rel_tol > 0.0;
    if (!std::isfinite(rel_tol)) throw Error{DTD_EINVAL, "rel_tol must be finite"};
    // The non-empty grids take part; the others get 0 iterations and residual 0.
    std::vector<dtd_grid> gs;
    std::vector<size_t> idx;
    for (size_t k = 0; k < count; k++) {
        const dtd_grid &g = gs_in[k];
        check_grid(g.w, g.h);
        if (iterations) iterations[k] = 0;
        if (rels) rels[k] = 0.0;
        if (g.w * g.h == 0) continue;
        if (!g.h_dev || !g.kd_dev) throw Error{DTD_EINVAL, "null array in grid " + std::to_string(k)};
        if (!(std::isfinite(g.dx) && g.dx > 0.0 && std::isfinite(g.dt)))
            throw Error{DTD_EINVAL, "dx and dt must be finite (grid " + std::to_string(k) + ")"};
        if (!relative && !(std::isfinite(g.tol_m) && g.tol_m > 0.0))
            throw Error{DTD_EINVAL, "tol_m must be finite and positive (grid " +
                                        std::to_string(k) + ")"};
        gs.push_back(g);
        idx.push_back(k);
    }
    const size_t n = gs.size();
    if (n == 0) return;
    if (n > (size_t)INT32_MAX / 2) throw Error{DTD_ETOOBIG, "too many grids"};
    if (!ctx->batch_tuning) {
        auto it = ctx->batch_plans.find(batch_key(ctx, n, gs.data()));
        ctx->batch_policy = it != ctx->batch_plans.end() ? it->second : BatchPolicy{};
    }
    plan_batch(ctx, n, gs.data());
    BatchWs &bw = *ctx->batch;
    hipStream_t s = ctx->stream;
    for (size_t g = 0; g < n; g++) {
        BGrid &q = bw.grids[g];
        q.wrap_x = gs[g].wrap_x;
        q.dt = gs[g].dt;
        q.s2 = gs[g].dx * gs[g].dx;
        q.kd = gs[g].kd_dev;
        q.b = gs[g].h_dev;
        q.h = gs[g].h_dev;
    }
    HIPCHK(hipMemcpyAsync(bw.d_grids, bw.grids.data(), n * sizeof(BGrid), hipMemcpyHostToDevice,
                          s));
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    // The finest records only for the grids of one level (their coarsest solve reads them; the
    // others' finest smoothers form theirs from kd): the others' level 1 straight from kd.
    if (bw.oper_l.nblk) {
        kb_operator<<<bw.oper_l.nblk, kBlock, 0, s>>>(bw.d_grids, bw.oper_l.blk);
        LAUNCH_CHECK();
    }
    for (int l = 0; l + 1 < bw.nlevels; l++) {
        if (!bw.coarsen_l[l].nblk) continue;
        if (l == 0)
            kb_coarsen_kd<<<bw.coarsen_l[0].nblk, kBlock, 0, s>>>(
                bw.d_grids, bw.d_lv[0], bw.coarsen_l[0].blk, ctx->opts.scale);
        else
            kb_coarsen<<<bw.coarsen_l[l].nblk, kBlock, 0, s>>>(bw.d_lv[l], bw.coarsen_l[l].blk,
                                                               ctx->opts.scale);
        LAUNCH_CHECK();
    }
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    // Every grid active for the first pass.
    for (size_t g = 0; g < n; g++) {
        const double nn = (double)(gs[g].w * gs[g].h);
        bw.h_init[g] = relative ? BInit{0.0, rel_tol * rel_tol, 1, 0}
                                : BInit{gs[g].tol_m * gs[g].tol_m * nn, 0.0, 1, 0};
    }
    const bool lod_guess = ctx->opts.guess == DTD_GUESS_LOD;
    if (lod_guess) {
        if (bw.rows4_l.nblk) {
            kb_lod_rows_wave<<<bw.rows4_l.nblk, kBlock, 0, s>>>(bw.d_grids, bw.rows4_l.blk);
            LAUNCH_CHECK();
        }
        if (bw.rows_l.nblk) {
            for (int mode = 0; mode < 2; mode++) {
                kb_transpose<<<bw.tiles_l.nblk, kBlock, 0, s>>>(bw.d_grids, bw.tiles_l.blk, mode);
                LAUNCH_CHECK();
            }
            kb_lod_rows<<<bw.rows_l.nblk, kLineBlock, 0, s>>>(bw.d_grids, bw.rows_l.blk);
            LAUNCH_CHECK();
            kb_transpose<<<bw.tiles_l.nblk, kBlock, 0, s>>>(bw.d_grids, bw.tiles_l.blk, 2);
            LAUNCH_CHECK();
        }
        if (bw.colseg_l.nblk) {
            kb_col_seg<<<bw.colseg_l.nblk, kBlock, 0, s>>>(bw.d_grids, bw.colseg_l.blk, 0);
            LAUNCH_CHECK();
        }
        kb_lod_cols<<<bw.cols_l.nblk, kLineBlock, 0, s>>>(bw.d_grids, bw.cols_l.blk);
        LAUNCH_CHECK();
        if (bw.colseg_l.nblk) {
            kb_col_seg<<<bw.colseg_l.nblk, kBlock, 0, s>>>(bw.d_grids, bw.colseg_l.blk, 1);
            LAUNCH_CHECK();
        }
    }
    HIPCHK(hipEventRecord(ctx->ev[2], s));
    batch_pass(ctx, bw, lod_guess ? 0 : 1, false);
    // Per grid: iterations, the true relative residual; then the grids that missed the target
    // solve again from their last answer, aiming below it by the factor they missed it by (at most
    // four times), the others left alone.
    std::vector<size_t> its(n, 0);
    std::vector<double> rel(n, 0.0), tolm(n, 0.0);
    size_t restarts = 0;
    auto collect = [&](bool first) {
        for (size_t g = 0; g < n; g++) {
            if (!first && !bw.h_init[g].active) continue;
            its[g] += (size_t)bw.h_st[g].its;
            const double rr = bw.h_pairs[2 * g], bb = bw.h_pairs[2 * g + 1];
            rel[g] = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
            if (first)
                tolm[g] = relative ? rel_tol * std::sqrt(bb) / std::sqrt((double)(gs[g].w * gs[g].h))
                                   : gs[g].tol_m;
        }
    };
    collect(true);
    std::vector<int> again_of(n, 0);
    for (int round = 0; relative && round < 4; round++) {
        bool any = false;
        for (size_t g = 0; g < n; g++) {
            const double bb = bw.h_pairs[2 * g + 1];
            const bool again = rel[g] > rel_tol && bb > 0.0;
            if (again) {
                tolm[g] *= 0.5 * rel_tol / rel[g];
                restarts++;
                any = true;
            }
            const double nn = (double)(gs[g].w * gs[g].h);
            bw.h_init[g] = BInit{tolm[g] * tolm[g] * nn, 0.0, again ? 1 : 0, 0};
            again_of[g] = again;
        }
        if (!any) break;
        // Only the grids solved again have blocks in the pass's launches, with short bands.
        bw.grids_r = bw.grids;
        for (BGrid &q : bw.grids_r) q.th = kRestartTh;
        HIPCHK(hipMemcpyAsync(bw.d_grids_r, bw.grids_r.data(), n * sizeof(BGrid),
                              hipMemcpyHostToDevice, s));
        part_tables(ctx, bw, again_of.data(), true);
        batch_pass(ctx, bw, 0, true);
        collect(false);
    }
    HIPCHK(hipEventRecord(ctx->ev[3], s));
    // Every grid's answer into the caller's h.
    kb_copy_out<<<bw.cells_l.nblk, kBlock, 0, s>>>(bw.d_grids, bw.cells_l.blk);
    LAUNCH_CHECK();
    HIPCHK(hipEventRecord(ctx->ev[4], s));
    HIPCHK(hipStreamSynchronize(s));
    dtd_batch_stats out{};
    out.grids = n;
    out.cells = bw.cells;
    out.restarts = restarts;
    for (size_t g = 0; g < n; g++) {
        if (iterations) iterations[idx[g]] = its[g];
        if (rels) rels[idx[g]] = rel[g];
        out.iterations_max = std::max(out.iterations_max, its[g]);
        out.iterations_sum += its[g];
        out.rel_residual_max = std::max(out.rel_residual_max, rel[g]);
    }
    out.ms_setup = elapsed(ctx->ev[0], ctx->ev[1]);
    out.ms_guess = elapsed(ctx->ev[1], ctx->ev[2]);
    out.ms_cg = elapsed(ctx->ev[2], ctx->ev[3]);
    out.ms_total = elapsed(ctx->ev[0], ctx->ev[4]);
    if (stats) *stats = out;
}

// dtd_tune_batch: the band-height policy of one shape list, by timing whole batched solves (the
// built-in policy first, then one class at a time over its candidates, keeping the best so far).
static void tune_batch(dtd_ctx *ctx, size_t count, const dtd_grid *gs, double rel_tol, int reps,
                       double *ms_rule, double *ms_tuned) {
    std::vector<dtd_grid> live;
    for (size_t k = 0; k < count; k++)
        if (gs[k].w * gs[k].h) live.push_back(gs[k]);
    if (live.empty()) return;
    const auto key = batch_key(ctx, live.size(), live.data());
    // Each timed solve starts from the grids' h0 (the solve writes its answers over them), and the
    // arrays get h0 back at the end.
    size_t total = 0;
    for (const dtd_grid &g : live) total += g.w * g.h;
    double *save = nullptr;
    HIPCHK(hipMalloc(&save, total * sizeof(double)));
    auto copy_all = [&](bool out) {
        size_t o = 0;
        for (const dtd_grid &g : live) {
            const size_t b = g.w * g.h * sizeof(double);
            HIPCHK(out ? hipMemcpyAsync(save + o, g.h_dev, b, hipMemcpyDeviceToDevice, ctx->stream)
                       : hipMemcpyAsync(g.h_dev, save + o, b, hipMemcpyDeviceToDevice, ctx->stream));
            o += g.w * g.h;
        }
    };
    ctx->batch_tuning = true;
    auto time_policy = [&](const BatchPolicy &p) {
        ctx->batch_policy = p;
        std::vector<double> t;
        for (int k = 0; k <= reps; k++) {
            copy_all(false);
            dtd_batch_stats st{};
            solve_batch(ctx, count, gs, rel_tol, nullptr, nullptr, &st);
            if (k) t.push_back(st.ms_total);
        }
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    };
    try {
        copy_all(true);
        BatchPolicy best{};
        const double t_rule = time_policy(best);
        double t_best = t_rule;
        const int tall[] = {8, 12, 24, 32}, mid[] = {4, 12, 16}, small[] = {8}, f64[] = {8, 12, 24, 32};
        const int fpre[] = {8, 24, 32, 48}, fpost[] = {8, 24, 32};
        auto sweep = [&](int BatchPolicy::*field, const int *cand, int nc) {
            for (int c = 0; c < nc; c++) {
                BatchPolicy p = best;
                p.*field = cand[c];
                const double t = time_policy(p);
                if (t < t_best) t_best = t, best = p;
            }
        };
        sweep(&BatchPolicy::fine_pre, fpre, 4);
        sweep(&BatchPolicy::fine_post, fpost, 3);
        sweep(&BatchPolicy::tall, tall, 4);
        sweep(&BatchPolicy::f64, f64, 4);
        sweep(&BatchPolicy::mid, mid, 3);
        sweep(&BatchPolicy::small, small, 1);
        ctx->batch_plans[key] = best;
        ctx->batch_policy = best;
        copy_all(false);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ms_rule) *ms_rule = t_rule;
        if (ms_tuned) *ms_tuned = t_best;
    } catch (...) {
        ctx->batch_tuning = false;
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(save);
        throw;
    }
    ctx->batch_tuning = false;
    (void)hipFree(save);
}

}  // namespace dtd
