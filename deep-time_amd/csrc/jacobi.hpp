// The Jacobi-preconditioned solve (`step_in`, D:262-283; `cg`, D:927-1026) on the device.
// Included once, by dtdiff.hip.
//
// Reference files (relative to the deep-time repository root):
//   D  crates/deeptime-core/src/surface/diffusion.rs
//
// The reference takes this solve's sums in a fixed order so that a run reproduces bit for bit at
// any thread count: per row in ascending column order from -0.0, then the row sums in row order.
// The device keeps that order.  Each row's sum is one chain of additions by one thread, over
// tiles of the row staged through LDS by the whole block (so the loads stay coalesced), and the
// row sums are added by one thread; every per-cell value is the reference's expression.  The
// iterations and every iterate are therefore the reference's.  The one place the order differs is
// the put-back (D:1000-1023), whose two sums run over all n cells in one chain in the reference:
// here they are block sums added in block order, so the put-back's step differs from the
// reference's by rounding only.
//
// Rows with nothing to diffuse (`Live`, D:285-367) are skipped by the reference only when doing so
// gives the bits of visiting them (their faces are +0 and their elevations normal, so the passes
// carry +0 through them and leave them as they are), and when a visited value ends up not finite
// it solves again on every row.  The device's loop skips the same strips of 8 rows (the first
// guess, the first residual and the passes after the loop visit every row, which leaves the
// skipped rows as they are): a skipped strip's row sums are the +0 that visiting it gives, its
// directions stay +0, its x and r are not touched.  A visited value that is not finite makes it
// solve again on every row, as the reference does.
#pragma once

#include "context.hpp"
#include "kernels.hpp"

namespace dtd {

constexpr int kRowTile = 8;     // rows per block of the row sums (16 and 4 measured slower)
static_assert(kRowTile == 8, "a block of the row sums is one strip of `Live` (STRIP_ROWS, D:502)");
constexpr int kColTile = 256;   // columns per tile (one per thread of the block)
constexpr double kCgTolM = 1.0e-6;  // the Jacobi solve's RMS tolerance, metres (D:495)

// The Jacobi loop's scalars (D:941-998).
struct JState {
    double rz, alpha, beta, tol2, lost, weight;
    int have_beta, pending, done, its, max_iter, pad;
};

// The strips of kRowTile rows the loop visits (`Live`, D:285-367): their indices in ascending
// order (list) and how many (*count), of ns strips of an h-row grid.
struct JLive {
    const int *list;
    const int *count;
    int ns, h;
};

// diag (D:585-600), stored once a solve: every pass divides by it.
__global__ __launch_bounds__(kBlock) void kj_diag(Grid g, const double2 *__restrict__ es,
                                                  double *__restrict__ diag) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const int j = c / g.w, i = c - j * g.w;
    diag[c] = diag_es(g, es, c, i, j);
}

// r = b - A x (pass_init, D:747-780).
__global__ __launch_bounds__(kBlock) void kj_resid(Grid g, const double2 *__restrict__ es,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ r) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const int j = c / g.w, i = c - j * g.w;
    r[c] = b[c] - apply_es(g, es, c, i, j, [&](int k) { return x[k]; });
}

// Per row, in ascending column order from -0.0 (`row_sums`/`row_dots`, D:666-745):
//   kRowsRZ:  s0[j] = sum r r,  s1[j] = sum r (r / g)    (a = r, b = diag)
//   kRowsRZU: r -= alpha A p (pass_update, D:883-905) written back, then the sums of kRowsRZ
//             on the new r                  (a = r, b = diag, c = A p; rout may be a)
// The block takes kRowTile rows; for each tile of kColTile columns its threads form the products
// (one column each) into LDS, then thread k adds row k's in column order.  A row's sum is one
// dependent chain of w adds, so the rows are spread thin (kRowTile rows a block, several blocks
// a CU): one block's chains then run while another's loads are in flight.
enum { kRowsRZ = 0, kRowsRZU = 1 };

template <int MODE>
__global__ __launch_bounds__(kColTile) void kj_rows(int w, int h, const double *a,
                                                    const double *__restrict__ b,
                                                    const double *__restrict__ c, double *rout,
                                                    double *__restrict__ s0,
                                                    double *__restrict__ s1,
                                                    const JState *__restrict__ st,
                                                    JLive lv) {
    __shared__ double t0[kRowTile][kColTile + 1];
    __shared__ double t1[kRowTile][kColTile + 1];
    if (st->done || (int)blockIdx.x >= *lv.count) return;
    const double alpha = MODE == kRowsRZU ? st->alpha : 0.0;
    // Block b takes the b-th visited strip (a skipped strip's r stays +0).
    const int j0 = lv.list[blockIdx.x] * kRowTile;
    const int tid = threadIdx.x;
    // The next tile's values are loaded into registers while the current tile is being added.
    double qa[kRowTile], qb[kRowTile], qc[MODE == kRowsRZU ? kRowTile : 1];
#define DTD_FETCH_ROWS(i0_)                                                                    \
    do {                                                                                       \
        const int i_ = (i0_) + tid;                                                            \
        _Pragma("unroll") for (int k = 0; k < kRowTile; k++) {                                 \
            const int j_ = j0 + k;                                                             \
            if (j_ < h && i_ < w) {                                                            \
                qa[k] = a[(size_t)j_ * w + i_];                                                \
                qb[k] = b[(size_t)j_ * w + i_];                                                \
                if (MODE == kRowsRZU) qc[k] = c[(size_t)j_ * w + i_];                          \
            }                                                                                  \
        }                                                                                      \
    } while (0)
    double acc0 = -0.0, acc1 = -0.0;
    DTD_FETCH_ROWS(0);
    for (int i0 = 0; i0 < w; i0 += kColTile) {
#pragma unroll
        for (int k = 0; k < kRowTile; k++) {
            if (MODE == kRowsRZU) {
                const int j_ = j0 + k, i_ = i0 + tid;
                const double rn = qa[k] - alpha * qc[k];
                if (j_ < h && i_ < w) rout[(size_t)j_ * w + i_] = rn;
                t0[k][tid] = rn * rn;
                t1[k][tid] = rn * (rn / qb[k]);
            } else {
                t0[k][tid] = qa[k] * qa[k];
                t1[k][tid] = qa[k] * (qa[k] / qb[k]);
            }
        }
        __syncthreads();
        if (i0 + kColTile < w) DTD_FETCH_ROWS(i0 + kColTile);
        const int cols = min(kColTile, w - i0);
        if (tid < kRowTile && j0 + tid < h) {
#pragma unroll 8
            for (int q = 0; q < cols; q++) {
                acc0 += t0[tid][q];
                acc1 += t1[tid][q];
            }
        }
        __syncthreads();
    }
#undef DTD_FETCH_ROWS
    if (tid < kRowTile && j0 + tid < h) {
        s0[j0 + tid] = acc0;
        s1[j0 + tid] = acc1;
    }
}

// A p (`Operator::apply_row`, D:609-649) into ap and, per row in ascending column order from -0.0,
// s0[j] = sum p Ap (`row_dots`, D:666-745), in one pass: the block takes kApplyRows rows; for each
// tile of kColTile columns its threads form A p for one column of each row, store it and put
// p Ap into LDS; thread k then adds row k's in column order.  The same values and sums as
// kj_apply followed by a row-sum pass over p and A p, without reading them back.
constexpr int kApplyRows = 4;  // rows per block of kj_apply_rows (8 measured slower, 2 the same)

// One cell's operands of A p (`apply_es`): its faces and the five values of p, with flags for
// the neighbours that exist; gathered a tile ahead, then combined in apply_es's order.
struct ApplyIn {
    double e, s, wv, nv, pc, pe, pw, ps, pn;
    bool he, hw;
};

__device__ __forceinline__ ApplyIn apply_gather(const Grid &g, const double2 *__restrict__ es,
                                                const double *__restrict__ p, int c, int i,
                                                int j) {
    ApplyIn q;
    const int c0 = j * g.w;
    const double2 me = es[c];
    q.e = me.x, q.s = me.y;
    q.wv = i > 0 ? es[c - 1].x : (g.wrap ? es[c0 + g.w - 1].x : 0.0);
    q.nv = j > 0 ? es[c - g.w].y : 0.0;
    q.pc = p[c];
    q.he = i + 1 < g.w || g.wrap;
    q.pe = i + 1 < g.w ? p[c + 1] : (g.wrap ? p[c0] : 0.0);
    q.hw = i > 0 || g.wrap;
    q.pw = i > 0 ? p[c - 1] : (g.wrap ? p[c0 + g.w - 1] : 0.0);
    q.ps = j + 1 < g.h ? p[c + g.w] : 0.0;
    q.pn = j > 0 ? p[c - g.w] : 0.0;
    return q;
}

__device__ __forceinline__ double apply_combine(const ApplyIn &q, int j, int h) {
    const double d = 1.0 + q.e + q.wv + q.s + q.nv;
    double o = d * q.pc;
    if (q.he) o -= q.e * q.pe;
    if (q.hw) o -= q.wv * q.pw;
    if (j + 1 < h) o -= q.s * q.ps;
    if (j > 0) o -= q.nv * q.pn;
    return o;
}

template <int RT>
__global__ __launch_bounds__(kColTile) void kj_apply_rows(Grid g, const double2 *__restrict__ es,
                                                          const double *__restrict__ p,
                                                          double *__restrict__ ap,
                                                          double *__restrict__ s0,
                                                          const JState *__restrict__ st,
                                                          JLive lv) {
    static_assert(kRowTile % RT == 0, "a block's rows lie in one strip");
    constexpr int per = kRowTile / RT;
    __shared__ double t0[RT][kColTile + 1];
    if (st->done || (int)blockIdx.x >= *lv.count * per) return;
    // Blocks per + k take the b-th visited strip's rows (a skipped strip's p and A p are +0).
    const int j0 = lv.list[blockIdx.x / per] * kRowTile + (blockIdx.x % per) * RT,
              tid = threadIdx.x;
    // The next tile's operands are gathered into registers while the current tile is being added.
    ApplyIn q[RT];
    auto gather = [&](int i0) {
        const int i = i0 + tid;
#pragma unroll
        for (int k = 0; k < RT; k++) {
            const int j = j0 + k;
            if (j < g.h && i < g.w) q[k] = apply_gather(g, es, p, j * g.w + i, i, j);
        }
    };
    double acc = -0.0;
    gather(0);
    for (int i0 = 0; i0 < g.w; i0 += kColTile) {
        const int i = i0 + tid;
#pragma unroll
        for (int k = 0; k < RT; k++) {
            const int j = j0 + k;
            if (j < g.h && i < g.w) {
                const double o = apply_combine(q[k], j, g.h);
                ap[j * g.w + i] = o;
                t0[k][tid] = q[k].pc * o;
            }
        }
        __syncthreads();
        if (i0 + kColTile < g.w) gather(i0 + kColTile);
        const int cols = min(kColTile, g.w - i0);
        if (tid < RT && j0 + tid < g.h) {
#pragma unroll 8
            for (int c = 0; c < cols; c++) acc += t0[tid][c];
        }
        __syncthreads();
    }
    if (tid < RT && j0 + tid < g.h) s0[j0 + tid] = acc;
}

// The scalar steps (D:940-998), one thread, the row sums added in row order from -0.0.
enum { kJInit = 0, kJPap = 1, kJUpdate = 2 };

// The row sums are staged through LDS by the whole block, then added by thread 0 in row order
// from -0.0.
constexpr int kStepBlock = 1024, kStepChunk = 4096;

// Only the visited strips' rows are added: a skipped row's sum is +0, and adding +0 to a running
// sum changes nothing but a -0 (all terms so far -0) into +0, so the rows in order give the
// visited rows' total in order, except that a -0 total becomes +0 when a row was skipped.  A row
// past h in the last strip reads -0, which adds nothing.
template <int STEP>
__device__ inline void kj_step_body(const double *__restrict__ s0, const double *__restrict__ s1,
                                    int h, JState *st, double *l0, double *l1, JLive lv) {
    if (st->done) return;
    double v0 = -0.0, v1 = -0.0;
    const int nl = *lv.count, rows = nl * kRowTile;
    for (int c0 = 0; c0 < rows; c0 += kStepChunk) {
        const int cnt = min(kStepChunk, rows - c0);
        for (int k = threadIdx.x; k < cnt; k += kStepBlock) {
            const int q = c0 + k, j = lv.list[q / kRowTile] * kRowTile + q % kRowTile;
            l0[k] = j < h ? s0[j] : -0.0;
            if (STEP != kJPap) l1[k] = j < h ? s1[j] : -0.0;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll 8
            for (int k = 0; k < cnt; k++) {
                v0 += l0[k];
                if (STEP != kJPap) v1 += l1[k];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (nl < lv.ns) v0 += 0.0, v1 += 0.0;  // the skipped rows' +0
    if (STEP == kJInit) {
        // rr0 > tol2 starts the loop with rz = sum r z (D:947-953).
        if (v0 > st->tol2)
            st->rz = v1;
        else
            st->done = 1;
    } else if (STEP == kJPap) {
        // iterations += 1; the direction pass took the pending update; the p.A p break
        // (D:955-977).
        st->its += 1;
        st->pending = 0;
        if (v0 <= 0.0) {
            st->done = 2;
            return;
        }
        st->alpha = st->rz / v0;
    } else {
        // After r -= alpha A p: the pending update, the r.r stop, beta (D:978-992); the cap.
        st->pending = 1;
        if (v0 <= st->tol2) {
            st->done = 1;
            return;
        }
        st->beta = v1 / st->rz;
        st->have_beta = 1;
        st->rz = v1;
        if (st->its >= st->max_iter) st->done = 4;
    }
}

// With flag given (host memory the device can write), the step leaves the stop flag there too.
template <int STEP>
__global__ __launch_bounds__(kStepBlock) void kj_step(const double *__restrict__ s0,
                                                      const double *__restrict__ s1, int h,
                                                      JState *st, JLive lv, int *flag = nullptr) {
    __shared__ double l0[kStepChunk], l1[STEP == kJPap ? 1 : kStepChunk];
    kj_step_body<STEP>(s0, s1, h, st, l0, l1, lv);
    if (flag && threadIdx.x == 0) *reinterpret_cast<volatile int *>(flag) = st->done;
}

// The direction pass (D:782-880) cell by cell: the pending x += alpha p_old, then
// p = r / g (+ beta p_old).
__global__ __launch_bounds__(kBlock) void kj_dir(int n, const double *__restrict__ r,
                                                 const double *__restrict__ diag,
                                                 const double *__restrict__ pold,
                                                 double *__restrict__ pnew,
                                                 double *__restrict__ x,
                                                 const JState *__restrict__ st, JLive lv,
                                                 int strip_cells) {
    if (st->done) return;
    // Chunks of kBlock cells of the visited strips, in turn over the grid's blocks.
    const int cps = (strip_cells + kBlock - 1) / kBlock;
    const int chunks = *lv.count * cps;
    const bool have_beta = st->have_beta, pending = st->pending;
    const double alpha = st->alpha, beta = st->beta;
    for (int q = blockIdx.x; q < chunks; q += gridDim.x) {
        const int li = q / cps, off = (q - li * cps) * kBlock + (int)threadIdx.x;
        const int c = lv.list[li] * strip_cells + off;
        if (off >= strip_cells || c >= n) continue;
        const double po = (have_beta || pending) ? pold[c] : 0.0;
        if (pending) x[c] += alpha * po;
        pnew[c] = have_beta ? r[c] / diag[c] + beta * po : r[c] / diag[c];
    }
}

// A p (`Operator::apply_row`, D:609-649).
__global__ __launch_bounds__(kBlock) void kj_apply(Grid g, const double2 *__restrict__ es,
                                                   const double *__restrict__ p,
                                                   double *__restrict__ ap,
                                                   const JState *__restrict__ st) {
    if (st && st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const int j = c / g.w, i = c - j * g.w;
    ap[c] = apply_es(g, es, c, i, j, [&](int k) { return p[k]; });
}

// After the loop: the pending x += alpha p (pass_x, D:907-919) with p the last direction
// (iteration k wrote p0 when k is odd, p1 when even).
__global__ __launch_bounds__(kBlock) void kj_last(int n, const double *__restrict__ p0,
                                                  const double *__restrict__ p1,
                                                  double *__restrict__ x,
                                                  const JState *__restrict__ st) {
    if (!st->pending) return;
    const double *p = (st->its & 1) ? p0 : p1;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) x[c] += st->alpha * p[c];
}

// The put-back's sums per block (sum (b - A x), sum (g - 1)), A x already in ax.
__global__ __launch_bounds__(kBlock) void kj_putback_sums(int n, const double *__restrict__ b,
                                                          const double *__restrict__ ax,
                                                          const double *__restrict__ diag,
                                                          double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double l = 0.0, wg = 0.0;
    if (c < n) {
        l = b[c] - ax[c];
        wg = diag[c] - 1.0;
    }
    block_pair(make_double2(l, wg), parts, blockIdx.x, lds);
}

__global__ void kj_putback_step(const double *__restrict__ parts, int nparts, JState *st) {
    __shared__ double lds[kReduceBlock / 64];
    const double l = reduce_parts(parts, nparts, 2, 0, lds);
    const double wg = reduce_parts(parts, nparts, 2, 1, lds);
    if (threadIdx.x == 0) st->lost = l, st->weight = wg;
}

// x += (lost / weight) (g - 1) when weight > 0 and lost != 0 (D:1013-1023).
__global__ __launch_bounds__(kBlock) void kj_putback(int n, const double *__restrict__ diag,
                                                     double *__restrict__ x,
                                                     const JState *__restrict__ st) {
    const double lost = st->lost, weight = st->weight;
    if (!(weight > 0.0 && lost != 0.0)) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) x[c] += (lost / weight) * (diag[c] - 1.0);
}

__global__ void kj_init(JState *st, double tol2, int max_iter) {
    if (threadIdx.x != 0) return;
    JState s{};
    s.tol2 = tol2;
    s.max_iter = max_iter;
    *st = s;
}

// The first guess with the reference's recurrences (bit for bit), the rows on the transposed
// grid (a wave's rows side by side in memory): kd^T in kdt, b^T in xt (solved in place), cbuf
// and zbuf scratch, all w x h.
__global__ __launch_bounds__(kLineBlock) void kj_lod_rows_t(Grid g, int wrap_x, const double *kdt,
                                                        double *xt, double *cbuf, double *zbuf,
                                                        double dt, double s2) {
    const int j = blockIdx.x * kLineBlock + threadIdx.x;
    const int h = g.h;
    if (j < h)
        lod_line_x(Line<const double>{kdt + j, h}, Line<double>{xt + j, h},
                   Line<double>{cbuf + j, h}, Line<double>{zbuf + j, h}, g.w, wrap_x, dt, s2);
}

}  // namespace dtd

namespace dtd {

// ---- `Live` (D:285-367): the strips of 8 rows the loop visits -------------------------------

__device__ inline bool is_normal64(double v) {
    const double a = fabs(v);
    return a >= 2.2250738585072014e-308 && a <= 1.7976931348623157e308;
}

// One block a row: zero[j] = row j's diffusivity is +0 throughout (every bit clear; not -0 or
// NaN); flags[0] |= some row is; flags[1] |= a zero row has an elevation whose half is not normal
// (D:306-310, 326-334).
__global__ __launch_bounds__(kBlock) void kj_live_rows(int w, const double *__restrict__ kd,
                                                       const double *__restrict__ b,
                                                       unsigned char *__restrict__ zero,
                                                       int *flags) {
    const size_t c0 = (size_t)blockIdx.x * w;
    int z = 1;
    for (int i = threadIdx.x; i < w; i += kBlock) z &= __double_as_longlong(kd[c0 + i]) == 0;
    z = __syncthreads_and(z);
    int ok = 1;
    if (z) {
        for (int i = threadIdx.x; i < w; i += kBlock) ok &= is_normal64(0.5 * b[c0 + i]);
        ok = __syncthreads_and(ok);
    }
    if (threadIdx.x == 0) {
        zero[blockIdx.x] = (unsigned char)z;
        if (z) atomicOr(flags, 1);
        if (z && !ok) atomicOr(flags + 1, 1);
    }
}

// One block: live[s] = strip s holds a row that is not skipped (a row is skipped when it and its
// neighbours, where it has them, are zero rows; D:312-316); every strip when no row is a zero
// row, a zero row's elevations fail the normal test, or dt and dx^2 are not finite and positive
// (all_rows).  flags[2] = the visited strips' rows; list = the visited strips in order,
// flags[4] = how many.
__global__ __launch_bounds__(1024) void kj_live_strips(int h, const unsigned char *__restrict__ zero,
                                                       int *flags, int all_rows,
                                                       unsigned char *__restrict__ live,
                                                       int *__restrict__ list) {
    __shared__ int rows;
    if (threadIdx.x == 0) rows = 0;
    __syncthreads();
    const bool all = all_rows || !flags[0] || flags[1];
    const int ns = (h + kRowTile - 1) / kRowTile;
    int mine = 0;
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        const int a = s * kRowTile, e = min(a + kRowTile, h);
        bool v = all;
        for (int j = a; j < e && !v; j++) {
            const bool skip = zero[j] && (j == 0 || zero[j - 1]) && (j + 1 == h || zero[j + 1]);
            v = !skip;
        }
        live[s] = (unsigned char)v;
        if (v) mine += e - a;
    }
    atomicAdd(&rows, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        flags[2] = rows;
        int k = 0;
        for (int q = 0; q < ns; q++)
            if (live[q]) list[k++] = q;
        flags[4] = k;
    }
}

// The directions' +0 in the skipped strips (the loop never writes them there; the visited rows
// beside them read them across their +0 faces).
__global__ __launch_bounds__(kBlock) void kj_zero_dead(int n, int strip_cells,
                                                       const unsigned char *__restrict__ live,
                                                       double *__restrict__ p0,
                                                       double *__restrict__ p1) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n || live[c / strip_cells]) return;
    p0[c] = 0.0, p1[c] = 0.0;
}

// flags[3] |= a visited cell of x is not finite (D:272-279).
__global__ __launch_bounds__(kBlock) void kj_live_finite(int n, int strip_cells,
                                                         const unsigned char *__restrict__ live,
                                                         const double *__restrict__ x, int *flags) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int bad = c < n && live[c / strip_cells] && !isfinite(x[c]);
    if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(flags + 3, 1);
}

// Device buffers of the Jacobi solve beside the workspace's vectors (kept with the shape).
static void jacobi_fit(Workspace &ws) {
    if (ws.jdiag) return;
    ws.jdiag = dalloc<double>(ws.n);
    ws.jrows = dalloc<double>(2 * (size_t)ws.h);
    ws.jst = dalloc<JState>(1);
    ws.jzero = dalloc<unsigned char>((size_t)ws.h);
    ws.jlive = dalloc<unsigned char>((size_t)(ws.h + kRowTile - 1) / kRowTile);
    ws.jflags = dalloc<int>(5);
    ws.jlist = dalloc<int>((size_t)(ws.h + kRowTile - 1) / kRowTile);
}

// The direction pass's grid: its blocks take the visited strips' chunks of cells in turn.
constexpr int kDirBlocks = 16384;

static JLive jlive_of(const Workspace &ws) {
    return JLive{ws.jlist, ws.jflags + 4, (ws.h + kRowTile - 1) / kRowTile, ws.h};
}

// One iteration (D:955-992): parity 0 reads the direction from p1 and writes p0, parity 1 the
// other way.
static void jacobi_iteration(dtd_ctx *ctx, int parity) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    JState *st = (JState *)ws.jst;
    const Grid g = grid_of(ws);
    const int nb = nblocks(ws.n), nr = (ws.h + kRowTile - 1) / kRowTile;
    double *pold = parity == 0 ? ws.p1 : ws.p0, *pnew = parity == 0 ? ws.p0 : ws.p1;
    double *s0 = ws.jrows, *s1 = ws.jrows + ws.h;
    const JLive lv = jlive_of(ws);
    kj_dir<<<std::min(nb, kDirBlocks), kBlock, 0, s>>>((int)ws.n, ws.r, ws.jdiag, pold, pnew, ws.x,
                                                       st, lv, kRowTile * ws.w);
    LAUNCH_CHECK();
    kj_apply_rows<kApplyRows><<<nr * (kRowTile / kApplyRows), kColTile, 0, s>>>(
        g, ws.es, pnew, ws.ap, s0, st, lv);
    LAUNCH_CHECK();
    kj_step<kJPap><<<1, kStepBlock, 0, s>>>(s0, s1, ws.h, st, lv);
    LAUNCH_CHECK();
    kj_rows<kRowsRZU><<<nr, kColTile, 0, s>>>(ws.w, ws.h, ws.r, ws.jdiag, ws.ap, ws.r, s0, s1,
                                              st, lv);
    LAUNCH_CHECK();
    kj_step<kJUpdate><<<1, kStepBlock, 0, s>>>(s0, s1, ws.h, st, lv, ctx->dev_flag + parity);
    LAUNCH_CHECK();
}

// `step_in` (D:262-283) of b (device; only read) into ws.x: the operator, the splitting's first
// guess by the reference's recurrences, the loop to CG_TOL_M, the last update and the put-back.
static size_t jacobi_solve(dtd_ctx *ctx, const double *b, const double *kd, double dx, double dt,
                           int wrap_x, dtd_stats *stats, bool all_rows = false) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    jacobi_fit(ws);
    JState *st = (JState *)ws.jst;
    const Grid g = grid_of(ws);
    const int nb = nblocks(ws.n), nr = (ws.h + kRowTile - 1) / kRowTile;
    const double s2 = dx * dx;
    double *s0 = ws.jrows, *s1 = ws.jrows + ws.h;
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    k_operator<<<nb, kBlock, 0, s>>>(g, kd, dt, s2, ws.es, nullptr, nullptr);
    LAUNCH_CHECK();
    kj_diag<<<nb, kBlock, 0, s>>>(g, ws.es, ws.jdiag);
    LAUNCH_CHECK();
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    {
        // kd^T into r, b^T into ap; the rows; ap^T back into x; the columns in place.
        const dim3 tg((ws.w + 31) / 32, (ws.h + 31) / 32), tgt((ws.h + 31) / 32, (ws.w + 31) / 32);
        k_transpose<<<tg, 256, 0, s>>>(kd, ws.r, ws.h, ws.w);
        LAUNCH_CHECK();
        k_transpose<<<tg, 256, 0, s>>>(b, ws.ap, ws.h, ws.w);
        LAUNCH_CHECK();
        kj_lod_rows_t<<<(ws.h + kLineBlock - 1) / kLineBlock, kLineBlock, 0, s>>>(g, wrap_x, ws.r, ws.ap, ws.p0, ws.p1, dt,
                                                        s2);
        LAUNCH_CHECK();
        k_transpose<<<tgt, 256, 0, s>>>(ws.ap, ws.x, ws.w, ws.h);
        LAUNCH_CHECK();
        k_lod_cols<<<(ws.w + kLineBlock - 1) / kLineBlock, kLineBlock, 0, s>>>(g, kd, dt, s2, ws.x,
                                                                              ws.p0);
        LAUNCH_CHECK();
    }
    HIPCHK(hipEventRecord(ctx->ev[2], s));
    // The strips the loop visits (`Live::of`, D:298-357).
    HIPCHK(hipMemsetAsync(ws.jflags, 0, 5 * sizeof(int), s));
    const bool every = all_rows || !ctx->use_live ||
                       !(std::isfinite(dt) && dt > 0.0 && std::isfinite(s2) && s2 > 0.0);
    kj_live_rows<<<ws.h, kBlock, 0, s>>>(ws.w, kd, b, ws.jzero, ws.jflags);
    LAUNCH_CHECK();
    kj_live_strips<<<1, 1024, 0, s>>>(ws.h, ws.jzero, ws.jflags, every ? 1 : 0, ws.jlive,
                                      ws.jlist);
    LAUNCH_CHECK();
    kj_zero_dead<<<nb, kBlock, 0, s>>>((int)ws.n, kRowTile * ws.w, ws.jlive, ws.p0, ws.p1);
    LAUNCH_CHECK();
    kj_init<<<1, 64, 0, s>>>(st, kCgTolM * kCgTolM * (double)ws.n, (int)ctx->opts.max_iter);
    LAUNCH_CHECK();
    kj_resid<<<nb, kBlock, 0, s>>>(g, ws.es, b, ws.x, ws.r);
    LAUNCH_CHECK();
    kj_rows<kRowsRZ><<<nr, kColTile, 0, s>>>(ws.w, ws.h, ws.r, ws.jdiag, nullptr, nullptr, s0,
                                             s1, st, jlive_of(ws));
    LAUNCH_CHECK();
    kj_step<kJInit><<<1, kStepBlock, 0, s>>>(s0, s1, ws.h, st, jlive_of(ws));
    LAUNCH_CHECK();
    if (ctx->use_graph)
        for (int k = 0; k < 2; k++) {
            if (ws.jgraph[k]) continue;
            hipGraph_t gr = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                jacobi_iteration(ctx, k);
            } catch (...) {
                (void)hipStreamEndCapture(s, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            HIPCHK(hipStreamEndCapture(s, &gr));
            hipError_t e = hipGraphInstantiate(&ws.jgraph[k], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            HIPCHK(e);
        }
    // The loop, reading the stop flag one iteration behind (the flag before the loop counts as
    // iteration 0's); the state's cap stops it by max_iter.
    HIPCHK(hipMemcpyAsync(&ctx->host_flag[1], &st->done, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(ctx->flag_ev[1], s));
    const int max_iter = (int)ctx->opts.max_iter;
    for (int k = 0; k <= max_iter + 1; k++) {
        const int par = k & 1;
        if (ctx->use_graph)
            HIPCHK(hipGraphLaunch(ws.jgraph[par], s));
        else
            jacobi_iteration(ctx, par);
        HIPCHK(hipEventRecord(ctx->flag_ev[par], s));
        HIPCHK(hipEventSynchronize(ctx->flag_ev[par ^ 1]));
        if (ctx->host_flag[par ^ 1]) break;
    }
    kj_last<<<nb, kBlock, 0, s>>>((int)ws.n, ws.p0, ws.p1, ws.x, st);
    LAUNCH_CHECK();
    kj_apply<<<nb, kBlock, 0, s>>>(g, ws.es, ws.x, ws.ap, nullptr);
    LAUNCH_CHECK();
    kj_putback_sums<<<nb, kBlock, 0, s>>>((int)ws.n, b, ws.ap, ws.jdiag, ws.parts);
    LAUNCH_CHECK();
    kj_putback_step<<<1, kReduceBlock, 0, s>>>(ws.parts, nb, st);
    LAUNCH_CHECK();
    kj_putback<<<nb, kBlock, 0, s>>>((int)ws.n, ws.jdiag, ws.x, st);
    LAUNCH_CHECK();
    kj_live_finite<<<nb, kBlock, 0, s>>>((int)ws.n, kRowTile * ws.w, ws.jlive, ws.x, ws.jflags);
    LAUNCH_CHECK();
    HIPCHK(hipEventRecord(ctx->ev[3], s));
    JState hs{};
    int hflags[5] = {0, 0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(ctx->host_st, st, sizeof(JState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hflags, ws.jflags, sizeof(hflags), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memcpy(&hs, ctx->host_st, sizeof(JState));
    if (hflags[2] < ws.h && hflags[3]) {
        // A value that is not finite would reach the skipped rows in the full loop (0 * inf):
        // solve on every row instead (D:272-279); its count is the step's, as there.
        return jacobi_solve(ctx, b, kd, dx, dt, wrap_x, stats, true);
    }
    if (stats) {
        stats->iterations += (size_t)hs.its;
        stats->ms_setup += elapsed(ctx->ev[0], ctx->ev[1]);
        stats->ms_guess += elapsed(ctx->ev[1], ctx->ev[2]);
        stats->ms_cg += elapsed(ctx->ev[2], ctx->ev[3]);
        stats->ms_total += elapsed(ctx->ev[0], ctx->ev[3]);
        stats->rows_visited = (size_t)hflags[2];
    }
    return (size_t)hs.its;
}

}  // namespace dtd
