// The cube-grid diffusion solve (`surface::cube::diffusion::solve`) on the device: the operator on
// the cube's cells, the face-wise first guess, conjugate gradients preconditioned by the diagonal
// (`step_in`) or by the cube's V-cycle, and the residual's sum put back.  Included once, by
// dtdiff.hip.
//
// Reference files (relative to the deep-time repository root):
//   CD   crates/deeptime-core/src/surface/cube/diffusion.rs
//   CM   crates/deeptime-core/src/surface/cube/multigrid.rs
//   CUBE crates/deeptime-core/src/surface/cube.rs
//
// The grid comes from the caller as the reference's `CubeGrid` hands it to the solve: each cell's
// eight neighbour slots (D8 order in the cell's face frame, NO_CELL across a cube corner), the
// couplings of its Laplacian and the cells' areas with their mean.  dtd_cube_new uploads them and
// builds the V-cycle's levels (their neighbour lists and the lists each coarse value sums) once;
// a solve fills in the values.
//
// Every value is the reference's expression in its order, with FP contraction off, so each pass
// gives the reference's bits.  The sums are the reference's too: `ordered_sum` and the CG passes
// add each chunk of 4096 cells in cell order and then the chunks in order (CUBE:48-58, CD:441-470),
// and here each chunk's terms are added one by one in order by one 16-lane row of a wave, and the
// chunk sums in order by one thread.  The level sums (a coarse value over its finer entries, a
// block's residuals) run over fixed lists in the reference's order.  So the iterations and the
// answer are the reference's.
#pragma once

#include <utility>

#include "context.hpp"
#include "kernels.hpp"

namespace dtd {

constexpr uint32_t kNoCell = 0xFFFFFFFFu;  // sphere/cube.rs:65
constexpr int kCubeChunk = 4096;           // CUBE:48
// The raster cycle's constants, which the cube's cycle takes (CM:24;
// surface/diffusion/multigrid.rs:45-56).
constexpr int kCubeCoarsest = 64, kCubeNu = 2, kCubeCoarseSweeps = 16;

// The loop's scalars on the device.
struct CubeState {
    double rz, pap, alpha, beta, tol2, lost, weight;
    int have_beta, done, its, max_iter;
};

// ---- the operator (CD:94-224) ------------------------------------------------------------------

// `couplings`, CD:94-119: a side pair's mean kd, a diagonal's least of its quad's four side means
// (the two side cells exist: only diagonals go missing), times dt / mean area and the coupling.
__global__ __launch_bounds__(kBlock) void kc_couplings(int n6, const uint32_t *__restrict__ slots,
                                                      const double *__restrict__ lap,
                                                      const double *__restrict__ kd, double scale,
                                                      double *__restrict__ k,
                                                      unsigned char *__restrict__ any) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n6) return;
    const double kc = kd[c];
    uint32_t sl[8];
#pragma unroll
    for (int q = 0; q < 8; q++) sl[q] = slots[(size_t)q * n6 + c];
    bool nz = false;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        double v = 0.0;
        const uint32_t m = sl[q];
        if (m != kNoCell) {
            double d;
            if (q % 2 == 0) {
                d = 0.5 * (kc + kd[m]);
            } else {
                const uint32_t a = sl[q - 1], b = sl[(q + 1) % 8];
                const double ka = kd[a], km = kd[m], kb = kd[b];
                d = fmin(fmin(fmin(0.5 * (kc + ka), 0.5 * (ka + km)), 0.5 * (km + kb)),
                         0.5 * (kb + kc));
            }
            v = scale * d * lap[(size_t)q * n6 + c];
        }
        k[(size_t)q * n6 + c] = v;
        nz |= v != 0.0;
    }
    any[c] = nz;
}

// The operator at each position i of the cells taking part (cells[i]; the identity when every cell
// does): its couplings (gathered when a subset takes part), mu = a / mean area, s = the couplings'
// sum in slot order from -0.0, diag = mu + s (CD:206-215); rhs = mu h and x = h (CD:268-279).
__global__ __launch_bounds__(kBlock) void kc_opinit(int l, int n6, const uint32_t *__restrict__ cells,
                                                   const double *__restrict__ kfull,
                                                   double *__restrict__ kpos,
                                                   const double *__restrict__ area, double mean,
                                                   const double *__restrict__ h,
                                                   double *__restrict__ mu, double *__restrict__ s,
                                                   double *__restrict__ diag,
                                                   double *__restrict__ rhs,
                                                   double *__restrict__ x) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= l) return;
    const uint32_t c = cells ? cells[i] : (uint32_t)i;
    double sum = -0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const double v = kfull[(size_t)q * n6 + c];
        if (kpos != kfull) kpos[(size_t)q * l + i] = v;
        sum += v;
    }
    const double m = area[c] / mean;
    mu[i] = m;
    s[i] = sum;
    diag[i] = m + sum;
    const double v = h[c];
    rhs[i] = m * v;
    x[i] = v;
}

// `Op::row`, CD:79-91.
// Cell i's eight slot neighbours.  With side > 0 the table is known to be the structured one (each
// face row-major, slot q of an inner cell at i + DI[q] + DJ[q] side, checked by cube_new), and an
// inner cell forms them; other cells read the slot-major table (kNoCell where there is none).
__device__ __forceinline__ void cube_nbrs(int i, int l, int side,
                                          const uint32_t *__restrict__ cols, uint32_t *j) {
    bool inner = false;
    if (side > 0) {
        const int x = i % side, y = (i / side) % side;
        inner = x > 0 && x < side - 1 && y > 0 && y < side - 1;
    }
    if (inner) {
        const uint32_t c = (uint32_t)i, n = (uint32_t)side;
        j[0] = c + 1, j[1] = c + 1 - n, j[2] = c - n, j[3] = c - 1 - n;
        j[4] = c - 1, j[5] = c - 1 + n, j[6] = c + n, j[7] = c + 1 + n;
    } else {
#pragma unroll
        for (int q = 0; q < 8; q++) j[q] = cols[(size_t)q * l + i];
    }
}

__device__ __forceinline__ double cube_row(int i, int l, int side, const double *__restrict__ diag,
                                           const double *__restrict__ k,
                                           const uint32_t *__restrict__ cols,
                                           const double *__restrict__ p) {
    double acc = diag[i] * p[i];
    uint32_t j[8];
    cube_nbrs(i, l, side, cols, j);
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (j[q] != kNoCell) acc -= k[(size_t)q * l + i] * p[j[q]];
    return acc;
}

__global__ __launch_bounds__(kBlock) void kc_apply(int l, int side, const double *__restrict__ diag,
                                                  const double *__restrict__ k,
                                                  const uint32_t *__restrict__ cols,
                                                  const double *__restrict__ p,
                                                  double *__restrict__ out,
                                                  const double *__restrict__ sub) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= l) return;
    const double a = cube_row(i, l, side, diag, k, cols, p);
    out[i] = sub ? sub[i] - a : a;  // sub: r = b - A x (CD:386-390)
}

// ---- the face-wise first guess (CD:316-378) ----------------------------------------------------

// The lines' tridiagonal systems, formed first by one thread a cell.  Row line of cell
// (face, j, i): line face n + j, position i; column line: line face n + i, position j.  The
// systems are stored line-interleaved, position t of line L at t 6n + L, so that a wave of
// consecutive lines reads each position's values in one contiguous run.  Each position gets the
// reference's lo, di, up and right-hand side: di = 1 + (K_a + K_b) / mu, lo = -K_a / mu (0 first),
// up = -K_b / mu (0 last), v = b plus, at a face's edge, K b_beyond / mu (CD:326-345, 350-368).
// b is in cell order, or (b_rows) in the row lines' interleaved order the row pass left it in.
__device__ __forceinline__ size_t lod_row_at(uint32_t c, int n) {
    const uint32_t nn = (uint32_t)n * n, f = c / nn, rest = c - f * nn, j = rest / n;
    return (size_t)(rest - j * n) * (6 * (size_t)n) + (size_t)f * n + j;
}

__global__ __launch_bounds__(kBlock) void kc_lod_coef(int n, int cols, int b_rows,
                                                     const double *__restrict__ k,
                                                     const double *__restrict__ mu,
                                                     const uint32_t *__restrict__ slots,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ lo,
                                                     double *__restrict__ di,
                                                     double *__restrict__ up,
                                                     double *__restrict__ v) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int nn = n * n;
    if (c >= 6 * nn) return;
    const int f = c / nn, rest = c - f * nn, j = rest / n, i = rest - j * n;
    // Row lines: slots E (0) and W (4) along i; column lines: N (2) and S (6) along j.
    const int sa = cols ? 2 : 4, sb = cols ? 6 : 0, t = cols ? j : i;
    const size_t at = (size_t)t * (6 * (size_t)n) + (size_t)f * n + (cols ? i : j);
    const size_t n6 = 6 * (size_t)nn;
    auto bv = [&](uint32_t m) { return b[b_rows ? lod_row_at(m, n) : (size_t)m]; };
    const double ka = k[sa * n6 + c], kb = k[sb * n6 + c], m = mu[c];
    di[at] = 1.0 + (kb + ka) / m;
    double rhs = bv(c);
    double l = 0.0, u = 0.0;
    if (t > 0)
        l = -ka / m;
    else
        rhs += ka * bv(slots[sa * n6 + c]) / m;
    if (t + 1 < n)
        u = -kb / m;
    else
        rhs += kb * bv(slots[sb * n6 + c]) / m;
    lo[at] = l, up[at] = u, v[at] = rhs;
}

// `thomas` (D:1029-1046) on each of nl line-interleaved lines of n values (position t of line L at
// t nl + L), one thread a line: v solved in place, cp the lines' scratch.  The recurrence is one
// long dependent chain per line and only nl / 64 waves run, so the loads are software-pipelined:
// the next kThomasP positions' values are read while the current ones are used.
constexpr int kThomasP = 16;

__global__ __launch_bounds__(64) void kc_thomas(int nl, int n, const double *__restrict__ lo,
                                               const double *__restrict__ di,
                                               const double *__restrict__ up,
                                               double *__restrict__ v, double *__restrict__ cp) {
    constexpr int P = kThomasP;
    const int line = blockIdx.x * 64 + threadIdx.x;
    if (line >= nl) return;
    const size_t stride = (size_t)nl;
    const double *lo_ = lo + line, *di_ = di + line, *up_ = up + line;
    double *v_ = v + line, *cp_ = cp + line;
    double beta = di_[0];
    double prev = v_[0] / beta;
    v_[0] = prev;
    {
        // Forward: positions t = 1 .. n - 1 use up[t - 1], lo[t], di[t], v[t].
        double ua[P], la[P], da[P], ra[P], ub[P], lb[P], db[P], rb[P];
        auto load = [&](double *u_, double *l_, double *d_, double *r_, int t0) {
#pragma unroll
            for (int q = 0; q < P; q++) {
                const int t = t0 + q;
                if (t < n) {
                    u_[q] = up_[(t - 1) * stride];
                    l_[q] = lo_[t * stride];
                    d_[q] = di_[t * stride];
                    r_[q] = v_[t * stride];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        auto step = [&](const double *u_, const double *l_, const double *d_, const double *r_,
                        int t0) {
#pragma unroll
            for (int q = 0; q < P; q++) {
                const int t = t0 + q;
                if (t < n) {
                    const double ct = u_[q] / beta;
                    cp_[t * stride] = ct;
                    beta = d_[q] - l_[q] * ct;
                    prev = (r_[q] - l_[q] * prev) / beta;
                    v_[t * stride] = prev;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        int t0 = 1;
        load(ua, la, da, ra, t0);
        for (; t0 + P < n; t0 += 2 * P) {
            load(ub, lb, db, rb, t0 + P);
            step(ua, la, da, ra, t0);
            load(ua, la, da, ra, t0 + 2 * P);
            step(ub, lb, db, rb, t0 + P);
        }
        step(ua, la, da, ra, t0);
    }
    {
        // Backward: positions t = n - 2 .. 0 use v[t] and cp[t + 1].
        double va[P], ca[P], vb[P], cb[P];
        auto load = [&](double *v2, double *c2, int t0) {
#pragma unroll
            for (int q = 0; q < P; q++) {
                const int t = t0 - q;
                if (t >= 0) {
                    v2[q] = v_[t * stride];
                    c2[q] = cp_[(t + 1) * stride];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        auto step = [&](const double *v2, const double *c2, int t0) {
#pragma unroll
            for (int q = 0; q < P; q++) {
                const int t = t0 - q;
                if (t >= 0) {
                    prev = v2[q] - c2[q] * prev;
                    v_[t * stride] = prev;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        int t0 = n - 2;
        load(va, ca, t0);
        for (; t0 - P >= 0; t0 -= 2 * P) {
            load(vb, cb, t0 - P);
            step(va, ca, t0);
            load(va, ca, t0 - 2 * P);
            step(vb, cb, t0 - P);
        }
        step(va, ca, t0);
    }
}

// x[f, j, i] = the column lines' value at position j of line f n + i.
__global__ __launch_bounds__(kBlock) void kc_lod_back(int n, const double *__restrict__ tmp,
                                                     double *__restrict__ x) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int nn = n * n;
    if (c >= 6 * nn) return;
    const int f = c / nn, rest = c - f * nn, j = rest / n, i = rest - j * n;
    x[c] = tmp[(size_t)j * (6 * (size_t)n) + (size_t)f * n + i];
}

// ---- ordered sums -------------------------------------------------------------------------

// One thread adds t[0 .. cnt) from LDS in order from `init` (the scalar steps' sums of the
// chunk sums).
__device__ __forceinline__ double chunk_sum(const double *t, int cnt, double init) {
    constexpr int D = 16;
    double s = init, a[D], b[D];
    int q = 0;
    // Software-pipelined over two register sets: while the current D terms are added, one by one
    // in order, the next D are read from LDS, the reads placed between the dependent adds (one
    // 16-byte read after each pair of adds) so that they issue in the adds' latency.
    auto stage = [&](double *next, int at, const double *cur) {
#pragma unroll
        for (int u = 0; u < D; u++) next[u] = t[at + u];
#pragma unroll
        for (int u = 0; u < D; u++) s += cur[u];
#pragma unroll
        for (int u = 0; u < D / 2; u++) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
            __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // two VALU
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto add = [&](const double *cur) {
#pragma unroll
        for (int u = 0; u < D; u++) s += cur[u];
    };
    if (cnt >= D) {
#pragma unroll
        for (int u = 0; u < D; u++) a[u] = t[u];
        __builtin_amdgcn_sched_barrier(0);
        for (q = D; q + 2 * D <= cnt; q += 2 * D) {
            stage(b, q, a);
            stage(a, q + D, b);
        }
        // a holds terms [q - D, q), not yet added.
        if (q + D <= cnt) {
            stage(b, q, a);
            add(b);
            q += D;
        } else {
            add(a);
        }
    }
    for (; q < cnt; q++) s += t[q];
    return s;
}

// Each chunk's terms are added one by one in order from `init` (the reference's chunk sums start
// from +0.0 in the CG passes and from -0.0 in `ordered_sum`): one dependent chain of f64 adds a
// chunk.  A wave runs four chains, one per 16-lane row.  A row reads its chunk straight from
// global memory, 32 terms a batch (lane k of the row holding terms 2k and 2k + 1), kRowP batches
// ahead of the adds, and each term reaches the whole row by a DPP row broadcast
// (row_newbcast:k), so that the chain costs a broadcast and an add a term and no LDS.

// bc<K>(old, v): lane K's v across its row.  Every lane has a source lane, so `old` is never
// read; it only orders the move after the instruction that produced `old`.
template <int K>
__device__ __forceinline__ double row_bcast(double old, double v) {
    return __builtin_amdgcn_update_dpp(old, v, 0x150 + K, 0xF, 0xF, false);
}

struct RowChain {
    double s1, s2;  // the running sum and its value one term earlier
};

template <int J>
__device__ __forceinline__ double row_term(double old, double a, double b) {
    return row_bcast<(J >> 1)>(old, (J & 1) ? b : a);
}

// Term J's add.  Before it: s1 = the sum through term J - 1, s2 = through J - 2 (no longer read),
// x = term J.  The broadcast of term J + 1 is written before the add, into s2's register, so it
// issues while the previous add completes.
template <int J>
__device__ __forceinline__ void row_step(RowChain &c, double &x, double a, double b) {
    double xn = 0.0;
    if constexpr (J < 31) xn = row_term<J + 1>(c.s2, a, b);
    c.s2 = c.s1;
    c.s1 = c.s1 + x;
    x = xn;
}

template <int... J>
__device__ __forceinline__ void row_add32(RowChain &c, double a, double b,
                                          std::integer_sequence<int, J...>) {
    double x = row_term<0>(c.s2, a, b);
    (row_step<J>(c, x, a, b), ...);
}

// The chain of this lane's row over terms [0, cnt) of a chunk (cnt <= kCubeChunk, the same for
// the row's 16 lanes), from `init`; the sum is in every lane of the row.  Src gives
// `Raw load(i)` (the raw values behind terms i and i + 1, i even; only those below cnt are read)
// and `double2 terms(const Raw &, i)`; terms at or past cnt are taken as -0.0, which leaves
// every sum unchanged.
template <int P, typename Src>
__device__ __forceinline__ double row_chain(const Src &src, int cnt, double init) {
    constexpr int NB = kCubeChunk / 32;
    const int k2 = 2 * (threadIdx.x & 15);
    typename Src::Raw buf[P];
#pragma unroll
    for (int p = 0; p < P; p++) buf[p] = src.load(32 * p + k2, cnt);
    RowChain c{init, init};
    for (int b0 = 0; b0 < NB; b0 += P) {
#pragma unroll
        for (int p = 0; p < P; p++) {
            const typename Src::Raw raw = buf[p];
            if (b0 + P + p < NB) buf[p] = src.load(32 * (b0 + P + p) + k2, cnt);
            const int i = 32 * (b0 + p) + k2;
            double2 x = src.terms(raw, i);
            if (i >= cnt) x.x = -0.0;
            if (i + 1 >= cnt) x.y = -0.0;
            row_add32(c, x.x, x.y, std::make_integer_sequence<int, 32>{});
        }
    }
    return c.s1;
}

// Two consecutive f64 values of p from i (i even), only those below cnt read.
__device__ __forceinline__ double2 load_pair(const double *p, int i, int cnt) {
    if (i + 1 < cnt) return *reinterpret_cast<const double2 *>(p + i);
    return make_double2(i < cnt ? p[i] : 0.0, 0.0);
}

__device__ __forceinline__ float2 load_pair(const float *p, int i, int cnt) {
    if (i + 1 < cnt) return *reinterpret_cast<const float2 *>(p + i);
    return make_float2(i < cnt ? p[i] : 0.0f, 0.0f);
}

// The chunk-sum kernels: one wave a block, chain c of a kernel on row c % 4 of wave c / 4.
constexpr int kRowBlock = 64, kRowChains = 4;

struct RowChunk {
    int chain, c0, cnt;  // this row's chain, its chunk's first cell and its term count (0: none)
};

__device__ __forceinline__ RowChunk row_chunk(int l, int per_chunk) {
    RowChunk r;
    r.chain = blockIdx.x * kRowChains + (threadIdx.x >> 4);
    r.c0 = (r.chain / per_chunk) * kCubeChunk;
    r.cnt = max(0, min(kCubeChunk, l - r.c0));
    return r;
}

// parts[chunk] = sum of the chunk's terms from -0.0 (`ordered_sum`): a[i] * b[i] (kDotF64),
// a[i] * (f64) b32[i] (kDotF32) or a[i] * b[i] / div[i] (kDotDiv).
enum { kDotF64 = 0, kDotF32 = 1, kDotDiv = 2 };

template <int MODE>
struct DotSrc {
    const double *a, *b, *div;
    const float *b32;
    struct Raw {
        double2 a, b, d;
        float2 b32;
    };
    __device__ __forceinline__ Raw load(int i, int cnt) const {
        Raw r;
        r.a = load_pair(a, i, cnt);
        if constexpr (MODE == kDotF32) r.b32 = load_pair(b32, i, cnt);
        else r.b = load_pair(b, i, cnt);
        if constexpr (MODE == kDotDiv) r.d = load_pair(div, i, cnt);
        return r;
    }
    __device__ __forceinline__ double2 terms(const Raw &r, int) const {
        if constexpr (MODE == kDotF32)
            return make_double2(r.a.x * (double)r.b32.x, r.a.y * (double)r.b32.y);
        else if constexpr (MODE == kDotDiv)
            return make_double2(r.a.x * r.b.x / r.d.x, r.a.y * r.b.y / r.d.y);
        else
            return make_double2(r.a.x * r.b.x, r.a.y * r.b.y);
    }
};

template <int MODE>
__global__ __launch_bounds__(kRowBlock) void kc_dot(int l, const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   const float *__restrict__ b32,
                                                   const double *__restrict__ div,
                                                   double *__restrict__ parts,
                                                   const CubeState *__restrict__ st) {
    if (st && st->done) return;
    const RowChunk r = row_chunk(l, 1);
    const DotSrc<MODE> src{a + r.c0, b ? b + r.c0 : nullptr, div ? div + r.c0 : nullptr,
                           b32 ? b32 + r.c0 : nullptr};
    const double s = row_chain<MODE == kDotF32 ? 16 : 8>(src, r.cnt, -0.0);
    if ((threadIdx.x & 15) == 0 && r.cnt > 0) parts[r.chain] = s;
}

// The direction pass (CD:411-440), one thread a cell: p = z + beta p_old (z = r / diag for the
// diagonal), A p, and the term p A p into t (added per chunk by kc_chunks).  Every neighbour's p
// is formed again by the same expression.
__global__ __launch_bounds__(kBlock) void kc_dir(int l, int mg, int side,
                                                const double *__restrict__ r,
                                                const float *__restrict__ z32,
                                                const double *__restrict__ diag,
                                                const double *__restrict__ k,
                                                const uint32_t *__restrict__ cols,
                                                const double *__restrict__ pold,
                                                double *__restrict__ pnew, double *__restrict__ ap,
                                                double *__restrict__ t,
                                                const CubeState *__restrict__ st) {
    if (st->done) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= l) return;
    const bool hb = st->have_beta;
    const double beta = st->beta;
    auto dir = [&](uint32_t j) {
        const double zj = mg ? (double)z32[j] : r[j] / diag[j];
        return hb ? zj + beta * pold[j] : zj;
    };
    const double pi = dir(i);
    double acc = diag[i] * pi;
    uint32_t j[8];
    cube_nbrs(i, l, side, cols, j);
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (j[q] != kNoCell) acc -= k[(size_t)q * l + i] * dir(j[q]);
    pnew[i] = pi;
    ap[i] = acc;
    t[i] = pi * acc;
}

// parts[chunk] (and parts[2 chunk], parts[2 chunk + 1] with two arrays) = the chunk's terms
// added in order from `init`.
struct ArraySrc {
    const double *p;
    using Raw = double2;
    __device__ __forceinline__ Raw load(int i, int cnt) const { return load_pair(p, i, cnt); }
    __device__ __forceinline__ double2 terms(const Raw &r, int) const { return r; }
};

__global__ __launch_bounds__(kRowBlock) void kc_chunks(int l, const double *__restrict__ t0,
                                                      const double *__restrict__ t1, double init,
                                                      double *__restrict__ parts,
                                                      const CubeState *__restrict__ st) {
    if (st && st->done) return;
    const RowChunk r = row_chunk(l, t1 ? 2 : 1);
    const double *src = (t1 && (r.chain & 1)) ? t1 : t0;
    const double s = row_chain<16>(ArraySrc{src + r.c0}, r.cnt, init);
    if ((threadIdx.x & 15) == 0 && r.cnt > 0) parts[r.chain] = s;
}

// The update pass (CD:449-470), one thread a cell: x += alpha p, r -= alpha A p, and the terms
// r r and r r / diag into t0 and t1 (added per chunk from +0.0 by kc_chunks).
__global__ __launch_bounds__(kBlock) void kc_update(int l, const double *__restrict__ p,
                                                   const double *__restrict__ ap,
                                                   const double *__restrict__ diag,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ t0,
                                                   double *__restrict__ t1,
                                                   const CubeState *__restrict__ st) {
    if (st->done) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= l) return;
    const double alpha = st->alpha;
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * ap[i];
    r[i] = rv;
    t0[i] = rv * rv;
    if (t1) t1[i] = rv * rv / diag[i];  // the diagonal's rz (Jacobi only)
}

// The scalar steps, one thread: the parts added in order from -0.0 (`iter().sum()`).
enum { kCubeRR0 = 0, kCubeRZ0 = 1, kCubePAP = 2, kCubeRR = 3, kCubeRZ = 4, kCubePutback = 5 };

// The block stages the parts in LDS, a tile at a time, and thread 0 adds them in order.
constexpr int kStepTile = 2048;

template <int STEP>
__global__ __launch_bounds__(kBlock) void kc_step(const double *__restrict__ parts, int np,
                                                 CubeState *st, int mg, int *flag) {
    __shared__ double ta[kStepTile], tb[kStepTile];
    if (STEP != kCubeRR0 && STEP != kCubePutback && st->done) {
        if (flag && threadIdx.x == 0) *reinterpret_cast<volatile int *>(flag) = st->done;
        return;
    }
    // r.r with the diagonal's rz beside it (Jacobi), or alone (multigrid).
    const int stride = ((STEP == kCubeRR && !mg) || STEP == kCubePutback) ? 2 : 1;
    double a = -0.0, b = -0.0;
    for (int q0 = 0; q0 < np; q0 += kStepTile) {
        const int cnt = min(kStepTile, np - q0);
        constexpr int per = kStepTile / kBlock;
        double va[per], vb[per];
#pragma unroll
        for (int u = 0; u < per; u++) {
            const int q = (int)threadIdx.x + u * kBlock;
            if (q < cnt) {
                va[u] = parts[stride * (q0 + q)];
                if (stride == 2) vb[u] = parts[2 * (q0 + q) + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < per; u++) {
            const int q = (int)threadIdx.x + u * kBlock;
            if (q < cnt) {
                ta[q] = va[u];
                if (stride == 2) tb[q] = vb[u];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            a = chunk_sum(ta, cnt, a);
            if (stride == 2) b = chunk_sum(tb, cnt, b);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (STEP == kCubeRR0) {
        // rr0 <= tol2 ends the solve before the loop (CD:392-394).
        st->have_beta = 0, st->its = 0, st->done = a <= st->tol2 ? 1 : 0;
    } else if (STEP == kCubeRZ0) {
        st->rz = a;
    } else if (STEP == kCubePAP) {
        // The p.A p break (CD:443-446); alpha = rz / pap.
        st->pap = a;
        if (isnan(a) || a <= 0.0) {
            st->done = 2;
        } else {
            st->alpha = st->rz / a;
        }
    } else if (STEP == kCubeRR) {
        // iterations += 1; the r.r stop (CD:476-479); the diagonal's rz (CD:480-486).
        st->its += 1;
        if (a <= st->tol2) {
            st->done = 1;
        } else if (st->its >= st->max_iter) {
            st->done = 4;
        } else if (!mg) {
            st->beta = b / st->rz, st->have_beta = 1, st->rz = b;
        }
    } else if (STEP == kCubeRZ) {
        st->beta = a / st->rz, st->have_beta = 1, st->rz = a;
    } else {
        st->lost = a, st->weight = b;
    }
    if (flag) *reinterpret_cast<volatile int *>(flag) = st->done;
}

// The put-back's sums (CD:289-292): per chunk, sum (rhs - A x) into parts[2 chunk] and sum s
// into parts[2 chunk + 1], each from -0.0.
struct DiffSrc {
    const double *u, *v;  // terms u - v (v null: u)
    struct Raw {
        double2 u, v;
    };
    __device__ __forceinline__ Raw load(int i, int cnt) const {
        Raw r;
        r.u = load_pair(u, i, cnt);
        if (v) r.v = load_pair(v, i, cnt);
        return r;
    }
    __device__ __forceinline__ double2 terms(const Raw &r, int) const {
        return v ? make_double2(r.u.x - r.v.x, r.u.y - r.v.y) : r.u;
    }
};

__global__ __launch_bounds__(kRowBlock) void kc_putback_sums(int l, const double *__restrict__ rhs,
                                                            const double *__restrict__ ax,
                                                            const double *__restrict__ s,
                                                            double *__restrict__ parts) {
    const RowChunk r = row_chunk(l, 2);
    const DiffSrc src = (r.chain & 1) ? DiffSrc{s + r.c0, nullptr}
                                      : DiffSrc{rhs + r.c0, ax + r.c0};
    const double sum = row_chain<8>(src, r.cnt, -0.0);
    if ((threadIdx.x & 15) == 0 && r.cnt > 0) parts[r.chain] = sum;
}

// x += (lost / weight) s / mu when weight > 0 and lost != 0 (CD:293-298), then h[cells[i]] = x.
__global__ __launch_bounds__(kBlock) void kc_putback(int l, const uint32_t *__restrict__ cells,
                                                    const double *__restrict__ s,
                                                    const double *__restrict__ mu,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ h,
                                                    const CubeState *__restrict__ st) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= l) return;
    double v = x[i];
    if (st->weight > 0.0 && st->lost != 0.0) {
        const double kk = st->lost / st->weight;
        v += kk * s[i] / mu[i];
    }
    h[cells ? cells[i] : (uint32_t)i] = v;
}

__global__ void kc_init(CubeState *st, double tol2, int max_iter) {
    if (threadIdx.x == 0) *st = CubeState{0, 0, 0, 0, tol2, 0, 0, 0, 0, 0, max_iter};
}

// ---- the cube's V-cycle (CM), f32 ------------------------------------------------------------

// A level's couplings in ELL form, slot-major: entry q of cell c at q cells + c, the row's
// entries first in the reference's order (its CSR order) and kNoCell after them, so a wave's
// loads of one entry are contiguous.
struct CLvl {
    int cells, ne, w;                    // cells, CSR entries, ELL width
    int side;                            // > 0: the finest level of a structured table
    const uint32_t *col;                 // ELL neighbours
    const uint32_t *ell_of;              // per CSR entry, its ELL position (level >= 1)
    const uint32_t *src_ptr, *src;       // per CSR entry, the finer ELL positions (level >= 1)
    const uint32_t *kid_ptr, *kids;      // per cell, its finer block, ascending (level >= 1)
    const uint32_t *parent;              // per finer cell, its cell here (level >= 1)
    float *val, *mass, *diag, *r, *z, *t;
};

// The finest level's values (CM:205-230): the couplings in slot order (the CSR order), the mass
// and the f64 diagonal, in f32.
__global__ __launch_bounds__(kBlock) void kc_mg_finest(CLvl L, const uint32_t *__restrict__ slots,
                                                      const double *__restrict__ k,
                                                      const double *__restrict__ mu,
                                                      const double *__restrict__ diag) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= L.cells) return;
    size_t at = c;
    for (int q = 0; q < 8; q++)
        if (slots[(size_t)q * L.cells + c] != kNoCell) {
            L.val[at] = (float)k[(size_t)q * L.cells + c];
            at += L.cells;
        }
    L.mass[c] = (float)mu[c];
    L.diag[c] = (float)diag[c];
}

// A coarse entry: scale times the sum of its finer entries (CM:236-243).
__global__ __launch_bounds__(kBlock) void kc_mg_vals(CLvl F, CLvl C, float scale) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= C.ne) return;
    float s = 0.0f;
    for (uint32_t q = C.src_ptr[e]; q < C.src_ptr[e + 1]; q++) s += F.val[C.src[q]];
    C.val[C.ell_of[e]] = scale * s;
}

// A coarse cell's mass (its block's, summed) and diagonal, mass + its couplings (CM:244-258).
__global__ __launch_bounds__(kBlock) void kc_mg_diag(CLvl F, CLvl C) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= C.cells) return;
    float s = 0.0f, k = 0.0f;
    for (uint32_t q = C.kid_ptr[r]; q < C.kid_ptr[r + 1]; q++) s += F.mass[C.kids[q]];
    for (int q = 0; q < C.w; q++) {
        const size_t e = (size_t)q * C.cells + r;
        if (C.col[e] == kNoCell) break;
        k += C.val[e];
    }
    C.mass[r] = s;
    C.diag[r] = s + k;
}

// r32 = r (f64 -> f32, CM:263-268).
__global__ __launch_bounds__(kBlock) void kc_mg_r32(int n, const double *__restrict__ r,
                                                   float *__restrict__ r32,
                                                   const CubeState *__restrict__ st) {
    if (st->done) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) r32[i] = (float)r[i];
}

// diag z_i - sum of the row's couplings times z, in the row's order.  W = 8 reads all eight
// entries' columns at once (the rows are at most eight long); W = 0 walks a wider row.
template <int W>
__device__ __forceinline__ float cube_az(const CLvl &L, const float *z, int i) {
    float az = L.diag[i] * z[i];
    if constexpr (W == 8) {
        uint32_t j[W];
        cube_nbrs(i, L.cells, L.side, L.col, j);  // an inner cell's ELL row is its 8 slots
#pragma unroll
        for (int q = 0; q < W; q++)
            if (j[q] != kNoCell) az -= L.val[(size_t)q * L.cells + i] * z[j[q]];
    } else {
        for (int q = 0; q < L.w; q++) {
            const uint32_t jq = L.col[(size_t)q * L.cells + i];
            if (jq == kNoCell) break;
            az -= L.val[(size_t)q * L.cells + i] * z[jq];
        }
    }
    return az;
}

// `sweep`, CM:285-310: z = omega r / diag from zero (first), else zout = z + omega (r - A z) / diag.
template <int W>
__global__ __launch_bounds__(kBlock) void kc_mg_sweep(CLvl L, const float *__restrict__ zin,
                                                     float *__restrict__ zout, float omega,
                                                     int first, const CubeState *__restrict__ st) {
    if (st->done) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= L.cells) return;
    if (first)
        zout[i] = omega * L.r[i] / L.diag[i];
    else
        zout[i] = zin[i] + omega * (L.r[i] - cube_az<W>(L, zin, i)) / L.diag[i];
}

// The residual r - A z of each cell (CM:328-341).
template <int W>
__global__ __launch_bounds__(kBlock) void kc_mg_resid(CLvl L, const float *__restrict__ z,
                                                     float *__restrict__ t,
                                                     const CubeState *__restrict__ st) {
    if (st->done) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < L.cells) t[i] = L.r[i] - cube_az<W>(L, z, i);
}

// The coarse r: each block's residuals summed, its cells ascending (CM:343-351).
__global__ __launch_bounds__(kBlock) void kc_mg_restrict(CLvl C, const float *__restrict__ t,
                                                        const CubeState *__restrict__ st) {
    if (st->done) return;
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= C.cells) return;
    float s = 0.0f;
    for (uint32_t q = C.kid_ptr[r]; q < C.kid_ptr[r + 1]; q++) s += t[C.kids[q]];
    C.r[r] = s;
}

// z += the coarse correction of the cell's block (CM:353-358).
__global__ __launch_bounds__(kBlock) void kc_mg_prolong(int cells, float *__restrict__ z,
                                                       const float *__restrict__ zc,
                                                       const uint32_t *__restrict__ parent,
                                                       const CubeState *__restrict__ st) {
    if (st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < cells) z[c] += zc[parent[c]];
}

// ---- the cube's coarse levels in one block ----------------------------------------------------

// The levels of at most kCubeBottomCells cells in all are cycled by one block of
// kCubeBottomThreads, with each level's z, its second buffer and r in LDS: the same per-cell
// arithmetic as the per-level kernels, with block barriers in place of launches.
constexpr int kCubeBottomLevels = 8, kCubeBottomThreads = 1024;

struct CBottom {
    int nlev;
    int off[kCubeBottomLevels];  // each level's first cell in the LDS arrays
    CLvl lv[kCubeBottomLevels];
};

template <int W>
__global__ __launch_bounds__(kCubeBottomThreads) void kc_mg_bottom(
    CBottom B, float *__restrict__ ztop, float omega, const CubeState *__restrict__ st) {
    extern __shared__ float cb_mem[];
    if (st->done) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    unsigned cur = 0;  // bit l: which of level l's two buffers holds z
    auto zbuf = [&](int l, unsigned which) {
        return cb_mem + 3 * B.off[l] + (which ? B.lv[l].cells : 0);
    };
    auto rbuf = [&](int l) { return cb_mem + 3 * B.off[l] + 2 * B.lv[l].cells; };
    auto bit = [&](int l) { return (cur >> l) & 1u; };
    for (int i = tid; i < B.lv[0].cells; i += nt) rbuf(0)[i] = B.lv[0].r[i];
    __syncthreads();
    // `sweep`, CM:285-310.
    auto sweep = [&](int l, bool first) {
        const CLvl &L = B.lv[l];
        const float *zi = zbuf(l, bit(l)), *r = rbuf(l);
        float *zo = zbuf(l, first ? bit(l) : bit(l) ^ 1u);
        for (int i = tid; i < L.cells; i += nt) {
            if (first)
                zo[i] = omega * r[i] / L.diag[i];
            else
                zo[i] = zi[i] + omega * (r[i] - cube_az<W>(L, zi, i)) / L.diag[i];
        }
        __syncthreads();
        if (!first) cur ^= 1u << l;
    };
    for (int l = 0; l + 1 < B.nlev; l++) {
        const CLvl &L = B.lv[l], &C = B.lv[l + 1];
        for (int k = 0; k < kCubeNu; k++) sweep(l, k == 0);
        // The residual into the second buffer (CM:328-341), then the coarse r (CM:343-351).
        const float *z = zbuf(l, bit(l)), *r = rbuf(l);
        float *t = zbuf(l, bit(l) ^ 1u);
        for (int i = tid; i < L.cells; i += nt) t[i] = r[i] - cube_az<W>(L, z, i);
        __syncthreads();
        float *rc = rbuf(l + 1);
        for (int c = tid; c < C.cells; c += nt) {
            float s = 0.0f;
            for (uint32_t q = C.kid_ptr[c]; q < C.kid_ptr[c + 1]; q++) s += t[C.kids[q]];
            rc[c] = s;
        }
        __syncthreads();
    }
    for (int k = 0; k < kCubeCoarseSweeps; k++) sweep(B.nlev - 1, k == 0);
    for (int l = B.nlev - 2; l >= 0; l--) {
        // z += the coarse correction of the cell's block (CM:353-358), then the post-sweeps.
        float *z = zbuf(l, bit(l));
        const float *zc = zbuf(l + 1, bit(l + 1));
        const uint32_t *parent = B.lv[l + 1].parent;
        for (int i = tid; i < B.lv[l].cells; i += nt) z[i] += zc[parent[i]];
        __syncthreads();
        for (int k = 0; k < kCubeNu; k++) sweep(l, false);
    }
    const float *z = zbuf(0, bit(0));
    for (int i = tid; i < B.lv[0].cells; i += nt) ztop[i] = z[i];
}

// ---- host side -----------------------------------------------------------------------------

}  // namespace dtd

struct dtd_cube {
    int n = 0, n6 = 0;
    double mean_area = 0.0;
    std::vector<uint32_t> slots_h;  // host copy (the subset's neighbour positions)
    bool structured = false;        // each face row-major with the D8 slots of the reference
    std::vector<void *> owned;
    uint32_t *slots = nullptr;
    double *lap = nullptr, *area = nullptr;
    // The V-cycle's levels (patterns on the device; values filled per solve).
    std::vector<dtd::CLvl> lv;
    // Levels bottom .. end are cycled by kc_mg_bottom (bottom == lv.size(): none).
    size_t bottom = 0;
    int bottom_w8 = 1;
    size_t bottom_lds = 0;
    dtd::CBottom cb{};
    // Per-solve buffers (n6 cells).
    double *kfull = nullptr, *kpos = nullptr, *mu = nullptr, *s = nullptr, *diag = nullptr;
    double *rhs = nullptr, *x = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr;
    double *ap = nullptr, *b = nullptr, *tmp = nullptr, *parts = nullptr;
    uint32_t *cells = nullptr, *cols = nullptr;
    unsigned char *any = nullptr;
    dtd::CubeState *st = nullptr;
    dtd::CubeState *h_st = nullptr;  // pinned
    double *hd = nullptr, *kdd = nullptr;  // device copies for the host entry (first use)
};

namespace dtd {

// The cube's levels of at most this many cells in all are cycled by one block, kc_mg_bottom
// (DTD_CUBE_BOTTOM_CELLS overrides it; 0 turns it off).
static int cube_bottom_cells() {
    static const int v = [] {
        const char *e = getenv("DTD_CUBE_BOTTOM_CELLS");
        return e ? atoi(e) : 2048;
    }();
    return v;
}

// One level's pattern on the host (CM:29-45, 65-165).
struct CPat {
    size_t side = 0, cells = 0;
    std::vector<uint32_t> ptr, col, src_ptr, src, kid_ptr, kids, parent;
};

// `finest`, CM:65-90: the slots in order, the missing ones skipped.
static CPat cube_finest(int n, const uint32_t *slots) {
    CPat p;
    p.side = n, p.cells = 6 * (size_t)n * n;
    p.ptr.reserve(p.cells + 1);
    p.ptr.push_back(0);
    for (size_t c = 0; c < p.cells; c++) {
        for (int q = 0; q < 8; q++)
            if (slots[8 * c + q] != kNoCell) p.col.push_back(slots[8 * c + q]);
        p.ptr.push_back((uint32_t)p.col.size());
    }
    return p;
}

// `coarsen`, CM:93-165: 2 x 2 blocks of each face; each coarse cell's neighbours ascending, each
// with its finer entries in (kid, entry) order.
static CPat cube_coarsen(const CPat &f) {
    CPat c;
    const size_t nf = f.side, m = (nf + 1) / 2;
    c.side = m, c.cells = 6 * m * m;
    c.parent.resize(f.cells);
    for (size_t q = 0; q < f.cells; q++) {
        const size_t face = q / (nf * nf), rest = q % (nf * nf), i = rest % nf, j = rest / nf;
        c.parent[q] = (uint32_t)(face * m * m + (j / 2) * m + i / 2);
    }
    c.kid_ptr.assign(c.cells + 1, 0);
    for (uint32_t p : c.parent) c.kid_ptr[p + 1]++;
    for (size_t q = 0; q < c.cells; q++) c.kid_ptr[q + 1] += c.kid_ptr[q];
    c.kids.resize(f.cells);
    std::vector<uint32_t> fill(c.kid_ptr.begin(), c.kid_ptr.end() - 1);
    for (size_t q = 0; q < f.cells; q++) c.kids[fill[c.parent[q]]++] = (uint32_t)q;
    c.ptr.push_back(0);
    c.src_ptr.push_back(0);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (size_t r = 0; r < c.cells; r++) {
        pairs.clear();
        for (uint32_t q = c.kid_ptr[r]; q < c.kid_ptr[r + 1]; q++) {
            const uint32_t k = c.kids[q];
            for (uint32_t e = f.ptr[k]; e < f.ptr[k + 1]; e++) {
                const uint32_t p = c.parent[f.col[e]];
                if (p != r) pairs.emplace_back(p, e);
            }
        }
        std::stable_sort(pairs.begin(), pairs.end(),
                         [](const std::pair<uint32_t, uint32_t> &a,
                            const std::pair<uint32_t, uint32_t> &b) { return a.first < b.first; });
        for (size_t a = 0; a < pairs.size(); a++) {
            if (a == 0 || pairs[a].first != pairs[a - 1].first) {
                if (a > 0) c.src_ptr.push_back((uint32_t)c.src.size());
                c.col.push_back(pairs[a].first);
            }
            c.src.push_back(pairs[a].second);
        }
        if (!pairs.empty()) c.src_ptr.push_back((uint32_t)c.src.size());
        c.ptr.push_back((uint32_t)c.col.size());
    }
    return c;
}

static void free_cube(dtd_cube *g) {
    if (!g) return;
    for (void *p : g->owned) (void)hipFree(p);
    if (g->h_st) (void)hipHostFree(g->h_st);
    delete g;
}

// The grid on the device and its levels' patterns (`Hierarchy::new`, CM:170-186).
static dtd_cube *cube_new(dtd_ctx *ctx, int n, const uint32_t *slots, const double *lap,
                          const double *area, double mean_area) {
    dtd_cube *g = new dtd_cube();
    try {
        g->n = n, g->n6 = 6 * n * n, g->mean_area = mean_area;
        const size_t n6 = (size_t)g->n6;
        g->slots_h.assign(slots, slots + 8 * n6);
        // Is the table the structured one (each face row-major, slot q of an inner cell at
        // c + DI[q] + DJ[q] n, D8 order as surface/grid.rs:21-31)?  Then inner cells form their
        // neighbours instead of reading them.
        {
            static const int DI[8] = {1, 1, 0, -1, -1, -1, 0, 1}, DJ[8] = {0, -1, -1, -1, 0, 1, 1, 1};
            bool st = n >= 3;
            for (size_t f = 0; st && f < 6; f++)
                for (int j = 1; st && j + 1 < n; j++)
                    for (int i = 1; st && i + 1 < n; i++) {
                        const size_t c = f * (size_t)n * n + (size_t)j * n + i;
                        for (int q = 0; q < 8; q++)
                            if (slots[8 * c + q] != (uint32_t)((long)c + DI[q] + (long)DJ[q] * n)) {
                                st = false;
                                break;
                            }
                    }
            g->structured = st;
        }
        auto own = [&](size_t bytes) {
            void *p = nullptr;
            HIPCHK(hipMalloc(&p, bytes ? bytes : 1));
            g->owned.push_back(p);
            return p;
        };
        auto put = [&](const void *src, size_t bytes) {
            void *p = own(bytes);
            HIPCHK(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
            return p;
        };
        // Slot-major on the device (slot q of cell c at q n6 + c), so a wave's loads of one slot
        // are contiguous.
        std::vector<uint32_t> slots_t(8 * n6);
        std::vector<double> lap_t(8 * n6);
        for (size_t c = 0; c < n6; c++)
            for (int q = 0; q < 8; q++)
                slots_t[q * n6 + c] = slots[8 * c + q], lap_t[q * n6 + c] = lap[8 * c + q];
        g->slots = (uint32_t *)put(slots_t.data(), 8 * n6 * 4);
        g->lap = (double *)put(lap_t.data(), 8 * n6 * 8);
        HIPCHK(hipStreamSynchronize(ctx->stream));  // the host copies end with this block
        g->area = (double *)put(area, n6 * 8);
        std::vector<CPat> pats;
        pats.push_back(cube_finest(n, slots));
        while (!(pats.back().cells <= (size_t)kCubeCoarsest || pats.back().side <= 1))
            pats.push_back(cube_coarsen(pats.back()));
        // Each level's ELL form; a coarse entry's finer entries are named by their ELL position.
        std::vector<std::vector<uint32_t>> ell_of(pats.size()), ecol(pats.size());
        std::vector<int> width(pats.size());
        for (size_t l = 0; l < pats.size(); l++) {
            const CPat &p = pats[l];
            int w = 1;
            for (size_t c = 0; c < p.cells; c++) w = std::max(w, (int)(p.ptr[c + 1] - p.ptr[c]));
            width[l] = w;
            ecol[l].assign((size_t)w * p.cells, kNoCell);
            ell_of[l].resize(p.col.size());
            for (size_t c = 0; c < p.cells; c++)
                for (uint32_t e = p.ptr[c]; e < p.ptr[c + 1]; e++) {
                    const size_t at = (size_t)(e - p.ptr[c]) * p.cells + c;
                    if (at > 0xFFFFFFFEu)
                        throw Error{DTD_ETOOBIG, "the cube's levels exceed 2^32 entries"};
                    ecol[l][at] = p.col[e];
                    ell_of[l][e] = (uint32_t)at;
                }
        }
        for (size_t l = 0; l < pats.size(); l++) {
            const CPat &p = pats[l];
            CLvl L{};
            L.cells = (int)p.cells, L.ne = (int)p.col.size(), L.w = width[l];
            L.side = l == 0 && g->structured ? n : 0;
            L.col = (const uint32_t *)put(ecol[l].data(), ecol[l].size() * 4);
            if (!p.parent.empty()) {
                std::vector<uint32_t> src(p.src.size());
                for (size_t q = 0; q < src.size(); q++) src[q] = ell_of[l - 1][p.src[q]];
                L.ell_of = (const uint32_t *)put(ell_of[l].data(), ell_of[l].size() * 4);
                L.src_ptr = (const uint32_t *)put(p.src_ptr.data(), p.src_ptr.size() * 4);
                L.src = (const uint32_t *)put(src.data(), src.size() * 4);
                L.kid_ptr = (const uint32_t *)put(p.kid_ptr.data(), p.kid_ptr.size() * 4);
                L.kids = (const uint32_t *)put(p.kids.data(), p.kids.size() * 4);
                L.parent = (const uint32_t *)put(p.parent.data(), p.parent.size() * 4);
                HIPCHK(hipStreamSynchronize(ctx->stream));  // src ends with this block
            }
            L.val = (float *)own(ecol[l].size() * 4);
            L.mass = (float *)own(p.cells * 4);
            L.diag = (float *)own(p.cells * 4);
            L.r = (float *)own(p.cells * 4);
            L.z = (float *)own(p.cells * 4);
            L.t = (float *)own(p.cells * 4);
            g->lv.push_back(L);
        }
        // The bottom levels: from the coarsest up while they hold at most cube_bottom_cells().
        g->bottom = g->lv.size();
        size_t tot = 0;
        for (size_t l = g->lv.size(); l-- > 0;) {
            tot += (size_t)g->lv[l].cells;
            if (tot > (size_t)cube_bottom_cells() || 3 * tot * sizeof(float) > 65536 ||
                g->lv.size() - l > (size_t)kCubeBottomLevels)
                break;
            g->bottom = l;
        }
        if (g->bottom < g->lv.size()) {
            CBottom &B = g->cb;
            B.nlev = (int)(g->lv.size() - g->bottom);
            int off = 0;
            for (int q = 0; q < B.nlev; q++) {
                B.lv[q] = g->lv[g->bottom + q];
                B.off[q] = off;
                off += B.lv[q].cells;
                if (B.lv[q].w > 8) g->bottom_w8 = 0;
            }
            g->bottom_lds = 3 * (size_t)off * sizeof(float);
        }
        for (double **v : {&g->kfull, &g->kpos})
            *v = (double *)own(8 * n6 * 8);
        for (double **v : {&g->mu, &g->s, &g->diag, &g->rhs, &g->x, &g->r, &g->p0, &g->p1, &g->ap,
                           &g->b, &g->tmp})
            *v = (double *)own(n6 * 8);
        g->parts = (double *)own(2 * ((n6 + kCubeChunk - 1) / kCubeChunk + 1) * 8);
        g->cells = (uint32_t *)own(n6 * 4);
        g->cols = (uint32_t *)own(8 * n6 * 4);
        g->any = (unsigned char *)own(n6);
        g->st = (CubeState *)own(sizeof(CubeState));
        HIPCHK(hipHostMalloc((void **)&g->h_st, sizeof(CubeState), hipHostMallocDefault));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return g;
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        free_cube(g);
        throw;
    }
}

// One V-cycle of the cube's levels from level l (`vcycle`, CM:312-363); the levels' z buffers
// alternate with their t buffers as the reference swaps them, so zs[l] tracks which holds z.
static void cube_vcycle(dtd_ctx *ctx, dtd_cube *g, size_t l, std::vector<float *> &zs,
                        std::vector<float *> &ts) {
    hipStream_t s = ctx->stream;
    const float omega = ctx->opts.omega;
    CLvl L = g->lv[l];
    L.z = zs[l], L.t = ts[l];
    const int nb = nblocks(L.cells);
    if (l == g->bottom) {
        const size_t lds = g->bottom_lds;
        if (g->bottom_w8)
            kc_mg_bottom<8><<<1, kCubeBottomThreads, lds, s>>>(g->cb, zs[l], omega, g->st);
        else
            kc_mg_bottom<0><<<1, kCubeBottomThreads, lds, s>>>(g->cb, zs[l], omega, g->st);
        LAUNCH_CHECK();
        return;
    }
    const bool w8 = L.w <= 8;
    auto sweep = [&](bool first) {
        float *zout = first ? zs[l] : ts[l];
        if (w8)
            kc_mg_sweep<8><<<nb, kBlock, 0, s>>>(L, zs[l], zout, omega, first, g->st);
        else
            kc_mg_sweep<0><<<nb, kBlock, 0, s>>>(L, zs[l], zout, omega, first, g->st);
        LAUNCH_CHECK();
        if (!first) std::swap(zs[l], ts[l]);
    };
    if (l + 1 == g->lv.size()) {
        for (int k = 0; k < kCubeCoarseSweeps; k++) sweep(k == 0);
        return;
    }
    for (int k = 0; k < kCubeNu; k++) sweep(k == 0);
    if (w8)
        kc_mg_resid<8><<<nb, kBlock, 0, s>>>(L, zs[l], ts[l], g->st);
    else
        kc_mg_resid<0><<<nb, kBlock, 0, s>>>(L, zs[l], ts[l], g->st);
    LAUNCH_CHECK();
    const CLvl &C = g->lv[l + 1];
    kc_mg_restrict<<<nblocks(C.cells), kBlock, 0, s>>>(C, ts[l], g->st);
    LAUNCH_CHECK();
    cube_vcycle(ctx, g, l + 1, zs, ts);
    kc_mg_prolong<<<nb, kBlock, 0, s>>>(L.cells, zs[l], zs[l + 1], C.parent, g->st);
    LAUNCH_CHECK();
    for (int k = 0; k < kCubeNu; k++) sweep(false);
}

// z = V-cycle(r) into the finest level's z (zs[0] after the call).
static const float *cube_precondition(dtd_ctx *ctx, dtd_cube *g, const double *r) {
    std::vector<float *> zs, ts;
    for (const CLvl &L : g->lv) zs.push_back(L.z), ts.push_back(L.t);
    kc_mg_r32<<<nblocks(g->lv[0].cells), kBlock, 0, ctx->stream>>>(g->lv[0].cells, r, g->lv[0].r,
                                                                  g->st);
    LAUNCH_CHECK();
    cube_vcycle(ctx, g, 0, zs, ts);
    // The next cycle starts from the same buffers: hand the final z back as level 0's z.
    return zs[0];
}

// `solve` / `step_in` / `run`, CD:227-314, on device arrays h (in/out) and kd.
static size_t cube_solve(dtd_ctx *ctx, dtd_cube *g, double *h, const double *kd, double dt,
                         int solver, double tol_m) {
    if (std::isnan(dt) || dt <= 0.0) return 0;
    hipStream_t s = ctx->stream;
    const bool mg = solver == DTD_SOLVER_MULTIGRID;
    const int n = g->n, n6 = g->n6;
    kc_couplings<<<nblocks(n6), kBlock, 0, s>>>(n6, g->slots, g->lap, kd, dt / g->mean_area,
                                                g->kfull, g->any);
    LAUNCH_CHECK();
    // Every cell taking part (the multigrid solve, or every cell coupled) is the grid's own
    // operator; else the cells with a coupling, ascending, with their neighbours as positions.
    bool all = mg;
    int l = n6;
    if (!mg) {
        std::vector<unsigned char> any(n6);
        download(ctx, any.data(), g->any, n6);
        all = std::all_of(any.begin(), any.end(), [](unsigned char v) { return v != 0; });
        if (!all) {
            std::vector<uint32_t> pos(n6, kNoCell), cells;
            for (int c = 0; c < n6; c++)
                if (any[c]) pos[c] = (uint32_t)cells.size(), cells.push_back((uint32_t)c);
            l = (int)cells.size();
            if (l == 0) return 0;
            std::vector<uint32_t> cols(8 * (size_t)l);
            for (int i = 0; i < l; i++)
                for (int q = 0; q < 8; q++) {
                    const uint32_t m = g->slots_h[8 * (size_t)cells[i] + q];
                    cols[(size_t)q * l + i] = m == kNoCell ? kNoCell : pos[m];
                }
            upload(ctx, g->cells, cells.data(), (size_t)l * 4);
            upload(ctx, g->cols, cols.data(), cols.size() * 4);
            HIPCHK(hipStreamSynchronize(s));  // the host lists end with this block
        }
    }
    const uint32_t *cells = all ? nullptr : g->cells;
    const uint32_t *cols = all ? g->slots : g->cols;
    const int side = all && g->structured ? n : 0;  // inner cells form their neighbours
    double *k = all ? g->kfull : g->kpos;
    const int nb = nblocks(l), nc = (l + kCubeChunk - 1) / kCubeChunk;
    // The chunk-sum kernels' grids: four chains a wave (one or two chains a chunk).
    const int bc = (nc + kRowChains - 1) / kRowChains;
    const int bc2 = (2 * nc + kRowChains - 1) / kRowChains;
    kc_opinit<<<nb, kBlock, 0, s>>>(l, n6, cells, g->kfull, k, g->area, g->mean_area, h, g->mu,
                                    g->s, g->diag, g->rhs, g->x);
    LAUNCH_CHECK();
    if (all) {
        // The first guess from b = h (CD:280-284).
        HIPCHK(hipMemcpyAsync(g->b, g->x, (size_t)n6 * 8, hipMemcpyDeviceToDevice, s));
        const int nl = (6 * n + 63) / 64;
        // kpos is free when every cell takes part: it holds the lines' coefficients.
        double *lo = g->kpos, *di = lo + n6, *up = di + n6, *cp = up + n6;
        kc_lod_coef<<<nblocks(n6), kBlock, 0, s>>>(n, 0, 0, k, g->mu, g->slots, g->b, lo, di, up,
                                                   g->x);
        LAUNCH_CHECK();
        kc_thomas<<<nl, 64, 0, s>>>(6 * n, n, lo, di, up, g->x, cp);
        LAUNCH_CHECK();
        kc_lod_coef<<<nblocks(n6), kBlock, 0, s>>>(n, 1, 1, k, g->mu, g->slots, g->x, lo, di, up,
                                                   g->tmp);
        LAUNCH_CHECK();
        kc_thomas<<<nl, 64, 0, s>>>(6 * n, n, lo, di, up, g->tmp, cp);
        LAUNCH_CHECK();
        kc_lod_back<<<nblocks(n6), kBlock, 0, s>>>(n, g->tmp, g->x);
        LAUNCH_CHECK();
    }
    if (mg) {
        kc_mg_finest<<<nblocks(n6), kBlock, 0, s>>>(g->lv[0], g->slots, k, g->mu, g->diag);
        LAUNCH_CHECK();
        for (size_t q = 1; q < g->lv.size(); q++) {
            kc_mg_vals<<<nblocks(g->lv[q].ne), kBlock, 0, s>>>(g->lv[q - 1], g->lv[q],
                                                                ctx->opts.scale);
            LAUNCH_CHECK();
            kc_mg_diag<<<nblocks(g->lv[q].cells), kBlock, 0, s>>>(g->lv[q - 1], g->lv[q]);
            LAUNCH_CHECK();
        }
    }
    // The loop (CD:380-500).
    const double tol = mg ? tol_m : kCgTolM;
    kc_init<<<1, 64, 0, s>>>(g->st, tol * tol * (double)n6, (int)ctx->opts.max_iter);
    LAUNCH_CHECK();
    kc_apply<<<nb, kBlock, 0, s>>>(l, side, g->diag, k, cols, g->x, g->r, g->rhs);
    LAUNCH_CHECK();
    kc_dot<kDotF64><<<bc, kRowBlock, 0, s>>>(l, g->r, g->r, nullptr, nullptr, g->parts,
                                             nullptr);
    LAUNCH_CHECK();
    kc_step<kCubeRR0><<<1, kBlock, 0, s>>>(g->parts, nc, g->st, mg, nullptr);
    LAUNCH_CHECK();
    const float *z32 = nullptr;
    if (mg) {
        z32 = cube_precondition(ctx, g, g->r);
        kc_dot<kDotF32><<<bc, kRowBlock, 0, s>>>(l, g->r, nullptr, z32, nullptr, g->parts,
                                                 g->st);
    } else {
        kc_dot<kDotDiv><<<bc, kRowBlock, 0, s>>>(l, g->r, g->r, nullptr, g->diag, g->parts,
                                                 g->st);
    }
    LAUNCH_CHECK();
    kc_step<kCubeRZ0><<<1, kBlock, 0, s>>>(g->parts, nc, g->st, mg, nullptr);
    LAUNCH_CHECK();
    // Each iteration's last step leaves the stop flag in host memory; the host reads iteration
    // k's flag while iteration k + 1 is queued.
    const int max_iter = (int)ctx->opts.max_iter;
    double *po = g->p0, *pn = g->p1;
    for (int it = 0; it < max_iter; it++) {
        const int par = it & 1;
        kc_dir<<<nb, kBlock, 0, s>>>(l, mg, side, g->r, z32, g->diag, k, cols, po, pn, g->ap,
                                     g->tmp, g->st);
        LAUNCH_CHECK();
        kc_chunks<<<bc, kRowBlock, 0, s>>>(l, g->tmp, nullptr, 0.0, g->parts, g->st);
        LAUNCH_CHECK();
        std::swap(po, pn);
        kc_step<kCubePAP><<<1, kBlock, 0, s>>>(g->parts, nc, g->st, mg, nullptr);
        LAUNCH_CHECK();
        kc_update<<<nb, kBlock, 0, s>>>(l, po, g->ap, g->diag, g->x, g->r, g->tmp,
                                        mg ? nullptr : g->b, g->st);
        LAUNCH_CHECK();
        if (mg) {
            // r.r alone (the multigrid's rz comes from the cycle's z).
            kc_chunks<<<bc, kRowBlock, 0, s>>>(l, g->tmp, nullptr, 0.0, g->parts, g->st);
            LAUNCH_CHECK();
            kc_step<kCubeRR><<<1, kBlock, 0, s>>>(g->parts, nc, g->st, mg, nullptr);
            LAUNCH_CHECK();
            z32 = cube_precondition(ctx, g, g->r);
            kc_dot<kDotF32><<<bc, kRowBlock, 0, s>>>(l, g->r, nullptr, z32, nullptr, g->parts,
                                                     g->st);
            LAUNCH_CHECK();
            kc_step<kCubeRZ><<<1, kBlock, 0, s>>>(g->parts, nc, g->st, mg, ctx->dev_flag + par);
        } else {
            kc_chunks<<<bc2, kRowBlock, 0, s>>>(l, g->tmp, g->b, 0.0, g->parts, g->st);
            LAUNCH_CHECK();
            kc_step<kCubeRR><<<1, kBlock, 0, s>>>(g->parts, nc, g->st, mg, ctx->dev_flag + par);
        }
        LAUNCH_CHECK();
        HIPCHK(hipEventRecord(ctx->flag_ev[par], s));
        if (it >= 1) {
            HIPCHK(hipEventSynchronize(ctx->flag_ev[par ^ 1]));
            if (ctx->host_flag[par ^ 1]) break;
        }
    }
    // The put-back (CD:286-301) and the answer into h.
    kc_apply<<<nb, kBlock, 0, s>>>(l, side, g->diag, k, cols, g->x, g->ap, nullptr);
    LAUNCH_CHECK();
    kc_putback_sums<<<bc2, kRowBlock, 0, s>>>(l, g->rhs, g->ap, g->s, g->parts);
    LAUNCH_CHECK();
    kc_step<kCubePutback><<<1, kBlock, 0, s>>>(g->parts, nc, g->st, mg, nullptr);
    LAUNCH_CHECK();
    kc_putback<<<nb, kBlock, 0, s>>>(l, cells, g->s, g->mu, g->x, h, g->st);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(g->h_st, g->st, sizeof(CubeState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return (size_t)g->h_st->its;
}

}  // namespace dtd
