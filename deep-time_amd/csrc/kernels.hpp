// Device kernels of the diffusion solve (gfx950, wave64).  Included once, by dtdiff.hip (through
// context.hpp and batch.hpp as well).
//
// Reference files (relative to the deep-time repository root):
//   D  crates/deeptime-core/src/surface/diffusion.rs
//   MG crates/deeptime-core/src/surface/diffusion/multigrid.rs
//
// Every kernel does each cell's arithmetic in the reference's order (the build passes
// -ffp-contract=off, so no multiply-add is fused and every operation rounds once, as on the
// CPU), so a kernel's per-cell results equal the reference's bit for bit.  Sums are taken per
// block in a fixed shuffle/LDS order and the block partials are added in ascending block order by
// one block: run to run they are identical, though not in the reference's strip order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace dtd {

constexpr int kBlock = 256;       // threads per block of the cell-parallel kernels
constexpr int kReduceBlock = 1024;  // threads of the single-block reductions

// Conjugate-gradient scalars, resident on the device so an iteration needs no host round trip.
struct CgState {
    double rr;      // r.r of the current residual
    double rz;      // r.z of the current residual and its correction
    double pap;     // p.A p of the current direction
    double alpha;   // step of the last update (the pending x += alpha p)
    double beta;    // rz_new / rz
    double tol2;    // tol_m^2 * n
    double rr0;     // r.r of the first residual
    double lost;    // put-back: sum (b - A x)
    double weight;  // put-back: sum (diag - 1)
    double bb;      // b.b of the right-hand side
    double rel2;    // > 0: a relative target, tol2 = rel2 * b.b (set by the first residual)
    int have_beta;  // an iteration has formed beta (the first direction is z itself)
    int have_pending;  // an update has set alpha
    int done;       // the loop has stopped (1 converged, 2 p.A p, 3 r.z, 4 cap)
    int its;        // iterations started
    int max_iter;
    int skip;       // a batched solve leaves this grid alone (its answer is final)
    int pad[2];
    // This solve's kd (device) and dt, dx^2: the direction pass forms the operator's faces from
    // them (set per solve, so a captured iteration stays valid for new fields of the shape).
    const double *kd;
    double dt, s2;
};

// ---- reductions ----------------------------------------------------------------------------

// Sum of v over a block of BS threads in a fixed order (shuffles within each wave, then the
// wave sums in ascending order); the result is valid in thread 0.
template <int BS>
__device__ inline double block_sum(double v, double *lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        s = lds[0];
        for (int k = 1; k < BS / 64; k++) s += lds[k];
    }
    __syncthreads();
    return s;
}

// Sum of parts[k * stride + slot] over k < n by one block, in a fixed order: thread t adds
// k = t, t + 1024, ... in turn (the loads unrolled so several are in flight), then block_sum.
__device__ inline double reduce_parts(const double *parts, int n, int stride, int slot,
                                      double *lds) {
    double v = 0.0;
    int k = threadIdx.x;
    for (; k + 3 * kReduceBlock < n; k += 4 * kReduceBlock) {
        const double a = parts[(size_t)k * stride + slot];
        const double b = parts[(size_t)(k + kReduceBlock) * stride + slot];
        const double c = parts[(size_t)(k + 2 * kReduceBlock) * stride + slot];
        const double d = parts[(size_t)(k + 3 * kReduceBlock) * stride + slot];
        v += a;
        v += b;
        v += c;
        v += d;
    }
    for (; k < n; k += kReduceBlock) v += parts[(size_t)k * stride + slot];
    return block_sum<kReduceBlock>(v, lds);
}

// Cells per thread of the conjugate-gradient passes: a block covers kCgCells consecutive cells,
// thread t taking t, t + 256, t + 512 and t + 768.
constexpr int kCgPer = 4;
constexpr int kCgCells = kBlock * kCgPer;

// ---- the f64 operator (D:532-663) ----------------------------------------------------------

struct Grid {
    int w, h, wrap;  // wrap: wrap_x && w > 1
    int n;
};

// Face between cells a and b: 0.5 * (K_a + K_b) * dt / dx^2 (D:556-572).
__device__ inline double face(double ka, double kb, double dt, double s2) {
    return 0.5 * (ka + kb) * dt / s2;
}

// q / b correctly rounded from z = 1 / b (itself correctly rounded), without a division: y0 = q z
// is within 1.5 ulp of q / b, one correction y1 = y0 + (q - b y0) z makes it faithful, and a second
// gives the rounded quotient exactly (Markstein's theorem: Muller et al., Handbook of
// Floating-Point Arithmetic, 2010, section 5.3); each remainder q - b y is one fma.  The theorem
// asks for no underflow: a quotient below 2^-960 (or q = 0) is taken from the division itself.
__device__ __forceinline__ double div_by(double q, double b, double z) {
    double y = q * z;
    double r = fma(-b, y, q);
    y = fma(r, z, y);
    r = fma(-b, y, q);
    y = fma(r, z, y);
    if (__builtin_expect(fabs(y) < 0x1p-960, 0)) y = q / b;
    return y;
}

// `face` with the division by dx^2 as div_by (rs2 = 1 / dx^2): the same bits.  P2: dx^2 is a
// power of two, so rs2 is exact and q rs2 is the correctly rounded q / dx^2 (one rounding of the
// same real number), with no correction steps.
template <int P2 = 0>
__device__ __forceinline__ double face_rz(double ka, double kb, double dt, double s2, double rs2) {
    if constexpr (P2)
        return 0.5 * (ka + kb) * dt * rs2;
    else
        return div_by(0.5 * (ka + kb) * dt, s2, rs2);
}

// dx^2 is a normal power of two (its reciprocal is then exact).
__host__ __device__ inline bool pow2_exact(double s2) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, s2);
    const unsigned e = (unsigned)(b >> 52) & 0x7ffu;
    return (b >> 63) == 0 && (b & ((1ull << 52) - 1)) == 0 && e >= 1 && e <= 2046;
}

// A x at one cell of an f64 lane in the order of `Operator::apply_row` (D:609-649), from its own
// and its neighbours' values; ew is the west neighbour's east face, sn the north neighbour's
// south face.  `inner`: every term present, no wrap (most cells).
__device__ inline double lane_apply64(double xc, double xe, double xw, double xs, double xn,
                                      double e, double ew, double s, double sn, int gi, int gj,
                                      int w, int h, int wrap, bool inner) {
    if (inner) {
        const double d = 1.0 + e + ew + s + sn;
        double o = d * xc;
        o -= e * xe;
        o -= ew * xw;
        o -= s * xs;
        o -= sn * xn;
        return o;
    }
    const double wv = (gi > 0 || wrap) ? ew : 0.0;
    const double nv = gj > 0 ? sn : 0.0;
    const double d = 1.0 + e + wv + s + nv;
    double o = d * xc;
    if (gi + 1 < w || wrap) o -= e * xe;
    if (gi > 0 || wrap) o -= wv * xw;
    if (gj + 1 < h) o -= s * xs;
    if (gj > 0) o -= nv * xn;
    return o;
}

// The diagonal of one cell of an f64 lane (the sum Operator::new takes).
__device__ inline double lane_diag64(double e, double ew, double s, double sn, int gi, int gj,
                                     int wrap) {
    const double wv = (gi > 0 || wrap) ? ew : 0.0;
    const double nv = gj > 0 ? sn : 0.0;
    return 1.0 + e + wv + s + nv;
}

// A cycle level's operator record: (east, south, diag) of a cell side by side.  omega / diag, the
// smoother's weight, is one correctly rounded division of diag, formed where it is used: the same
// bits as a stored copy, without its 4 bytes a cell of traffic.
constexpr int kRec = 3;

// `Operator::new` (D:532-603) with `Level::finest` (MG:96-114) in one pass.  The f64 couplings
// go out packed, es[c] = (east, south); the diagonal ((((1 + e) + w) + s) + n) takes the west and
// north faces from the same expression as the neighbours' own east and south, and is never
// stored (it is re-formed the same way wherever it is used).  The finest level of the cycle goes
// out as one record a cell, (east, south, diag): in f32 (coef32, the reference's cycle) or in f64
// (coef64, the f64 cycle).
__device__ __forceinline__ void operator_faces(const Grid &g, const double *__restrict__ kd,
                                               double dt, double s2, int c, double &e, double &s,
                                               double &d) {
    const int j = c / g.w, i = c - j * g.w, c0 = j * g.w;
    const double k = kd[c];
    if (i + 1 < g.w)
        e = face(k, kd[c + 1], dt, s2);
    else if (g.wrap)
        e = face(k, kd[c0], dt, s2);
    else
        e = 0.0;
    s = j + 1 < g.h ? face(k, kd[c + g.w], dt, s2) : 0.0;
    double wv;
    if (i > 0)
        wv = face(kd[c - 1], k, dt, s2);
    else if (g.wrap)
        wv = face(kd[c0 + g.w - 1], kd[c0], dt, s2);
    else
        wv = 0.0;
    const double nv = j > 0 ? face(kd[c - g.w], k, dt, s2) : 0.0;
    d = 1.0 + e + wv + s + nv;
}

__device__ __forceinline__ void operator_cell(const Grid &g, const double *__restrict__ kd,
                                              double dt, double s2, double2 *__restrict__ es,
                                              float *__restrict__ coef32,
                                              double *__restrict__ coef64, int c) {
    double e, s, d;
    operator_faces(g, kd, dt, s2, c, e, s, d);
    if (es) es[c] = make_double2(e, s);
    if (coef32) {
        float *q = coef32 + kRec * (size_t)c;
        q[0] = (float)e, q[1] = (float)s, q[2] = (float)d;
    }
    if (coef64) {
        double *q = coef64 + kRec * (size_t)c;
        q[0] = e, q[1] = s, q[2] = d;
    }
}

// The block's kRec * kBlock record values gathered in LDS, then written out as whole 16-byte
// words (the records of a full block start 16-byte aligned).
template <typename T>
__device__ __forceinline__ void store_records(T *__restrict__ out, const T *buf, int cells) {
    constexpr int per = 16 / sizeof(T);
    if (cells == kBlock) {
        using W = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
        for (int q = threadIdx.x; q < kRec * kBlock / per; q += kBlock)
            reinterpret_cast<W *>(out)[q] = reinterpret_cast<const W *>(buf)[q];
    } else {
        for (int q = threadIdx.x; q < kRec * cells; q += kBlock) out[q] = buf[q];
    }
}

__global__ __launch_bounds__(kBlock) void k_operator(Grid g, const double *__restrict__ kd,
                                                     double dt, double s2,
                                                     double2 *__restrict__ es,
                                                     float *__restrict__ coef32,
                                                     double *__restrict__ coef64) {
    __shared__ double buf64[kRec * kBlock];
    float *buf32 = reinterpret_cast<float *>(buf64);
    const int c0 = blockIdx.x * kBlock, t = threadIdx.x, c = c0 + t;
    if (c < g.n) {
        double e, s, d;
        operator_faces(g, kd, dt, s2, c, e, s, d);
        if (es) es[c] = make_double2(e, s);
        if (coef32) buf32[kRec * t] = (float)e, buf32[kRec * t + 1] = (float)s,
                    buf32[kRec * t + 2] = (float)d;
        if (coef64) buf64[kRec * t] = e, buf64[kRec * t + 1] = s, buf64[kRec * t + 2] = d;
    }
    if (!coef32 && !coef64) return;
    __syncthreads();
    const int cells = min(kBlock, g.n - c0);
    if (coef32) store_records<float>(coef32 + kRec * (size_t)c0, buf32, cells);
    if (coef64) store_records<double>(coef64 + kRec * (size_t)c0, buf64, cells);
}

// The f64 diagonal of cell c from the packed faces (the sum Operator::new takes).
__device__ inline double diag_es(const Grid &g, const double2 *es, int c, int i, int j) {
    const int c0 = j * g.w;
    const double2 me = es[c];
    const double wv = i > 0 ? es[c - 1].x : (g.wrap ? es[c0 + g.w - 1].x : 0.0);
    const double nv = j > 0 ? es[c - g.w].y : 0.0;
    return 1.0 + me.x + wv + me.y + nv;
}

// One cell of y = A x in the order of `Operator::apply_row` (D:609-649): diagonal, east (the
// wrapped one for the last cell), west (the wrapped one for the first), south, north.
template <class X>
__device__ inline double apply_es(const Grid &g, const double2 *es, int c, int i, int j, X x) {
    const int c0 = j * g.w;
    const double2 me = es[c];
    const double e = me.x, s = me.y;
    const double wv = i > 0 ? es[c - 1].x : (g.wrap ? es[c0 + g.w - 1].x : 0.0);
    const double nv = j > 0 ? es[c - g.w].y : 0.0;
    const double d = 1.0 + e + wv + s + nv;
    double o = d * x(c);
    if (i + 1 < g.w)
        o -= e * x(c + 1);
    else if (g.wrap)
        o -= e * x(c0);
    if (i > 0)
        o -= wv * x(c - 1);
    else if (g.wrap)
        o -= wv * x(c0 + g.w - 1);
    if (j + 1 < g.h) o -= s * x(c + g.w);
    if (j > 0) o -= nv * x(c - g.w);
    return o;
}

// y = A x (`Operator::apply`, D:652-663).
__global__ void k_apply(Grid g, const double2 *__restrict__ es, const double *__restrict__ x,
                        double *__restrict__ y) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const int j = c / g.w, i = c - j * g.w;
    y[c] = apply_es(g, es, c, i, j, [&](int k) { return x[k]; });
}

// east, south and diag of the f64 operator, unpacked (for dtd_operator).
__global__ void k_unpack_es(Grid g, const double2 *__restrict__ es, double *__restrict__ east,
                            double *__restrict__ south, double *__restrict__ diag) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const int j = c / g.w, i = c - j * g.w;
    east[c] = es[c].x;
    south[c] = es[c].y;
    diag[c] = diag_es(g, es, c, i, j);
}

// Second pass of an iteration (MG:703-723): r -= alpha A p, its copy in the cycle's precision,
// per block r*r.
template <typename T>
__device__ __forceinline__ double update_block(int n, const double *__restrict__ ap,
                                               double *__restrict__ r, T *__restrict__ rc,
                                               double alpha, int blk) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < kCgPer; q++) {
        const int c = blk * kCgCells + q * kBlock + threadIdx.x;
        if (c < n) {
            const double rv = r[c] - alpha * ap[c];
            r[c] = rv;
            if (rc) rc[c] = (T)rv;
            acc += rv * rv;
        }
    }
    return acc;
}

template <typename T>
__global__ void k_update(int n, const double *__restrict__ ap, double *__restrict__ r,
                         T *__restrict__ rc, double *__restrict__ parts,
                         const CgState *__restrict__ st) {
    __shared__ double lds[kBlock / 64];
    if (st->done) return;
    const double acc = update_block<T>(n, ap, r, rc, st->alpha, blockIdx.x);
    const double s = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// The scalar steps between the passes, one block each (MG:624-730).
enum Step { kStepRR0 = 0, kStepPAP = 1, kStepRR = 2, kStepRZ = 3, kStepRZ0 = 4 };

// The step itself, from the pass's sum v (and b.b for the first residual); one thread.
template <int STEP>
__device__ inline void cg_step_apply(double v, double bb, CgState *st) {
    if (STEP == kStepRR0) {
        // A relative target sets tol2 = rel^2 b.b here; then stop before the loop when the
        // first guess already meets the tolerance (MG:624-626).
        st->bb = bb;
        if (st->rel2 > 0.0) st->tol2 = st->rel2 * bb;
        st->rr0 = v;
        st->rr = v;
        st->have_beta = 0;
        st->have_pending = 0;
        st->its = 0;
        st->done = v > st->tol2 ? 0 : 1;
    } else if (STEP == kStepRZ0) {
        st->rz = v;
    } else if (STEP == kStepPAP) {
        // iterations += 1 and the p.A p check (MG:631, 697-699); alpha = rz / pAp (MG:700-702).
        st->its += 1;
        st->pap = v;
        if (v <= 0.0 || !isfinite(v)) {
            st->done = 2;
            return;
        }
        st->alpha = st->rz / v;
        st->have_pending = 1;
    } else if (STEP == kStepRR) {
        st->rr = v;
        if (v <= st->tol2) st->done = 1;  // MG:724-726
    } else {
        // rz_new, the break when it or the old rz is unusable, beta (MG:727-733); the cap of
        // the loop (MG:630).
        if (!(isfinite(v) && st->rz > 0.0)) {
            st->done = 3;
            return;
        }
        st->beta = v / st->rz;
        st->have_beta = 1;
        st->rz = v;
        if (st->its >= st->max_iter) st->done = 4;
    }
}

// The step's reduction over one block: the first residual's pass leaves (r.r, b.b) per block,
// the others one sum.
template <int STEP>
__device__ inline void cg_step_block(const double *__restrict__ parts, int nparts, CgState *st,
                                     double *lds) {
    if (STEP != kStepRR0 && st->done) return;
    const int stride = STEP == kStepRR0 ? 2 : 1;
    const double v = reduce_parts(parts, nparts, stride, 0, lds);
    double bb = 0.0;
    if (STEP == kStepRR0) bb = reduce_parts(parts, nparts, 2, 1, lds);
    if (threadIdx.x == 0) cg_step_apply<STEP>(v, bb, st);
}

// With flag given (host memory the device can write), the step leaves the stop flag there too:
// the host reads it once the iteration's event has completed, with no copy in the stream.
template <int STEP>
__global__ void k_cg_step(const double *__restrict__ parts, int nparts, CgState *st,
                          int *flag = nullptr) {
    __shared__ double lds[kReduceBlock / 64];
    cg_step_block<STEP>(parts, nparts, st, lds);
    if (flag && threadIdx.x == 0) *reinterpret_cast<volatile int *>(flag) = st->done;
}

// The put-back's two sums (MG:737-755), then its step k = lost / weight.
__global__ void k_putback_step(const double *__restrict__ parts, int nparts, CgState *st) {
    __shared__ double lds[kReduceBlock / 64];
    const double l = reduce_parts(parts, nparts, 2, 0, lds);
    const double wg = reduce_parts(parts, nparts, 2, 1, lds);
    if (threadIdx.x == 0) {
        st->lost = l;
        st->weight = wg;
    }
}

// (sum of slot 0, sum of slot 1) of pairs of partial sums.
__global__ void k_pair_sum(const double *__restrict__ parts, int nparts, double *out) {
    __shared__ double lds[kReduceBlock / 64];
    const double a = reduce_parts(parts, nparts, 2, 0, lds);
    const double b = reduce_parts(parts, nparts, 2, 1, lds);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// ---- the V-cycle's levels (MG:73-523) -------------------------------------------------------

// A level's operator record is kRec values a cell, (east, south, diag), side by side.
template <typename T>
struct Lvl {
    int w, h, wrap, n;
    T omega;  // the smoother's damping
    T *coef;  // kRec values a cell
    T *mass;  // the identity's share of the diagonal; nullptr on the finest level (1 a cell)
    T *r, *z[2];
    // The same records rounded to bf16 (8 bytes a cell), which the fused smoothing kernels read
    // instead when given (the bf16-storage variant of the cycle); nullptr otherwise.
    unsigned short *coefh;
    // Band heights of the level's fused pre- and post-smoothing launches chosen by the tuner
    // (dtd_tune); 0: the occupancy rule.
    int th_pre, th_post;
    // Options of the cycle with no reference counterpart.  bil: this level restricts to and is
    // corrected from the next by the bilinear transfers instead of 2 x 2 aggregation.  chol (the
    // coarsest level only): its Cholesky factor in f64 (n x n, lower triangle), for the exact
    // solve that replaces the coarsest level's sweeps.
    int bil;
    double *chol;
};

// f32 -> bf16 rounded to nearest, ties to even (a NaN stays a quiet NaN); bf16 -> f32 is exact.
__host__ __device__ inline unsigned short bf16_of(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float float_of_bf16(unsigned b) { return __uint_as_float(b << 16); }

// coefh = the level's records with omega / diag, rounded to bf16: 4 per cell.
__global__ void k_pack_bf16(int n, const float *__restrict__ coef, float omega,
                            unsigned short *__restrict__ coefh) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const float *q = coef + kRec * (size_t)c;
    const unsigned lo = bf16_of(q[0]) | ((unsigned)bf16_of(q[1]) << 16);
    const unsigned hi = bf16_of(q[2]) | ((unsigned)bf16_of(omega / q[2]) << 16);
    *reinterpret_cast<uint2 *>(coefh + 4 * (size_t)c) = make_uint2(lo, hi);
}

template <typename T> __device__ inline T c_east(const Lvl<T> &L, int c) { return L.coef[kRec * (size_t)c]; }
template <typename T> __device__ inline T c_south(const Lvl<T> &L, int c) { return L.coef[kRec * (size_t)c + 1]; }
template <typename T> __device__ inline T c_diag(const Lvl<T> &L, int c) { return L.coef[kRec * (size_t)c + 2]; }
template <typename T> __device__ inline T c_wd(const Lvl<T> &L, int c) { return L.omega / c_diag(L, c); }

// `Level::finest` (MG:96-114) from a given f64 operator: the f32 record.
__global__ void k_finest32(int n, const double *__restrict__ east,
                           const double *__restrict__ south, const double *__restrict__ diag,
                           float *__restrict__ coef) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    float *q = coef + kRec * (size_t)c;
    q[0] = (float)east[c], q[1] = (float)south[c], q[2] = (float)diag[c];
}

// The f64 cycle's finest level from a given f64 operator.
__global__ void k_finest64(int n, const double *__restrict__ east,
                           const double *__restrict__ south, const double *__restrict__ diag,
                           double *__restrict__ coef) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    double *q = coef + kRec * (size_t)c;
    q[0] = east[c], q[1] = south[c], q[2] = diag[c];
}

// `Level::coarsen` (MG:117-199) for coarse cell (ic, jc): the mass of the 2x2 block (rows then
// columns), the coarse east face = scale * the fine east faces leaving the block's last column
// (rows in order; none on a ring of one block), the coarse south face likewise over the block's
// last row's columns.
template <typename T>
__device__ inline T coarse_mass(const Lvl<T> &f, int ic, int jc) {
    const int r1 = min(2 * jc + 2, f.h), c1 = min(2 * ic + 2, f.w);
    T m = (T)0;
    for (int j = 2 * jc; j < r1; j++)
        for (int i = 2 * ic; i < c1; i++) m += f.mass ? f.mass[j * f.w + i] : (T)1;
    return m;
}
// east(c), south(c): the finer level's east and south faces of cell c.
template <typename T, class E>
__device__ inline T coarse_east(const Lvl<T> &f, int wc, int cwrap, T scale, int ic, int jc,
                                E east) {
    if (!(ic + 1 < wc || cwrap)) return (T)0;
    const int r1 = min(2 * jc + 2, f.h), last_col = min(2 * ic + 2, f.w) - 1;
    T e = (T)0;
    for (int j = 2 * jc; j < r1; j++) e += east(j * f.w + last_col);
    return scale * e;
}
template <typename T, class S>
__device__ inline T coarse_south(const Lvl<T> &f, int hc, T scale, int ic, int jc, S south) {
    if (!(jc + 1 < hc)) return (T)0;
    const int last_row = min(2 * jc + 2, f.h) - 1, c1 = min(2 * ic + 2, f.w);
    T s = (T)0;
    for (int i = 2 * ic; i < c1; i++) s += south(last_row * f.w + i);
    return scale * s;
}

// One coarse level from the finer one: mass, faces, and the diagonal mass + e + w + s + n
// (MG:189-198); the west and north faces are formed by the same sums as the
// neighbours' own, so every value equals the reference's.
template <typename T, class E, class S>
__device__ __forceinline__ void coarsen_cell(const Lvl<T> &f, const Lvl<T> &c, T scale, int k,
                                             E east, S south) {
    const int jc = k / c.w, ic = k - jc * c.w;
    const T m = coarse_mass(f, ic, jc);
    const T e = coarse_east(f, c.w, c.wrap, scale, ic, jc, east);
    const T s = coarse_south(f, c.h, scale, ic, jc, south);
    const T west = ic > 0 ? coarse_east(f, c.w, c.wrap, scale, ic - 1, jc, east)
                          : (c.wrap ? coarse_east(f, c.w, c.wrap, scale, c.w - 1, jc, east) : (T)0);
    const T north = jc > 0 ? coarse_south(f, c.h, scale, ic, jc - 1, south) : (T)0;
    const T d = m + e + west + s + north;
    c.mass[k] = m;
    T *q = c.coef + kRec * (size_t)k;
    q[0] = e, q[1] = s, q[2] = d;
}
template <typename T>
__device__ __forceinline__ void coarsen_cell(const Lvl<T> &f, const Lvl<T> &c, T scale, int k) {
    coarsen_cell(f, c, scale, k, [&](int q) { return c_east(f, q); },
                 [&](int q) { return c_south(f, q); });
}

template <typename T>
__global__ void k_coarsen(Lvl<T> f, Lvl<T> c, T scale) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < c.n) coarsen_cell(f, c, scale, k);
}

// Level 1 of the f32 cycle straight from kd, with no finest records in between (the fused finest
// smoothers form theirs from kd): each fine face is the f32 rounding of the f64 face
// Operator::new forms (D:556-572), exactly the value `Level::finest` would store (MG:96-114),
// summed as from stored records (coarse_east / coarse_south's sums).  f is the finest level (its
// sizes; mass 1).  A wave takes 64 adjacent coarse columns and walks kCoarsenRows coarse rows:
// each coarse cell forms its own east and south faces; its west face is the lane before's east
// face (the wave's first lane and the first coarse column form theirs), its north face the south
// face of the row before (the walk's first row forms it).  The same sums, so the same bits, with
// half the faces formed.
constexpr int kCoarsenRows = 8;

template <typename T>
__device__ inline T from_west(T v);  // lane - 1's value (a DPP move, defined below)

__device__ __forceinline__ void coarsen_kd_walk(const Lvl<float> &f, const Lvl<float> &c,
                                                float scale, const double *__restrict__ kd,
                                                double dt, double s2, int ic, int jc0) {
    const int w = f.w, h = f.h, lane = threadIdx.x & 63;
    const bool ok = ic < c.w;
    const int icc = ok ? ic : c.w - 1;  // past the last column: in bounds, nothing stored
    auto east_f = [&](int j, int i) {
        const size_t q = (size_t)j * w + i;
        if (i + 1 < w) return (float)face(kd[q], kd[q + 1], dt, s2);
        return f.wrap ? (float)face(kd[q], kd[(size_t)j * w], dt, s2) : 0.0f;
    };
    auto south_f = [&](int j, int i) {
        const size_t q = (size_t)j * w + i;
        return j + 1 < h ? (float)face(kd[q], kd[q + w], dt, s2) : 0.0f;
    };
    auto ce = [&](int x, int jc) {
        if (!(x + 1 < c.w || c.wrap)) return 0.0f;
        const int r1 = min(2 * jc + 2, h), last_col = min(2 * x + 2, w) - 1;
        float e = 0.0f;
        for (int j = 2 * jc; j < r1; j++) e += east_f(j, last_col);
        return scale * e;
    };
    auto cs = [&](int x, int jc) {
        if (!(jc + 1 < c.h)) return 0.0f;
        const int last_row = min(2 * jc + 2, h) - 1, c1 = min(2 * x + 2, w);
        float v = 0.0f;
        for (int i = 2 * x; i < c1; i++) v += south_f(last_row, i);
        return scale * v;
    };
    const int jc1 = min(jc0 + kCoarsenRows, c.h);
    float north = jc0 > 0 ? cs(icc, jc0 - 1) : 0.0f;
    for (int jc = jc0; jc < jc1; jc++) {
        const float e = ce(icc, jc), s = cs(icc, jc);
        float west = from_west(e);
        if (lane == 0 || icc == 0) west = icc > 0 ? ce(icc - 1, jc) : (c.wrap ? ce(c.w - 1, jc) : 0.0f);
        const float m = coarse_mass(f, icc, jc);
        const float d = m + e + west + s + north;
        if (ok) {
            const size_t k = (size_t)jc * c.w + ic;
            c.mass[k] = m;
            float *q = c.coef + kRec * k;
            q[0] = e, q[1] = s, q[2] = d;
        }
        north = s;
    }
}

// Grid: ceil(c.w / 64) x ceil(c.h / (4 kCoarsenRows)) blocks of kBlock threads.
__global__ __launch_bounds__(kBlock) void k_coarsen_kd(Lvl<float> f, Lvl<float> c, float scale,
                                                       const double *__restrict__ kd, double dt,
                                                       double s2) {
    const int ic = blockIdx.x * 64 + (threadIdx.x & 63);
    const int jc0 = (blockIdx.y * (kBlock / 64) + (int)(threadIdx.x >> 6)) * kCoarsenRows;
    if (jc0 >= c.h) return;  // the whole wave
    coarsen_kd_walk(f, c, scale, kd, dt, s2, ic, jc0);
}

// One cell of A x on a level in the order of `Level::apply_row` (MG:203-240): diagonal, east
// and west (non-wrapped), the wrapped east term of the last cell, the wrapped west term of the
// first, south, north.
template <typename T, class X>
__device__ inline T lvl_apply(const Lvl<T> &L, int c, int i, int j, X x) {
    const int c0 = j * L.w;
    T o = c_diag(L, c) * x(c);
    if (i + 1 < L.w) o -= c_east(L, c) * x(c + 1);
    if (i > 0) o -= c_east(L, c - 1) * x(c - 1);
    if (L.wrap) {
        if (i == L.w - 1) o -= c_east(L, c0 + L.w - 1) * x(c0);
        if (i == 0) o -= c_east(L, c0 + L.w - 1) * x(c0 + L.w - 1);
    }
    if (j + 1 < L.h) o -= c_south(L, c) * x(c + L.w);
    if (j > 0) o -= c_south(L, c - L.w) * x(c - L.w);
    return o;
}

// `Level::first_sweep` (MG:243-256): z = wd * r.
template <typename T>
__global__ void k_first(Lvl<T> L, T *__restrict__ zout, const CgState *__restrict__ st) {
    if (st && st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= L.n) return;
    zout[c] = c_wd(L, c) * L.r[c];
}

// One cell of `Level::sweep` (MG:260-329): z' + wd (r - A z') with z' = zin + zc[j/2][i/2] when
// zc is given.
template <typename T>
__device__ inline T sweep_cell(const Lvl<T> &L, const T *__restrict__ zin,
                               const T *__restrict__ zc, int wc, int c) {
    const int j = c / L.w, i = c - j * L.w;
    auto zp = [&](int k) -> T {
        if (!zc) return zin[k];
        const int jj = k / L.w, ii = k - jj * L.w;
        return zin[k] + zc[(jj >> 1) * wc + (ii >> 1)];
    };
    const T ax = lvl_apply(L, c, i, j, zp);
    return zp(c) + c_wd(L, c) * (L.r[c] - ax);
}

// ---- the bilinear transfers (an option of the cycle; no reference counterpart) ---------------
//
// Along one axis, fine index i lies in coarse cell i / 2 and nearest to coarse cell i / 2 + 1 (i
// odd) or i / 2 - 1 (i even); that neighbour wraps on a periodic axis of nc cells and otherwise
// falls back to the cell itself.  P takes 3/4 of the first and 1/4 of the second along each axis;
// the restriction is P's transpose.
__device__ inline void bil_axis(int i, int nc, int wrap, int &c, int &nb) {
    c = i >> 1;
    int m = (i & 1) ? c + 1 : c - 1;
    if (wrap)
        m = (m + nc) % nc;
    else if (m < 0 || m >= nc)
        m = c;
    nb = m;
}
template <typename T>
__device__ inline T bil_weight(int i, int C, int nc, int wrap) {
    int c, nb;
    bil_axis(i, nc, wrap, c, nb);
    return (c == C ? (T)0.75 : (T)0) + (nb == C ? (T)0.25 : (T)0);
}
// (P zc) at fine cell (i, j) of a level whose next coarser level is C:
// 3/4 (3/4 zc[cy][cx] + 1/4 zc[cy][nx]) + 1/4 (3/4 zc[ny][cx] + 1/4 zc[ny][nx]).
template <typename T>
__device__ inline T bil_at(const T *__restrict__ zc, int wc, int hc, int cwrap, int i, int j) {
    int cx, nx, cy, ny;
    bil_axis(i, wc, cwrap, cx, nx);
    bil_axis(j, hc, 0, cy, ny);
    const T a = (T)0.75 * zc[cy * wc + cx] + (T)0.25 * zc[cy * wc + nx];
    const T b = (T)0.75 * zc[ny * wc + cx] + (T)0.25 * zc[ny * wc + nx];
    return (T)0.75 * a + (T)0.25 * b;
}
// (P^T (r - A z)) at coarse cell k: over the fine rows 2J-1 .. 2J+2 on the level and the fine
// columns 2I-1 .. 2I+2 (wrapped on a periodic level, each column once), the terms
// (weight_y * weight_x) * (r - A z) with a non-zero weight, added from +0, rows outer.
template <typename T>
__device__ inline T bil_restrict_cell(const Lvl<T> &f, const T *__restrict__ z, const Lvl<T> &c,
                                      int k) {
    const int J = k / c.w, I = k - J * c.w;
    T acc = (T)0;
    for (int t = -1; t <= 2; t++) {
        const int j = 2 * J + t;
        if (j < 0 || j >= f.h) continue;
        const T wy = bil_weight<T>(j, J, c.h, 0);
        if (wy == (T)0) continue;
        int seen[4], ns = 0;
        for (int u = -1; u <= 2; u++) {
            int i = 2 * I + u;
            if (f.wrap)
                i = ((i % f.w) + f.w) % f.w;
            else if (i < 0 || i >= f.w)
                continue;
            bool dup = false;
            for (int q = 0; q < ns; q++) dup |= seen[q] == i;
            if (dup) continue;
            seen[ns++] = i;
            const T wx = bil_weight<T>(i, I, c.w, c.wrap);
            if (wx == (T)0) continue;
            const int cf = j * f.w + i;
            acc += (wy * wx) * (f.r[cf] - lvl_apply(f, cf, i, j, [&](int q) { return z[q]; }));
        }
    }
    return acc;
}

// z += P zc on level L from its next coarser level C, in place (the corrected sweep then starts
// from z).
template <typename T>
__global__ void k_add_prolong(Lvl<T> L, Lvl<T> C, T *__restrict__ z, const CgState *__restrict__ st) {
    if (st && st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= L.n) return;
    const int j = c / L.w, i = c - j * L.w;
    z[c] = z[c] + bil_at(C.z[1], C.w, C.h, C.wrap, i, j);
}

// ---- the exact coarsest solve (an option of the cycle; no reference counterpart) -------------
//
// The coarsest level's matrix in f64, row c holding the terms of lvl_apply in its order from 0,
// factorised A = L L^T in LDS by one wave (row c by lane c: per column k the pivot's square root,
// the column scaled by it, the trailing rows updated), the lower triangle written to chol.
template <typename T>
__global__ __launch_bounds__(64) void k_chol(Lvl<T> L) {
    __shared__ double A[64][64];
    const int c = threadIdx.x, n = L.n;
    if (c < n) {
        for (int q = 0; q < n; q++) A[c][q] = 0.0;
        const int j = c / L.w, i = c - j * L.w, c0 = j * L.w;
        A[c][c] += (double)c_diag(L, c);
        if (i + 1 < L.w) A[c][c + 1] -= (double)c_east(L, c);
        if (i > 0) A[c][c - 1] -= (double)c_east(L, c - 1);
        if (L.wrap) {
            if (i == L.w - 1) A[c][c0] -= (double)c_east(L, c0 + L.w - 1);
            if (i == 0) A[c][c0 + L.w - 1] -= (double)c_east(L, c0 + L.w - 1);
        }
        if (j + 1 < L.h) A[c][c + L.w] -= (double)c_south(L, c);
        if (j > 0) A[c][c - L.w] -= (double)c_south(L, c - L.w);
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
        if (c == k) A[k][k] = sqrt(A[k][k]);
        __syncthreads();
        if (c > k && c < n) A[c][k] = A[c][k] / A[k][k];
        __syncthreads();
        if (c > k && c < n)
            for (int q = k + 1; q <= c; q++) A[c][q] -= A[c][k] * A[q][k];
        __syncthreads();
    }
    if (c < n)
        for (int q = 0; q <= c; q++) L.chol[c * n + q] = A[c][q];
}

// z = (L L^T)^-1 r by the block's first wave alone (lanes = cells; the caller's other waves do
// not take part): r in f64, forward then backward substitution by columns, rounded to T.  ys: 64
// doubles of LDS.  One wave's LDS accesses take effect in program order; the fences keep the
// compiler from moving them across a step.
template <typename T>
__device__ inline void coarsest_exact_wave(const Lvl<T> &L, const T *r, T *zout, double *ys) {
    const int c = threadIdx.x, n = L.n;
    const double *ch = L.chol;
    auto step = [] {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    };
    if (c < n) ys[c] = (double)r[c];
    step();
    for (int k = 0; k < n; k++) {
        if (c == k) ys[k] = ys[k] / ch[k * n + k];
        step();
        if (c > k && c < n) ys[c] -= ch[c * n + k] * ys[k];
        step();
    }
    for (int k = n - 1; k >= 0; k--) {
        if (c == k) ys[k] = ys[k] / ch[k * n + k];
        step();
        if (c < k) ys[c] -= ch[k * n + c] * ys[k];
        step();
    }
    if (c < n) zout[c] = (T)ys[c];
}

// One coarse cell of `Level::restrict_into` (MG:332-356): the sum over its 2 x 2 fine block of
// (r - A z), fine row j0 then j0 + 1, columns ascending, from +0 (the bilinear transpose when the
// level takes that option).
template <typename T>
__device__ inline T restrict_cell(const Lvl<T> &f, const T *__restrict__ z, const Lvl<T> &c,
                                  int k) {
    if (f.bil) return bil_restrict_cell(f, z, c, k);
    const int jc = k / c.w, ic = k - jc * c.w;
    const int r1 = min(2 * jc + 2, f.h), c1 = min(2 * ic + 2, f.w);
    T acc = (T)0;
    for (int j = 2 * jc; j < r1; j++)
        for (int i = 2 * ic; i < c1; i++) {
            const int cf = j * f.w + i;
            acc += f.r[cf] - lvl_apply(f, cf, i, j, [&](int q) { return z[q]; });
        }
    return acc;
}

// `Level::sweep` (MG:260-329): zout = z' + wd (r - A z') with z' = zin + zc[j/2][i/2] when zc is
// given; with dot, per block sum dot * zout.
template <typename T>
__global__ void k_sweep(Lvl<T> L, const T *__restrict__ zin, T *__restrict__ zout,
                        const T *__restrict__ zc, int wc, const double *__restrict__ dot,
                        double *__restrict__ parts, const CgState *__restrict__ st) {
    __shared__ double lds[kBlock / 64];
    if (st && st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double acc = 0.0;
    if (c < L.n) {
        const int j = c / L.w, i = c - j * L.w;
        auto zp = [&](int k) -> T {
            if (!zc) return zin[k];
            const int jj = k / L.w, ii = k - jj * L.w;
            return zin[k] + zc[(jj >> 1) * wc + (ii >> 1)];
        };
        const T ax = lvl_apply(L, c, i, j, zp);
        const T v = zp(c) + c_wd(L, c) * (L.r[c] - ax);
        zout[c] = v;
        if (dot) acc = dot[c] * (double)v;
    }
    if (dot) {
        const double s = block_sum<kBlock>(acc, lds);
        if (threadIdx.x == 0) parts[blockIdx.x] = s;
    }
}

// `Level::restrict_into` (MG:332-356): the coarse r = sum over the block of (r - A z), fine row
// j0 then j0 + 1, columns ascending, from +0.
template <typename T>
__global__ void k_restrict(Lvl<T> f, const T *__restrict__ z, Lvl<T> c,
                           const CgState *__restrict__ st) {
    if (st && st->done) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= c.n) return;
    if (f.bil) {
        c.r[k] = bil_restrict_cell(f, z, c, k);
        return;
    }
    const int jc = k / c.w, ic = k - jc * c.w;
    const int r1 = min(2 * jc + 2, f.h), c1 = min(2 * ic + 2, f.w);
    T acc = (T)0;
    for (int j = 2 * jc; j < r1; j++)
        for (int i = 2 * ic; i < c1; i++) {
            const int cf = j * f.w + i;
            acc += f.r[cf] - lvl_apply(f, cf, i, j, [&](int q) { return z[q]; });
        }
    c.r[k] = acc;
}

// The coarsest level (MG:502-510): z = wd r, then COARSE_SWEEPS - 1 sweeps, in one wave (the
// level has at most 64 cells).  With dot (the coarsest level is the finest), sum dot * z in the
// reference's order: per strip of 8 rows, then the strips in order.
// One wave of 64 threads; zs is 2 x 64 values of LDS; with dot, *part gets the sum.
template <typename T>
__device__ __forceinline__ void coarsest_wave(const Lvl<T> &L, T *__restrict__ zout, int sweeps,
                                              const double *__restrict__ dot, double *part,
                                              T (*zs)[64]) {
    const int c = threadIdx.x;
    const bool live = c < L.n;
    int j = 0, i = 0;
    if (live) {
        j = c / L.w;
        i = c - j * L.w;
        zs[0][c] = c_wd(L, c) * L.r[c];
    }
    __syncthreads();
    int cur = 0;
    for (int k = 1; k < sweeps; k++) {
        if (live) {
            const T *zz = zs[cur];
            const T ax = lvl_apply(L, c, i, j, [&](int q) { return zz[q]; });
            zs[cur ^ 1][c] = zz[c] + c_wd(L, c) * (L.r[c] - ax);
        }
        __syncthreads();
        cur ^= 1;
    }
    if (live) zout[c] = zs[cur][c];
    if (dot && c == 0) {
        double total = -0.0;
        for (int s0 = 0; s0 < L.n; s0 += 8 * L.w) {
            double a = -0.0;
            const int s1 = min(s0 + 8 * L.w, L.n);
            for (int q = s0; q < s1; q++) a += dot[q] * (double)zs[cur][q];
            total += a;
        }
        *part = total;
    }
}

template <typename T>
__global__ void k_coarsest(Lvl<T> L, T *__restrict__ zout, int sweeps,
                           const double *__restrict__ dot, double *__restrict__ parts,
                           const CgState *__restrict__ st) {
    __shared__ T zs[2][64];
    __shared__ double ys[64];
    if (st && st->done) return;
    if (!L.chol) {
        coarsest_wave<T>(L, zout, sweeps, dot, parts, zs);
        return;
    }
    coarsest_exact_wave<T>(L, L.r, zout, ys);
    if (dot && threadIdx.x == 0) {
        // sum dot * z per strip of 8 rows, the strips in order (as the sweeps' dot).
        double total = -0.0;
        for (int s0 = 0; s0 < L.n; s0 += 8 * L.w) {
            double a = -0.0;
            const int s1 = min(s0 + 8 * L.w, L.n);
            for (int q = s0; q < s1; q++) a += dot[q] * (double)(T)ys[q];
            total += a;
        }
        *parts = total;
    }
}

// The bottom of the V-cycle (nu = 2) in one block: from level b0 down, each level's first sweep,
// second sweep and restriction; the coarsest level's 16 sweeps; then back up, each level's
// corrected sweep and last sweep (MG:500-523).  Every phase is one pass of the block over the
// level's cells with the one-sweep kernels' arithmetic, so the result is theirs bit for bit; a
// barrier separates the phases.  The levels live in LDS for the launch (records, r, z[0], z[1]:
// (kRec + 3) values a cell, kBottomLds bytes at most, dynamic shared memory): they are loaded once,
// and only level b0's result goes back, into its z[1], where the launch-per-sweep cycle leaves
// it.  lv holds the hierarchy's levels (device copy); nl is their count.
constexpr int kBottomBlock = 1024;
constexpr int kBottomLevels = 16;
constexpr size_t kBottomLds = 65536;

// One sweep of cell (i, j) of level L from z (z + the prolonged correction, already added, on the
// way up): z + wd (r - A z).
template <typename T>
__device__ inline T bottom_sweep(const Lvl<T> &L, const T *z, int c, int i, int j) {
    const T ax = lvl_apply(L, c, i, j, [&](int k) { return z[k]; });
    return z[c] + c_wd(L, c) * (L.r[c] - ax);
}

// Levels b0 .. nl - 1 of a cycle (every one of at most kBottomLevels, their records, r and z
// at most kBottomLds bytes) by one block of kBottomBlock threads in LDS: the top level's
// residual in, its correction out to lv[b0].z[1].
// row and column of cell c of a level w wide (c < 2^22, w < 2^22: the float quotient is within
// one of the integer one, and the corrections fix it).
__device__ inline int bottom_row(int c, int w, float rw) {
    int j = (int)((float)c * rw);
    if (j * w > c) j--;
    if ((j + 1) * w <= c) j++;
    return j;
}

// n values from global g into LDS l by the block, four loads in flight a thread (16-byte words
// when both ends allow).
template <typename T>
__device__ inline void bottom_copy(const T *__restrict__ g, T *l, int n) {
    const int t = threadIdx.x;
    constexpr int per = 16 / sizeof(T);
    if (n % per == 0 && ((reinterpret_cast<size_t>(g) | reinterpret_cast<size_t>(l)) & 15) == 0) {
        using W = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
        const W *gw = reinterpret_cast<const W *>(g);
        W *lw = reinterpret_cast<W *>(l);
        const int nw = n / per;
        int k = t;
        for (; k + kBottomBlock < nw; k += 2 * kBottomBlock) {
            const W a = gw[k], b = gw[k + kBottomBlock];
            lw[k] = a, lw[k + kBottomBlock] = b;
        }
        for (; k < nw; k += kBottomBlock) lw[k] = gw[k];
        return;
    }
    int k = t;
    for (; k + 3 * kBottomBlock < n; k += 4 * kBottomBlock) {
        const T a = g[k], b = g[k + kBottomBlock], c = g[k + 2 * kBottomBlock],
                d = g[k + 3 * kBottomBlock];
        l[k] = a, l[k + kBottomBlock] = b, l[k + 2 * kBottomBlock] = c, l[k + 3 * kBottomBlock] = d;
    }
    for (; k < n; k += kBottomBlock) l[k] = g[k];
}

// Levels b0 .. nl - 1 of a cycle (every one of at most kBottomLevels, their records, r and z
// at most kBottomLds bytes) by one block of kBottomBlock threads in LDS: the top level's
// residual in, its correction out to lv[b0].z[1].  The restriction first forms each fine cell's
// r - A z (into z[0], free by then) by its own thread, then sums each coarse block's four in
// restrict_cell's order; the coarsest level's neighbours come by lane shuffles.
template <typename T>
__device__ __forceinline__ void bottom_cycle(const Lvl<T> *__restrict__ lv, int b0, int nl) {
    extern __shared__ double bottom_mem[];
    __shared__ Lvl<T> loc[kBottomLevels];
    __shared__ int cells[kBottomLevels];
    const int t = threadIdx.x, nb = nl - b0;
    // The level table: one thread a level reads it, then each level's LDS place follows the
    // sizes of the levels before it.
    Lvl<T> mine{};
    if (t < nb) {
        mine = lv[b0 + t];
        cells[t] = mine.n;
    }
    __syncthreads();
    if (t < nb) {
        size_t off = 0;
        for (int q = 0; q < t; q++) off += (size_t)(kRec + 3) * cells[q];
        T *p = reinterpret_cast<T *>(bottom_mem) + off;
        Lvl<T> L = mine;
        L.coef = p, p += kRec * L.n;
        L.r = p, p += L.n;
        L.z[0] = p, p += L.n;
        L.z[1] = p;
        L.mass = nullptr, L.coefh = nullptr;
        loc[t] = L;
    }
    __syncthreads();
    for (int q = 0; q < nb; q++) bottom_copy<T>(lv[b0 + q].coef, loc[q].coef, kRec * cells[q]);
    bottom_copy<T>(lv[b0].r, loc[0].r, cells[0]);
    __syncthreads();
    for (int q = 0; q + 1 < nb; q++) {
        const Lvl<T> L = loc[q], C = loc[q + 1];
        const float rw = 1.0f / (float)L.w;
        for (int c = t; c < L.n; c += kBottomBlock) L.z[0][c] = c_wd(L, c) * L.r[c];
        __syncthreads();
        for (int c = t; c < L.n; c += kBottomBlock) {
            const int j = bottom_row(c, L.w, rw);
            L.z[1][c] = bottom_sweep(L, L.z[0], c, c - j * L.w, j);
        }
        __syncthreads();
        if (L.bil) {
            for (int k = t; k < C.n; k += kBottomBlock) C.r[k] = restrict_cell(L, L.z[1], C, k);
        } else {
            for (int c = t; c < L.n; c += kBottomBlock) {
                const int j = bottom_row(c, L.w, rw), i = c - j * L.w;
                L.z[0][c] = L.r[c] - lvl_apply(L, c, i, j, [&](int k) { return L.z[1][k]; });
            }
            __syncthreads();
            for (int k = t; k < C.n; k += kBottomBlock) {
                const int jc = bottom_row(k, C.w, 1.0f / (float)C.w), ic = k - jc * C.w;
                const int r1 = min(2 * jc + 2, L.h), c1 = min(2 * ic + 2, L.w);
                T acc = (T)0;
                for (int j = 2 * jc; j < r1; j++)
                    for (int i = 2 * ic; i < c1; i++) acc += L.z[0][j * L.w + i];
                C.r[k] = acc;
            }
        }
        __syncthreads();
    }
    if (lv[nl - 1].chol) {
        // The coarsest level's exact solve (an option), by the first wave.
        const Lvl<T> L = loc[nb - 1];
        __shared__ double ys[64];
        Lvl<T> G = L;
        G.chol = lv[nl - 1].chol;
        if (t < 64) coarsest_exact_wave<T>(G, L.r, nb == 1 ? lv[b0].z[1] : L.z[0], ys);
        __syncthreads();
    } else {
        // The coarsest level (at most 64 cells): z = wd r, then 15 sweeps, by the first wave with
        // its stencil's records in registers and its neighbours' z by lane shuffles.
        const Lvl<T> L = loc[nb - 1];
        if (t < 64) {
            const int c = t;
            const bool live = c < L.n;
            int j = 0, i = 0;
            T d = 0, wd = 0, rr = 0, ce = 0, cw = 0, cs = 0, cn = 0;
            int ke = c, kw = c, ks = c, kn = c;
            bool he = false, hw = false, hs = false, hn = false, hew = false, hww = false;
            if (live) {
                j = c / L.w, i = c - j * L.w;
                const int c0 = j * L.w;
                d = c_diag(L, c), wd = c_wd(L, c), rr = L.r[c];
                he = i + 1 < L.w, hw = i > 0, hs = j + 1 < L.h, hn = j > 0;
                hew = L.wrap && i == L.w - 1, hww = L.wrap && i == 0;
                if (he) ce = c_east(L, c), ke = c + 1;
                if (hw) cw = c_east(L, c - 1), kw = c - 1;
                if (hew) ce = c_east(L, c0 + L.w - 1), ke = c0;
                if (hww) cw = c_east(L, c0 + L.w - 1), kw = c0 + L.w - 1;
                if (hs) cs = c_south(L, c), ks = c + L.w;
                if (hn) cn = c_south(L, c - L.w), kn = c - L.w;
            }
            T z = live ? wd * rr : (T)0;
            for (int k = 1; k < 16; k++) {
                const T zE = __shfl(z, ke), zW = __shfl(z, kw), zS = __shfl(z, ks),
                        zN = __shfl(z, kn);
                // lvl_apply's terms in its order: diag, east, west, the wrapped east / west,
                // south, north.
                T o = d * z;
                if (he) o -= ce * zE;
                if (hw) o -= cw * zW;
                if (hew) o -= ce * zE;
                if (hww) o -= cw * zW;
                if (hs) o -= cs * zS;
                if (hn) o -= cn * zN;
                if (live) z = z + wd * (rr - o);
            }
            if (live) (nb == 1 ? lv[b0].z[1] : L.z[0])[c] = z;
        }
        __syncthreads();
    }
    // Back up: each level's result in its z[0] (the coarsest's too).
    for (int q = nb - 2; q >= 0; q--) {
        const Lvl<T> L = loc[q], C = loc[q + 1];
        const float rw = 1.0f / (float)L.w;
        // z' = z + the coarse correction, once a cell (the sum the corrected sweep takes for
        // every point of its stencil); then the corrected sweep and the last.
        for (int c = t; c < L.n; c += kBottomBlock) {
            const int j = bottom_row(c, L.w, rw), i = c - j * L.w;
            L.z[0][c] = L.z[1][c] + (L.bil ? bil_at(C.z[0], C.w, C.h, C.wrap, i, j)
                                           : C.z[0][(j >> 1) * C.w + (i >> 1)]);
        }
        __syncthreads();
        for (int c = t; c < L.n; c += kBottomBlock) {
            const int j = bottom_row(c, L.w, rw);
            L.z[1][c] = bottom_sweep(L, L.z[0], c, c - j * L.w, j);
        }
        __syncthreads();
        T *out = q > 0 ? L.z[0] : lv[b0].z[1];
        for (int c = t; c < L.n; c += kBottomBlock) {
            const int j = bottom_row(c, L.w, rw);
            out[c] = bottom_sweep(L, L.z[1], c, c - j * L.w, j);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBottomBlock) void k_bottom(const Lvl<T> *__restrict__ lv, int b0,
                                                         int nl, const CgState *__restrict__ st) {
    if (st && st->done) return;
    bottom_cycle<T>(lv, b0, nl);
}

// ---- fused, register-streaming smoothing of a level (nu = 2) -------------------------------
//
// Each wave owns a band of 60 columns (lanes 2..61; lanes 0, 1, 62, 63 hold the halo columns,
// wrapped when the level wraps) and 64 rows, and walks down the band one row per step keeping a
// few rows of every quantity in registers; the east and west neighbours come from the adjacent
// lanes by shuffles.  The pre-smoothing kernel forms, per step, the first sweep of the newest
// row, the second sweep of the row above it and the residual of the row above that, restricted
// to the coarse level in 2 x 2 blocks (`vcycle`, MG:511-514); the post-smoothing kernel forms the
// corrected sweep and the last sweep with its dot the same way (MG:516-521).  Halo lanes and the
// two rows above the band are recomputed by the neighbouring bands, so each cell's arithmetic is
// the unfused kernels', operation for operation, and the results are identical.

// Cells per lane: V consecutive columns (8 bytes a lane: two f32 or one f64 cell).  Halo lanes
// on each side: enough for two columns.
template <int V, int H = (V >= 2 ? 1 : 2)> struct LaneGeom {
    static constexpr int halo = H;                   // halo lanes on each side
    static constexpr int own = (64 - 2 * halo) * V;  // owned columns per wave
};

// A value one lane over, by a DPP wave shift (wave_shr:1 = 0x138: lane i gets lane i - 1's;
// wave_shl:1 = 0x130: lane i gets lane i + 1's) folded into a VALU move, instead of a ds_bpermute
// through the LDS crossbar.  The lane with no source (0, resp. 63) keeps its own value, as
// __shfl_up / __shfl_down give it; 8-byte values move as two 4-byte halves.
template <int CTRL>
__device__ inline int dpp_shift(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL, typename T>
__device__ inline T dpp_move(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte lanes");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp_shift<CTRL>(__builtin_bit_cast(int, v)));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = dpp_shift<CTRL>((int)(b & 0xffffffff)), hi = dpp_shift<CTRL>((int)(b >> 32));
        return __builtin_bit_cast(T, ((long long)hi << 32) | (unsigned)lo);
    }
}
template <typename T>
__device__ inline T from_west(T v) { return dpp_move<0x138>(v); }  // lane - 1's value
template <typename T>
__device__ inline T from_east(T v) { return dpp_move<0x130>(v); }  // lane + 1's value

// A load or store through the global address space: the streamed kernels' arrays are all device
// memory, and a pointer that the compiler cannot place (one read from the state, say) would
// otherwise take a flat access.
template <typename T>
__device__ __forceinline__ T gld(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ void gst(T *p, T v) {
    *(__attribute__((address_space(1))) T *)p = v;
}

// Plain vector types for the wide loads and stores (HIP's float2 & co. are structs, whose copies
// go through a generic pointer).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <typename T, int V>
__device__ inline void loadv(const T *p, T (&d)[V]) {
    if constexpr (V * sizeof(T) == 8 && V == 2) {
        const f32x2 q = gld(reinterpret_cast<const f32x2 *>(p));
        d[0] = q.x, d[1] = q.y;
    } else if constexpr (V * sizeof(T) == 16) {
        if constexpr (sizeof(T) == 4) {
            const f32x4 q = gld(reinterpret_cast<const f32x4 *>(p));
            d[0] = q.x, d[1] = q.y, d[2] = q.z, d[3] = q.w;
        } else {
            const f64x2 q = gld(reinterpret_cast<const f64x2 *>(p));
            d[0] = q.x, d[1] = q.y;
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) d[v] = gld(p + v);
    }
}
// The operator records of V consecutive cells (kRec values each): e, s, diag.
template <typename T, int V>
__device__ inline void load_coef(const T *coef, size_t g, T (&e)[V], T (&s)[V], T (&d)[V]) {
    const T *q = coef + kRec * g;
    if constexpr (V == 2 && sizeof(T) == 4) {
        // 6 floats from an 8-byte boundary (g is even): three float2 loads.
        const float2 a = reinterpret_cast<const float2 *>(q)[0];
        const float2 b = reinterpret_cast<const float2 *>(q)[1];
        const float2 c = reinterpret_cast<const float2 *>(q)[2];
        e[0] = a.x, s[0] = a.y, d[0] = b.x, e[1] = b.y, s[1] = c.x, d[1] = c.y;
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) e[v] = q[kRec * v], s[v] = q[kRec * v + 1], d[v] = q[kRec * v + 2];
    }
}

// The bf16 records of V consecutive cells (8 or 16 bytes), widened to f32.
template <int V>
__device__ inline void load_coef_bf16(const unsigned short *coefh, size_t g, float (&e)[V],
                                      float (&s)[V], float (&d)[V], float (&wd)[V]) {
    if constexpr (V == 2) {
        const uint4 q = *reinterpret_cast<const uint4 *>(coefh + 4 * g);
        const unsigned a[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int v = 0; v < 2; v++) {
            e[v] = float_of_bf16(a[2 * v] & 0xffffu), s[v] = float_of_bf16(a[2 * v] >> 16);
            d[v] = float_of_bf16(a[2 * v + 1] & 0xffffu), wd[v] = float_of_bf16(a[2 * v + 1] >> 16);
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) {
            const uint2 q = *reinterpret_cast<const uint2 *>(coefh + 4 * (g + v));
            e[v] = float_of_bf16(q.x & 0xffffu), s[v] = float_of_bf16(q.x >> 16);
            d[v] = float_of_bf16(q.y & 0xffffu), wd[v] = float_of_bf16(q.y >> 16);
        }
    }
}

// The records of V cells from the level: its bf16 copy when CB (with omega / diag), else its own
// (e, s, diag; the weight is formed by weights() when the row is consumed).
template <typename T, int V, int CB>
__device__ inline void load_rec(const Lvl<T> &L, size_t g, T (&e)[V], T (&s)[V], T (&d)[V],
                                T (&wd)[V]) {
    if constexpr (CB && sizeof(T) == 4)
        load_coef_bf16<V>(L.coefh, g, e, s, d, wd);
    else
        load_coef<T, V>(L.coef, g, e, s, d);
}

// The smoother's weights omega / diag of a row of V cells (the bf16 copy's own when CB; 0 on a row
// outside the level, whose diag reads 0).
template <typename T, int V, int CB>
__device__ inline void weights(const Lvl<T> &L, const T (&d)[V], const T (&wdq)[V], T (&wd)[V]) {
#pragma unroll
    for (int v = 0; v < V; v++) {
        if constexpr (CB && sizeof(T) == 4)
            wd[v] = wdq[v];
        else
            wd[v] = d[v] != (T)0 ? L.omega / d[v] : (T)0;
    }
}

template <typename T, int V>
__device__ inline void storev(T *p, const T (&d)[V]) {
    if constexpr (V * sizeof(T) == 8 && V == 2) {
        gst(reinterpret_cast<f32x2 *>(p), f32x2{d[0], d[1]});
    } else if constexpr (V * sizeof(T) == 16) {
        if constexpr (sizeof(T) == 4)
            gst(reinterpret_cast<f32x4 *>(p), f32x4{d[0], d[1], d[2], d[3]});
        else
            gst(reinterpret_cast<f64x2 *>(p), f64x2{d[0], d[1]});
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) gst(p + v, d[v]);
    }
}

// A x at one cell from its own and its four neighbours' values, in the order of
// `Level::apply_row` (MG:203-240); ew is the west neighbour's east coupling, sn the north
// neighbour's south coupling.
template <typename T>
__device__ inline T lane_apply(T xc, T xe, T xw, T xs, T xn, T d, T e, T ew, T s, T sn, int gi,
                               int gj, int w, int h, int wrap) {
    T o = d * xc;
    if (gi + 1 < w) o -= e * xe;
    if (gi > 0) o -= ew * xw;
    if (wrap) {
        if (gi == w - 1) o -= e * xe;
        if (gi == 0) o -= ew * xw;
    }
    if (gj + 1 < h) o -= s * xs;
    if (gj > 0) o -= sn * xn;
    return o;
}

// A x for the V cells of a lane.  `inner`: the row has rows above and below and the lane's
// cells have non-wrapped east and west neighbours, so every term is present and the five are
// subtracted in `apply_row`'s order without tests (most cells take this path).
template <typename T, int V>
__device__ inline void lane_ax(const T (&x)[V], const T (&xs)[V], const T (&xn)[V],
                               const T (&d)[V], const T (&e)[V], const T (&s)[V],
                               const T (&sn)[V], int gi0, int gj, int w, int h, int wrap,
                               bool inner, T (&ax)[V]) {
    const T xe0 = from_east(x[0]), xwl = from_west(x[V - 1]), ewl = from_west(e[V - 1]);
    if (inner) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            const T xe = v + 1 < V ? x[v + 1] : xe0;
            const T xw = v > 0 ? x[v - 1] : xwl;
            const T ew = v > 0 ? e[v - 1] : ewl;
            T o = d[v] * x[v];
            o -= e[v] * xe;
            o -= ew * xw;
            o -= s[v] * xs[v];
            o -= sn[v] * xn[v];
            ax[v] = o;
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) {
            const T xe = v + 1 < V ? x[v + 1] : xe0;
            const T xw = v > 0 ? x[v - 1] : xwl;
            const T ew = v > 0 ? e[v - 1] : ewl;
            ax[v] = lane_apply(x[v], xe, xw, xs[v], xn[v], d[v], e[v], ew, s[v], sn[v], gi0 + v,
                               gj, w, h, wrap);
        }
    }
}

// One sweep of a row of V cells per lane: out = x + wd (r - A x) for the cells of row gj (xs,
// xn: the rows below and above; the row's own and the north row's couplings); 0 where !ok.
template <typename T, int V>
__device__ inline void lane_sweep(const T (&x)[V], const T (&xs)[V], const T (&xn)[V],
                                  const T (&r)[V], const T (&wd)[V], const T (&d)[V],
                                  const T (&e)[V], const T (&s)[V], const T (&sn)[V], int gi0,
                                  int gj, int w, int h, int wrap, bool ok, bool inner,
                                  T (&out)[V]) {
    T ax[V];
    lane_ax<T, V>(x, xs, xn, d, e, s, sn, gi0, gj, w, h, wrap, inner, ax);
#pragma unroll
    for (int v = 0; v < V; v++) out[v] = ok ? x[v] + wd[v] * (r[v] - ax[v]) : (T)0;
}

// A level's bands: columns of LaneGeom<V>::own, rows of `th` (a multiple of 4, chosen by the
// host so that all bands run at once in one round of resident waves).
template <int V, int H = LaneGeom<V>::halo>
__host__ __device__ inline int bands_x(int w) {
    return (w + LaneGeom<V, H>::own - 1) / LaneGeom<V, H>::own;
}
template <int V, int H = LaneGeom<V>::halo>
__host__ __device__ inline int stream_waves(int w, int h, int th) {
    return bands_x<V, H>(w) * ((h + th - 1) / th);
}

struct Band {
    int inner;   // the lane's columns all have non-wrapped east and west neighbours
    int gi0;     // the lane's first column on the level (wrapped)
    int col_ok;  // the lane's columns are on the level
    int own;     // the lane owns its columns (writes them)
    int row0;    // first row of the band
    int rows;    // band height
    int active;  // the wave has a band
};

// The band of wave `wave` of a level (waves numbered across the launch's blocks for one grid).
template <int V, int H = LaneGeom<V>::halo>
__device__ inline Band band_at(int w, int h, int wrap, int th, int wave) {
    const int lane = threadIdx.x & 63;
    constexpr int OWN = LaneGeom<V, H>::own;
    const int nbx = bands_x<V, H>(w), nby = (h + th - 1) / th;
    Band b;
    b.rows = th;
    b.active = wave < nbx * nby;
    const int bx = b.active ? wave % nbx : 0, by = b.active ? wave / nbx : 0;
    const int raw = bx * OWN + (lane - H) * V;
    b.row0 = by * th;
    if (wrap) {
        int g = raw % w;
        b.gi0 = g < 0 ? g + w : g;
        b.col_ok = 1;
    } else {
        b.gi0 = raw;
        b.col_ok = raw >= 0 && raw < w;
    }
    b.own = b.active && lane >= H && lane < 64 - H && raw >= 0 && raw < w;
    b.inner = b.col_ok && b.gi0 >= 1 && b.gi0 + V <= w - 1;
    return b;
}

// The wave number of this thread's wave when a grid's waves start at block `block`.
// The same in every lane of the wave, and marked so (readfirstlane): band rows and row offsets
// derived from it stay in scalar registers, and tests on them are scalar branches.
__device__ inline int block_wave(int block) {
    return __builtin_amdgcn_readfirstlane(block * (kBlock / 64) + (int)(threadIdx.x >> 6));
}

// The launch's blocks go to the 8 XCDs in turn (block b to XCD b mod 8).  Numbering the bands so
// that each XCD takes one contiguous run of them puts a band and the band below it (which reads
// the first's last rows as its halo) on the same XCD at about the same time, so the halo rows come
// from that XCD's L2.
__device__ inline int xcd_block(int b, int nb) {
    const int per = nb / 8, rem = nb % 8, x = b % 8, k = b / 8;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}
__device__ inline int launch_wave() { return block_wave(xcd_block(blockIdx.x, gridDim.x)); }

template <int V, int H = LaneGeom<V>::halo>
__device__ inline Band band_of(int w, int h, int wrap, int th) {
    return band_at<V, H>(w, h, wrap, th, launch_wave());
}

template <typename T, int V>
__device__ inline void zero(T (&a)[V]) {
#pragma unroll
    for (int v = 0; v < V; v++) a[v] = (T)0;
}
template <typename T, int V>
__device__ inline void copy(T (&d)[V], const T (&s)[V]) {
#pragma unroll
    for (int v = 0; v < V; v++) d[v] = s[v];
}

// Where the finest level's smoothers take r from (the f32 cycle on the conjugate gradients'
// residual): kRLevel, the level's own r; kR64, the f64 residual, rounded to T as it is read;
// kRUpdate (pre-smoothing only), the f64 residual after this iteration's r -= alpha A p, formed
// from r_in and A p as the rows are read (the reference's update pass, MG:703-723, with the same
// expression), written to r_out for the owned cells, and summed as r.r.  The f64 residual
// alternates between two buffers, so a band's halo rows are read before any update.
enum { kRLevel = 0, kR64 = 1, kRUpdate = 2, kR64K = 3, kRUpdateK = 4 };
// kR64K and kRUpdateK are kR64 and kRUpdate on the finest level of the f32 cycle with its records
// (east, south, diag) formed from kd as the rows are read, with Operator::new's expressions and
// rounded to f32 (the bits k_operator stores), instead of read; kRUpdateK also forms A p from the
// direction p (the bits the direction pass forms) instead of reading it.
__host__ __device__ constexpr bool rm_r64(int rm) { return rm == kR64 || rm == kR64K; }
__host__ __device__ constexpr bool rm_update(int rm) { return rm == kRUpdate || rm == kRUpdateK; }
__host__ __device__ constexpr bool rm_kd(int rm) { return rm == kR64K || rm == kRUpdateK; }

// Rows of loads the fused smoothers keep in flight (a divisor of 4: bands are multiples of 4 rows
// and a walk takes the band's rows plus 4).  Measured at 8192x4096: 4 rows at 3 waves/SIMD
// (144-157 VGPRs) beat 2 rows at 4 waves (116-125 VGPRs; 308/220 vs 288/201 us) and 1 row.
#ifndef DTD_SMOOTH_AHEAD
#define DTD_SMOOTH_AHEAD 4
#endif
constexpr int kSmoothAhead = DTD_SMOOTH_AHEAD;
static_assert(4 % kSmoothAhead == 0, "rows in flight must divide 4");

struct R64 {
    const double *in;  // f64 residual (kR64, kRUpdate)
    const double *ap;  // A p (kRUpdate)
    double *out;       // updated residual (kRUpdate)
    double alpha;
    const double *p;   // the direction (kRUpdateK)
    const double *kd;  // the faces' diffusivity, dt and dx^2 (kR64K, kRUpdateK)
    double dt, s2;
    double rs2;        // 1 / dx^2
};

// The f64 faces and diagonal of a lane's V cells of row j, formed from kd as Operator::new forms
// them (D:556-603): K the row's kd, Ks the row below's (the south faces; none on the last row), kx
// kd of the column beyond the wave's end lane (lane 0: west of its first cell; lane 63: east of
// its last), sn the row above's south faces.  Every lane runs it (its shuffles).
template <int V>
struct KRow {
    double e[V], s[V], d[V];
    double ew;  // the west face of the lane's first cell
};
template <int V, int P2 = 0>
__device__ __forceinline__ void krow(const double (&K)[V], const double (&Ks)[V], double kx,
                                     const double (&sn)[V], int gi0, int j, int w, int h,
                                     int wrap, double dt, double s2, double rs2, KRow<V> &o) {
    const int lane = threadIdx.x & 63;
    const double kel = from_east(K[0]);
    const double ke = lane == 63 ? kx : kel;
#pragma unroll
    for (int v = 0; v < V; v++) {
        const bool has_e = gi0 + v + 1 < w || wrap;
        o.e[v] = has_e ? face_rz<P2>(K[v], v + 1 < V ? K[v + 1] : ke, dt, s2, rs2) : 0.0;
        o.s[v] = j + 1 < h ? face_rz<P2>(K[v], Ks[v], dt, s2, rs2) : 0.0;
    }
    const double ewl = from_west(o.e[V - 1]);
    o.ew = lane == 0 ? face_rz<P2>(kx, K[0], dt, s2, rs2) : ewl;
#pragma unroll
    for (int v = 0; v < V; v++)
        o.d[v] = lane_diag64(o.e[v], v > 0 ? o.e[v - 1] : o.ew, o.s[v], sn[v], gi0 + v, j, wrap);
}

// The column beyond the wave's end lane (lane 0: west of its first cell; lane 63: east of its
// last), wrapped on a periodic level; *ok: it is on the level (the other lanes load nothing).
template <int V>
__device__ __forceinline__ int end_column(const Band &bd, int w, int wrap, bool &ok) {
    const int lane = threadIdx.x & 63;
    int gx = lane == 0 ? bd.gi0 - 1 : bd.gi0 + V;
    ok = false;
    if (bd.active && (lane == 0 || lane == 63)) {
        if (wrap)
            gx = ((gx % w) + w) % w, ok = true;
        else
            ok = gx >= 0 && gx < w;
    }
    return gx;
}

// Pre-smoothing of level L (nu = 2): z1 = wd r, z2 = z1 + wd (r - A z1) -> zout, and the coarse
// residual C.r = sum over each 2x2 block of (r - A z2): fine row j0 then j0 + 1, columns
// ascending, from +0 (MG:332-356).  Rows are loaded kAhead steps before they are consumed.
// Returns the lane's share of r.r (kRUpdate, kRUpdateK; else 0).  With kR64K and kRUpdateK the
// records are formed from kd: kd (and for kRUpdateK the direction p) come one row ahead of r, the
// south faces and A p of a row needing the row below; kRUpdateK's A p is the direction pass's.
template <typename T, int V, int RM = kRLevel, int CB = 0, int PW2 = 0>
__device__ __forceinline__ double spre_band(const Lvl<T> &L, T *__restrict__ zout,
                                            const Lvl<T> &C, int th, int wave, R64 rio = R64{}) {
    // Rows in flight; with the records from kd (edge bands only) two, as the post-smoothing.
    constexpr int A = rm_kd(RM) ? 2 : kSmoothAhead;
    constexpr bool KD = rm_kd(RM), UPD = rm_update(RM), UPK = RM == kRUpdateK;
    const int w = L.w, h = L.h, wrap = L.wrap;
    const Band bd = band_at<V>(w, h, wrap, th, wave);
    const int gi0 = bd.gi0;
    constexpr int QR = KD ? 1 : A;  // the record queues, unused when the records come from kd
    T qr[A][V], qwd[QR][V], qdg[QR][V], qe[QR][V], qs[QR][V];
    double qr64[RM == kRLevel ? 1 : A][V], qap[RM == kRUpdate ? A : 1][V];
    double qk[KD ? A : 1][V], qkx[KD ? A : 1], qp[UPK ? A : 1][V], qpx[UPK ? A : 1];
    bool x_ok = false;
    const int gx = KD ? end_column<V>(bd, w, wrap, x_ok) : 0;
    double acc_rr = 0.0;
#define DTD_FETCH_PRE(j, u)                                                                   \
    do {                                                                                      \
        const int j_ = (j);                                                                   \
        if (j_ >= 0 && j_ < h && bd.col_ok) {                                                 \
            const size_t g_ = (size_t)j_ * w + gi0;                                           \
            if (RM == kRLevel) loadv<T, V>(L.r + g_, qr[u]);                                  \
            if (RM != kRLevel) loadv<double, V>(rio.in + g_, qr64[RM == kRLevel ? 0 : u]);    \
            if (RM == kRUpdate) loadv<double, V>(rio.ap + g_, qap[RM == kRUpdate ? u : 0]);   \
            if (!KD) load_rec<T, V, CB>(L, g_, qe[KD ? 0 : u], qs[KD ? 0 : u],                \
                                        qdg[KD ? 0 : u], qwd[KD ? 0 : u]);                    \
        } else {                                                                              \
            zero(qr[u]);                                                                      \
            if (!KD) zero(qwd[KD ? 0 : u]), zero(qdg[KD ? 0 : u]), zero(qe[KD ? 0 : u]),      \
                     zero(qs[KD ? 0 : u]);                                                    \
            if (RM != kRLevel) zero(qr64[RM == kRLevel ? 0 : u]);                             \
            if (RM == kRUpdate) zero(qap[RM == kRUpdate ? u : 0]);                            \
        }                                                                                     \
        if (KD) {                                                                             \
            const size_t gx_ = (size_t)j_ * w + gx;                                           \
            const bool rx_ = j_ >= 0 && j_ < h && x_ok;                                       \
            qkx[KD ? u : 0] = rx_ ? rio.kd[gx_] : 0.0;                                        \
            if (UPK) qpx[UPK ? u : 0] = rx_ ? rio.p[gx_] : 0.0;                               \
            const int j1_ = j_ + 1;                                                           \
            if (j1_ >= 0 && j1_ < h && bd.col_ok) {                                           \
                const size_t g1_ = (size_t)j1_ * w + gi0;                                     \
                loadv<double, V>(rio.kd + g1_, qk[KD ? u : 0]);                               \
                if (UPK) loadv<double, V>(rio.p + g1_, qp[UPK ? u : 0]);                      \
            } else {                                                                          \
                zero(qk[KD ? u : 0]);                                                         \
                if (UPK) zero(qp[UPK ? u : 0]);                                               \
            }                                                                                 \
        }                                                                                     \
    } while (0)
    // Rows by age: a = newest (ja), b = ja - 1, c = ja - 2, d = ja - 3.
    T z1a[V], z1b[V], z1c[V], z2b[V], z2c[V], z2d[V];
    T ra[V], rb[V], rc[V], wda[V], wdb[V], dga[V], dgb[V], dgc[V];
    T ea[V], eb[V], ec[V], sa[V], sb[V], sc[V], sd[V];
    zero(z1a), zero(z1b), zero(z2b), zero(z2c), zero(ra), zero(rb), zero(wda), zero(dga);
    zero(dgb), zero(ea), zero(eb), zero(sa), zero(sb), zero(sc);
    T acc[V];
    zero(acc);
    // kd-formed records: row a's kd, the row above's south faces (f64); kRUpdateK: p of rows
    // a - 1 and a.
    double K1[KD ? V : 1], S0[KD ? V : 1], P0[UPK ? V : 1], P1[UPK ? V : 1];
    zero(K1), zero(S0), zero(P0), zero(P1);
    const int steps = bd.active ? bd.rows + 4 : 0;  // a multiple of A
    if constexpr (KD) {
        // Row row0 - 2's kd and p, and the row above it (its south faces, its p for A p).
        if (bd.active) {
            const int j2 = bd.row0 - 2, j3 = bd.row0 - 3;
            double K3[V];
            zero(K3);
            if (j3 >= 0 && j3 < h && bd.col_ok) {
                loadv<double, V>(rio.kd + (size_t)j3 * w + gi0, K3);
                if constexpr (UPK) loadv<double, V>(rio.p + (size_t)j3 * w + gi0, P0);
            }
            if (j2 >= 0 && j2 < h && bd.col_ok) {
                loadv<double, V>(rio.kd + (size_t)j2 * w + gi0, K1);
                if constexpr (UPK) loadv<double, V>(rio.p + (size_t)j2 * w + gi0, P1);
            }
#pragma unroll
            for (int v = 0; v < V; v++)
                S0[v] = j3 >= 0 && j3 + 1 < h ? face_rz<PW2>(K3[v], K1[v], rio.dt, rio.s2, rio.rs2) : 0.0;
        }
    }
#pragma unroll
    for (int u = 0; u < A; u++) DTD_FETCH_PRE(bd.row0 - 2 + u, u);
    for (int k0 = 0; k0 < steps; k0 += A) {
#pragma unroll
        for (int u = 0; u < A; u++) {
            const int ja = bd.row0 - 2 + k0 + u, jb = ja - 1, jc = ja - 2;
            copy(z1c, z1b), copy(z1b, z1a), copy(z2d, z2c), copy(z2c, z2b);
            copy(rc, rb), copy(rb, ra), copy(wdb, wda), copy(dgc, dgb), copy(dgb, dga);
            copy(ec, eb), copy(eb, ea), copy(sd, sc), copy(sc, sb), copy(sb, sa);
            if constexpr (KD) {
                // Row a's records from kd (its south faces from row a + 1's, one slot ahead).
                double K2[V], r0[V];
                copy(K2, qk[KD ? u : 0]);
                copy(r0, qr64[RM == kRLevel ? 0 : u]);
                const double kx = qkx[KD ? u : 0];
                double P2[UPK ? V : 1], px = 0.0;
                if constexpr (UPK) copy(P2, qp[UPK ? u : 0]), px = qpx[UPK ? u : 0];
                if (k0 + u + A < steps) DTD_FETCH_PRE(ja + A, u);
                KRow<V> kr;
                krow<V, PW2>(K1, K2, kx, S0, gi0, ja, w, h, wrap, rio.dt, rio.s2, rio.rs2, kr);
                double rv[V];
                if constexpr (UPK) {
                    // A p of row a, the direction pass's expression; r -= alpha A p.
                    const int lane = threadIdx.x & 63;
                    const double pel = from_east(P1[0]), pwl = from_west(P1[V - 1]);
                    const double pe = lane == 63 ? px : pel, pw = lane == 0 ? px : pwl;
                    const bool inner = bd.inner && ja > 0 && ja + 1 < h;
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        const double o = lane_apply64(P1[v], v + 1 < V ? P1[v + 1] : pe,
                                                      v > 0 ? P1[v - 1] : pw, P2[v], P0[v],
                                                      kr.e[v], v > 0 ? kr.e[v - 1] : kr.ew,
                                                      kr.s[v], S0[v], gi0 + v, ja, w, h, wrap,
                                                      inner);
                        rv[v] = r0[v] - rio.alpha * o;
                    }
                    copy(P0, P1), copy(P1, P2);
                } else {
                    copy(rv, r0);
                }
#pragma unroll
                for (int v = 0; v < V; v++) {
                    ea[v] = (T)kr.e[v], sa[v] = (T)kr.s[v], dga[v] = (T)kr.d[v];
                    ra[v] = (T)rv[v];
                }
                weights<T, V, 0>(L, dga, dga, wda);
                copy(S0, kr.s), copy(K1, K2);
                if (UPD) {
                    const bool own_a = bd.own && ja >= bd.row0 && ja < bd.row0 + bd.rows && ja < h;
                    if (own_a) {
                        storev<double, V>(rio.out + (size_t)ja * w + gi0, rv);
#pragma unroll
                        for (int v = 0; v < V; v++) acc_rr += rv[v] * rv[v];
                    }
                }
            } else {
                weights<T, V, CB>(L, qdg[KD ? 0 : u], qwd[KD ? 0 : u], wda);
                copy(dga, qdg[KD ? 0 : u]), copy(ea, qe[KD ? 0 : u]);
                copy(sa, qs[KD ? 0 : u]);
                if (RM == kRLevel) {
                    copy(ra, qr[u]);
                } else {
                    // The f64 residual of row a (updated when asked), rounded for the cycle.
                    const bool own_a = bd.own && ja >= bd.row0 && ja < bd.row0 + bd.rows && ja < h;
                    double rv[V];
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        const double r0 = qr64[RM == kRLevel ? 0 : u][v];
                        rv[v] = RM == kRUpdate ? r0 - rio.alpha * qap[RM == kRUpdate ? u : 0][v] : r0;
                        ra[v] = (T)rv[v];
                    }
                    if (RM == kRUpdate && own_a) {
                        storev<double, V>(rio.out + (size_t)ja * w + gi0, rv);
#pragma unroll
                        for (int v = 0; v < V; v++) acc_rr += rv[v] * rv[v];
                    }
                }
                if (k0 + u + A < steps) DTD_FETCH_PRE(ja + A, u);
            }
            const bool oka = ja >= 0 && ja < h && bd.col_ok;
#pragma unroll
            for (int v = 0; v < V; v++) z1a[v] = oka ? wda[v] * ra[v] : (T)0;
            // Second sweep of row b.
            lane_sweep<T, V>(z1b, z1a, z1c, rb, wdb, dgb, eb, sb, sc, gi0, jb, w, h, wrap,
                             jb >= 0 && jb < h && bd.col_ok, bd.inner && jb > 0 && jb + 1 < h,
                             z2b);
            if (bd.own && jb >= bd.row0 && jb < bd.row0 + bd.rows && jb < h)
                storev<T, V>(zout + (size_t)jb * w + gi0, z2b);
            // Residual of row c, restricted in 2 x 2 blocks.
            const bool okc = jc >= 0 && jc < h && bd.col_ok;
            T res[V];
            lane_ax<T, V>(z2c, z2b, z2d, dgc, ec, sc, sd, gi0, jc, w, h, wrap,
                          bd.inner && jc > 0 && jc + 1 < h, res);
#pragma unroll
            for (int v = 0; v < V; v++) res[v] = okc ? rc[v] - res[v] : (T)0;
            const T rese = from_east(res[0]);
            if (bd.own && okc && jc >= bd.row0) {
#pragma unroll
                for (int v = 0; v < V; v += 2) {
                    const int gi = gi0 + v;
                    if (V == 1 && (gi & 1)) continue;
                    const T partner = v + 1 < V ? res[v + 1] : rese;
                    if ((jc & 1) == 0) acc[v] = (T)0;
                    acc[v] += res[v];
                    if (gi + 1 < w) acc[v] += partner;
                    if ((jc & 1) == 1 || jc == h - 1)
                        C.r[(size_t)(jc >> 1) * C.w + (gi >> 1)] = acc[v];
                }
            }
        }
    }
#undef DTD_FETCH_PRE
    return acc_rr;
}

// ---- interior bands of the finest level with kd-formed records -------------------------------
//
// Most waves of the finest level's fused smoothing own a band that touches no edge of the grid:
// every lane's cells have unwrapped east and west neighbours, the end lanes' outer columns lie on
// the grid, and the rows the walk reads (three beyond the band on either side) do too.  Such a
// band takes the code below: the same cells, operations and order as the general band code takes
// on its inner path, with every boundary test, zeroing and select folded away, the loads
// unconditional, and the end lanes' outer neighbours brought in by the lane moves themselves.

// v one lane over by a DPP wave shift; the lane with no source gets 0 (bound_ctrl) where the
// value is not used, or `old`.
template <int CTRL, typename T>
__device__ inline T dpp_zero(T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
        return __builtin_bit_cast(T, ((long long)hi << 32) | (unsigned)lo);
    }
}
__device__ inline double dpp_or(double v, double old, int east) {
    const long long b = __builtin_bit_cast(long long, v), o = __builtin_bit_cast(long long, old);
    int lo, hi;
    if (east) {
        lo = __builtin_amdgcn_update_dpp((int)(o & 0xffffffff), (int)(b & 0xffffffff), 0x130, 0xF, 0xF, false);
        hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), 0x130, 0xF, 0xF, false);
    } else {
        lo = __builtin_amdgcn_update_dpp((int)(o & 0xffffffff), (int)(b & 0xffffffff), 0x138, 0xF, 0xF, false);
        hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), 0x138, 0xF, 0xF, false);
    }
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// The wave's band needs no boundary rule (the same answer in every lane, so the branch on it is
// uniform): `margin` rows above and below the band are on the level.
template <int V>
__device__ inline bool interior_band(const Band &bd, int h, int margin) {
    const bool ok = bd.active && bd.inner && bd.row0 - margin >= 0 &&
                    bd.row0 + bd.rows + margin <= h;
    return __all(ok);
}

// A x for an interior row of two-cell lanes, `apply_row`'s five terms in order; the outer lanes'
// outer values are not used.
__device__ __forceinline__ void in_ax(const float (&x)[2], const float (&xs)[2],
                                      const float (&xn)[2], const float (&d)[2],
                                      const float (&e)[2], const float (&s)[2],
                                      const float (&sn)[2], float (&ax)[2]) {
    const float xe0 = dpp_zero<0x130>(x[0]), xwl = dpp_zero<0x138>(x[1]);
    const float ewl = dpp_zero<0x138>(e[1]);
#pragma unroll
    for (int v = 0; v < 2; v++) {
        const float xe = v == 0 ? x[1] : xe0;
        const float xw = v == 1 ? x[0] : xwl;
        const float ew = v == 1 ? e[0] : ewl;
        float o = d[v] * x[v];
        o -= e[v] * xe;
        o -= ew * xw;
        o -= s[v] * xs[v];
        o -= sn[v] * xn[v];
        ax[v] = o;
    }
}
__device__ __forceinline__ void in_sweep(const float (&x)[2], const float (&xs)[2],
                                         const float (&xn)[2], const float (&r)[2],
                                         const float (&wd)[2], const float (&d)[2],
                                         const float (&e)[2], const float (&s)[2],
                                         const float (&sn)[2], float (&out)[2]) {
    float ax[2];
    in_ax(x, xs, xn, d, e, s, sn, ax);
#pragma unroll
    for (int v = 0; v < 2; v++) out[v] = x[v] + wd[v] * (r[v] - ax[v]);
}

// An interior row's f64 faces and diagonal from kd (krow with every face present): K the row's
// kd, Ks the row below's, kx the end lane's outer column's kd (lane 0: west of its first cell;
// lane 63: east of its last), sn the row above's south faces.
template <int PW2>
__device__ __forceinline__ void in_krow(const double (&K)[2], const double (&Ks)[2], double kx,
                                        const double (&sn)[2], double dt, double s2, double rs2,
                                        KRow<2> &o) {
    const double ke = dpp_or(K[0], kx, 1);
    o.e[0] = face_rz<PW2>(K[0], K[1], dt, s2, rs2);
    o.e[1] = face_rz<PW2>(K[1], ke, dt, s2, rs2);
    o.s[0] = face_rz<PW2>(K[0], Ks[0], dt, s2, rs2);
    o.s[1] = face_rz<PW2>(K[1], Ks[1], dt, s2, rs2);
    // Lane 0's west face from its outer column; the others' from the lane to the west.
    o.ew = dpp_or(o.e[1], face_rz<PW2>(kx, K[0], dt, s2, rs2), 0);
    o.d[0] = 1.0 + o.e[0] + o.ew + o.s[0] + sn[0];
    o.d[1] = 1.0 + o.e[1] + o.e[0] + o.s[1] + sn[1];
}

// spre_band for an interior band of the finest level of the f32 cycle, two cells a lane, records
// formed from kd (RM = kR64K, kRUpdateK).
template <int RM, int PW2>
__device__ __forceinline__ double spre_band_in(const Lvl<float> &L, float *__restrict__ zout,
                                               const Lvl<float> &C, const Band &bd, R64 rio) {
    constexpr int V = 2, Q = 4;  // queue slots: a slot holds a row, refilled in place
    constexpr bool UPK = RM == kRUpdateK;
    const int w = L.w;
    const int lane = threadIdx.x & 63;
    const int gi0 = bd.gi0;
    // The end lanes' outer columns; the other lanes load the nearer end lane's (the same line),
    // unused.
    const int gx = lane < 32 ? gi0 - V * lane - 1 : gi0 + V * (63 - lane) + V;
    const bool own_l = bd.own;
    const float omega = L.omega;
    // Rows by slot, at the step of row ja (slot u): r, kd and the end columns' kd and p of rows
    // ja .. ja + 3 (row ja + k in slot (u + k) mod 4), p of rows ja - 1 .. ja + 2 (row ja - 1 + k
    // in slot (u + k) mod 4).  After the step, slot u takes rows ja + 4 (and p row ja + 3): no
    // loaded value outlives its slot, so every load lands in the register its reader uses.
    double qr64[Q][V], qk[Q][V], qkx[Q], qp[UPK ? Q : 1][V], qpx[UPK ? Q : 1];
    // The last rows the walk reads: r (and the end columns) to row0 + rows + 1, kd and p one
    // further.  The queue's loads past them take those rows again (cache hits, values unused):
    // unconditional loads keep the loop's counter waits exact.
    const int jr = bd.row0 + bd.rows + 1, jk = jr + 1;
    auto fetch = [&](int j, int u) {
        const size_t g = (size_t)min(j, jr) * w, gk = (size_t)min(j, jk) * w;
        loadv<double, V>(rio.in + g + gi0, qr64[u]);
        loadv<double, V>(rio.kd + gk + gi0, qk[u]);
        qkx[u] = gld(rio.kd + g + gx);
        if constexpr (UPK) qpx[UPK ? u : 0] = gld(rio.p + g + gx);
    };
    auto fetch_p = [&](int j, int u) {
        if constexpr (UPK) loadv<double, V>(rio.p + (size_t)min(j, jk) * w + gi0, qp[UPK ? u : 0]);
    };
    float z1a[V], z1b[V], z1c[V], z2b[V], z2c[V], z2d[V];
    float ra[V], rb[V], rc[V], wda[V], wdb[V], dga[V], dgb[V], dgc[V];
    float ea[V], eb[V], ec[V], sa[V], sb[V], sc[V], sd[V];
    zero(z1a), zero(z1b), zero(z2b), zero(z2c), zero(ra), zero(rb), zero(wda), zero(dga);
    zero(dgb), zero(ea), zero(eb), zero(sa), zero(sb), zero(sc);
    float acc[V];
    zero(acc);
    double acc_rr = 0.0;
    double S0[V];
    const int base = bd.row0 - 2;  // the walk's first row a
    {
        // The south faces of row base - 1.
        double K3[V], K2[V];
        loadv<double, V>(rio.kd + (size_t)(base - 1) * w + gi0, K3);
        loadv<double, V>(rio.kd + (size_t)base * w + gi0, K2);
#pragma unroll
        for (int v = 0; v < V; v++) S0[v] = face_rz<PW2>(K3[v], K2[v], rio.dt, rio.s2, rio.rs2);
    }
#pragma unroll
    for (int u = 0; u < Q; u++) fetch(base + u, u), fetch_p(base - 1 + u, u);
    // The queue's first rows in: from here the compiler's counter waits in the loop are exact
    // (vmcnt(0) expcnt(7) lgkmcnt(15)).
    __builtin_amdgcn_s_waitcnt(0x0F70);
    const int steps = bd.rows + 4;  // a multiple of Q
    for (int k0 = 0; k0 < steps; k0 += Q) {
#pragma unroll
        for (int u = 0; u < Q; u++) {
            const int ja = base + k0 + u, jb = ja - 1, jc = ja - 2;
            const int u1 = (u + 1) % Q, u2 = (u + 2) % Q;
            copy(z1c, z1b), copy(z1b, z1a), copy(z2d, z2c), copy(z2c, z2b);
            copy(rc, rb), copy(rb, ra), copy(wdb, wda), copy(dgc, dgb), copy(dgb, dga);
            copy(ec, eb), copy(eb, ea), copy(sd, sc), copy(sc, sb), copy(sb, sa);
            // Row a's records from kd (rows a and a + 1), then A p of row a (p rows a - 1 .. a + 1).
            KRow<V> kr;
            in_krow<PW2>(qk[u], qk[u1], qkx[u], S0, rio.dt, rio.s2, rio.rs2, kr);
            double rv[V];
            if constexpr (UPK) {
                // The direction pass's expression; r -= alpha A p.
                const double(&P0)[V] = qp[UPK ? u : 0];
                const double(&P1)[V] = qp[UPK ? u1 : 0];
                const double(&P2)[V] = qp[UPK ? u2 : 0];
                const double px = qpx[UPK ? u : 0];
                const double pe = dpp_or(P1[0], px, 1);
                const double pw = dpp_or(P1[1], px, 0);
#pragma unroll
                for (int v = 0; v < V; v++) {
                    const double ew = v > 0 ? kr.e[v - 1] : kr.ew;
                    double o = kr.d[v] * P1[v];
                    o -= kr.e[v] * (v + 1 < V ? P1[v + 1] : pe);
                    o -= ew * (v > 0 ? P1[v - 1] : pw);
                    o -= kr.s[v] * P2[v];
                    o -= S0[v] * P0[v];
                    rv[v] = qr64[u][v] - rio.alpha * o;
                }
            } else {
                copy(rv, qr64[u]);
            }
            copy(S0, kr.s);
            // Slot u is spent: rows a + 4 (p: a + 3) go into it.
            fetch(ja + Q, u), fetch_p(ja + Q - 1, u);
#pragma unroll
            for (int v = 0; v < V; v++) {
                ea[v] = (float)kr.e[v], sa[v] = (float)kr.s[v], dga[v] = (float)kr.d[v];
                ra[v] = (float)rv[v];
                wda[v] = dga[v] != 0.0f ? omega / dga[v] : 0.0f;
            }
            if (UPK && own_l && ja >= bd.row0 && ja < bd.row0 + bd.rows) {
                storev<double, V>(rio.out + (size_t)ja * w + gi0, rv);
#pragma unroll
                for (int v = 0; v < V; v++) acc_rr += rv[v] * rv[v];
            }
#pragma unroll
            for (int v = 0; v < V; v++) z1a[v] = wda[v] * ra[v];
            // Second sweep of row b.
            in_sweep(z1b, z1a, z1c, rb, wdb, dgb, eb, sb, sc, z2b);
            if (own_l && jb >= bd.row0 && jb < bd.row0 + bd.rows)
                storev<float, V>(zout + (size_t)jb * w + gi0, z2b);
            // Residual of row c, restricted in 2 x 2 blocks.
            float res[V];
            in_ax(z2c, z2b, z2d, dgc, ec, sc, sd, res);
#pragma unroll
            for (int v = 0; v < V; v++) res[v] = rc[v] - res[v];
            if (own_l && jc >= bd.row0) {
                if ((jc & 1) == 0) acc[0] = 0.0f;
                acc[0] += res[0];
                acc[0] += res[1];
                if (jc & 1) gst(C.r + (size_t)(jc >> 1) * C.w + (gi0 >> 1), acc[0]);
            }
        }
    }
    return acc_rr;
}

// One band of a level's fused pre-smoothing: an interior band of the finest kd-formed level
// takes spre_band_in (with the exact multiply when dx^2 is a power of two), any other the
// general band code.
template <typename T, int V, int RM, int CB>
__device__ __forceinline__ double spre_any(const Lvl<T> &L, T *__restrict__ zout, const Lvl<T> &C,
                                           int th, int wave, R64 rio) {
    if constexpr (rm_kd(RM) && V == 2 && CB == 0 && std::is_same<T, float>::value) {
        const Band bd = band_at<V>(L.w, L.h, L.wrap, th, wave);
        if (interior_band<V>(bd, L.h, 3)) {
            if (pow2_exact(rio.s2)) return spre_band_in<RM, 1>(L, zout, C, bd, rio);
            return spre_band_in<RM, 0>(L, zout, C, bd, rio);
        }
    }
    return spre_band<T, V, RM, CB, 0>(L, zout, C, th, wave, rio);
}

// With RM = kRUpdate or kRUpdateK the residual's update of this iteration rides on the pass
// (alpha from st) and parts[block] gets the block's r.r; kR64K and kRUpdateK take kd, dt and dx^2
// for the records from st.
template <typename T, int V, int RM = kRLevel, int CB = 0>
__global__ __launch_bounds__(kBlock) void k_spre(Lvl<T> L, T *__restrict__ zout, Lvl<T> C, int th,
                                                 const CgState *__restrict__ st, R64 rio,
                                                 double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    if (st && st->done) return;
    if (rm_update(RM)) rio.alpha = st->alpha;
    if (rm_kd(RM)) rio.kd = st->kd, rio.dt = st->dt, rio.s2 = st->s2, rio.rs2 = 1.0 / st->s2;
    const double acc = spre_any<T, V, RM, CB>(L, zout, C, th, launch_wave(), rio);
    if (rm_update(RM)) {
        const double sum = block_sum<kBlock>(acc, lds);
        if (threadIdx.x == 0) parts[blockIdx.x] = sum;
    }
}

// Post-smoothing of level L (nu = 2): z' = zin + zc[j/2][i/2], z3 = z' + wd (r - A z'),
// z4 = z3 + wd (r - A z3) -> zout; with dot, per block sum dot * z4 into parts[block].
// Returns the lane's share of sum dot * z4 (0 without dot).  With RM = kR64 (kR64K), r is the f64
// residual dot itself, rounded to T as it is read; kR64K forms the records from kd (rio.kd, dt,
// s2): row b's when row a's kd arrives (its south faces).
template <typename T, int V, int RM = kRLevel, int CB = 0, int PW2 = 0>
__device__ __forceinline__ double spost_band(const Lvl<T> &L, const T *__restrict__ zin,
                                             const T *__restrict__ zc, int wc,
                                             T *__restrict__ zout, int th,
                                             const double *__restrict__ dot, int wave,
                                             R64 rio = R64{}) {
    // Two rows in flight when the records come from kd: the f64 faces' registers would otherwise
    // cost a wave per SIMD.
    constexpr bool KD = rm_kd(RM), R64M = rm_r64(RM);
    constexpr int A = KD ? 2 : kSmoothAhead;
    const int w = L.w, h = L.h, wrap = L.wrap;
    const Band bd = band_at<V>(w, h, wrap, th, wave);
    const int gi0 = bd.gi0;
    constexpr int QR = KD ? 1 : A;
    T qr[A][V], qwd[QR][V], qdg[QR][V], qe[QR][V], qs[QR][V], qz[A][V], qc[A][V];
    double qr64[R64M ? A : 1][V], qk[KD ? A : 1][V], qkx[KD ? A : 1];
    bool x_ok = false;
    const int gx = KD ? end_column<V>(bd, w, wrap, x_ok) : 0;
    // z' = z + zc is formed when the row is consumed.
#define DTD_FETCH_POST(j, u)                                                                  \
    do {                                                                                      \
        const int j_ = (j);                                                                   \
        if (j_ >= 0 && j_ < h && bd.col_ok) {                                                 \
            const size_t g_ = (size_t)j_ * w + gi0;                                           \
            if (R64M)                                                                         \
                loadv<double, V>(dot + g_, qr64[R64M ? u : 0]);                               \
            else                                                                              \
                loadv<T, V>(L.r + g_, qr[u]);                                                 \
            loadv<T, V>(zin + g_, qz[u]);                                                     \
            if (KD)                                                                           \
                loadv<double, V>(rio.kd + g_, qk[KD ? u : 0]);                                \
            else                                                                              \
                load_rec<T, V, CB>(L, g_, qe[KD ? 0 : u], qs[KD ? 0 : u], qdg[KD ? 0 : u],    \
                                   qwd[KD ? 0 : u]);                                          \
            const T *zcr_ = zc + (size_t)(j_ >> 1) * wc;                                      \
            _Pragma("unroll") for (int v = 0; v < V; v++) qc[u][v] = zcr_[(gi0 + v) >> 1];    \
        } else {                                                                              \
            zero(qr[u]), zero(qz[u]), zero(qc[u]);                                            \
            if (!KD) zero(qwd[KD ? 0 : u]), zero(qdg[KD ? 0 : u]), zero(qe[KD ? 0 : u]),      \
                     zero(qs[KD ? 0 : u]);                                                    \
            if (KD) zero(qk[KD ? u : 0]);                                                     \
            if (R64M) zero(qr64[R64M ? u : 0]);                                               \
        }                                                                                     \
        if (KD) qkx[KD ? u : 0] = j_ >= 0 && j_ < h && x_ok ? rio.kd[(size_t)j_ * w + gx] : 0.0; \
    } while (0)
    T zpa[V], zpb[V], zpc[V], z3b[V], z3c[V], z3d[V];
    T ra[V], rb[V], rc[V], wda[V], wdb[V], wdc[V], dga[V], dgb[V], dgc[V];
    T ea[V], eb[V], ec[V], sa[V], sb[V], sc[V], sd[V];
    zero(zpa), zero(zpb), zero(z3b), zero(z3c), zero(ra), zero(rb), zero(wda), zero(wdb);
    zero(dga), zero(dgb), zero(ea), zero(eb), zero(sa), zero(sb), zero(sc);
    // With RM = kR64 the f64 residual of rows a, b, c stays in registers for the dot of row c.
    double r64a[R64M ? V : 1], r64b[R64M ? V : 1], r64c[R64M ? V : 1];
    zero(r64a), zero(r64b), zero(r64c);
    // kd-formed records: row b's kd and the end column's, row c's south faces (f64).
    double KB[KD ? V : 1], S0[KD ? V : 1], kxb = 0.0;
    zero(KB), zero(S0);
    double acc = 0.0;
    const int steps = bd.active ? bd.rows + 4 : 0;
#pragma unroll
    for (int u = 0; u < A; u++) DTD_FETCH_POST(bd.row0 - 2 + u, u);
    for (int k0 = 0; k0 < steps; k0 += A) {
#pragma unroll
        for (int u = 0; u < A; u++) {
            const int ja = bd.row0 - 2 + k0 + u, jb = ja - 1, jc = ja - 2;
            copy(zpc, zpb), copy(zpb, zpa), copy(z3d, z3c), copy(z3c, z3b);
            copy(rc, rb), copy(rb, ra), copy(wdc, wdb), copy(dgc, dgb), copy(ec, eb);
            copy(sd, sc), copy(sc, sb);
            if constexpr (KD) {
                // Row b's records, its south faces from row a's kd (just arrived).
                double KA[V];
                copy(KA, qk[KD ? u : 0]);
                KRow<V> kr;
                krow<V, PW2>(KB, KA, kxb, S0, gi0, jb, w, h, wrap, rio.dt, rio.s2, rio.rs2, kr);
#pragma unroll
                for (int v = 0; v < V; v++)
                    eb[v] = (T)kr.e[v], sb[v] = (T)kr.s[v], dgb[v] = (T)kr.d[v];
                weights<T, V, 0>(L, dgb, dgb, wdb);
                copy(S0, kr.s), copy(KB, KA);
                kxb = qkx[KD ? u : 0];
            } else {
                copy(wdb, wda), copy(dgb, dga), copy(eb, ea), copy(sb, sa);
                weights<T, V, CB>(L, qdg[KD ? 0 : u], qwd[KD ? 0 : u], wda);
                copy(dga, qdg[KD ? 0 : u]), copy(ea, qe[KD ? 0 : u]);
                copy(sa, qs[KD ? 0 : u]);
            }
            if (R64M) {
                copy(r64c, r64b), copy(r64b, r64a);
#pragma unroll
                for (int v = 0; v < V; v++) {
                    r64a[v] = qr64[R64M ? u : 0][v];
                    ra[v] = (T)r64a[v];
                }
            } else {
                copy(ra, qr[u]);
            }
            const bool oka = ja >= 0 && ja < h && bd.col_ok;
#pragma unroll
            for (int v = 0; v < V; v++) zpa[v] = oka ? qz[u][v] + qc[u][v] : (T)0;
            if (k0 + u + A < steps) DTD_FETCH_POST(ja + A, u);
            lane_sweep<T, V>(zpb, zpa, zpc, rb, wdb, dgb, eb, sb, sc, gi0, jb, w, h, wrap,
                             jb >= 0 && jb < h && bd.col_ok, bd.inner && jb > 0 && jb + 1 < h,
                             z3b);
            T z4[V];
            const bool okc = jc >= 0 && jc < h && bd.col_ok;
            lane_sweep<T, V>(z3c, z3b, z3d, rc, wdc, dgc, ec, sc, sd, gi0, jc, w, h, wrap, okc,
                             bd.inner && jc > 0 && jc + 1 < h, z4);
            if (bd.own && okc && jc >= bd.row0) {
                const size_t g = (size_t)jc * w + gi0;
                storev<T, V>(zout + g, z4);
                if (R64M) {
#pragma unroll
                    for (int v = 0; v < V; v++) acc += r64c[v] * (double)z4[v];
                } else if (dot) {
#pragma unroll
                    for (int v = 0; v < V; v++) acc += dot[g + v] * (double)z4[v];
                }
            }
        }
    }
#undef DTD_FETCH_POST
    return acc;
}

// spost_band for an interior band of the finest level of the f32 cycle, two cells a lane, records
// formed from kd (RM = kR64K): r is the f64 residual `dot`, rounded as it is read.  Two rows of
// loads in flight, four rows a trip.
template <int PW2, int A = 2>
__device__ __forceinline__ double spost_band_in(const Lvl<float> &L, const float *__restrict__ zin,
                                                const float *__restrict__ zc, int wc,
                                                float *__restrict__ zout,
                                                const double *__restrict__ dot, const Band &bd,
                                                R64 rio) {
    constexpr int V = 2;
    const int w = L.w;
    const int lane = threadIdx.x & 63;
    const int gi0 = bd.gi0;
    const int gx = lane < 32 ? gi0 - V * lane - 1 : gi0 + V * (63 - lane) + V;
    const bool own_l = bd.own;
    const float omega = L.omega;
    double qr64[A][V], qk[A][V], qkx[A];
    float qz[A][V], qc[A];
    auto fetch = [&](int j, int u) {
        const size_t g = (size_t)j * w;
        loadv<double, V>(dot + g + gi0, qr64[u]);
        loadv<float, V>(zin + g + gi0, qz[u]);
        loadv<double, V>(rio.kd + g + gi0, qk[u]);
        qkx[u] = gld(rio.kd + g + gx);
        qc[u] = gld(zc + (size_t)(j >> 1) * wc + (gi0 >> 1));
    };
    float zpa[V], zpb[V], zpc[V], z3b[V], z3c[V], z3d[V];
    float ra[V], rb[V], rc[V], wdb[V], wdc[V], dgb[V], dgc[V];
    float eb[V], ec[V], sb[V], sc[V], sd[V];
    zero(zpa), zero(zpb), zero(z3b), zero(z3c), zero(ra), zero(rb), zero(wdb);
    zero(dgb), zero(eb), zero(sb), zero(sc);
    double r64a[V], r64b[V], r64c[V];
    zero(r64a), zero(r64b), zero(r64c);
    double KB[V], S0[V], kxb = 0.0;
    zero(KB), zero(S0);
    double acc = 0.0;
    const int steps = bd.rows + 4;  // a multiple of 4
#pragma unroll
    for (int u = 0; u < A; u++) fetch(bd.row0 - 2 + u, u);
    for (int k0 = 0; k0 < steps; k0 += 4) {
#pragma unroll
        for (int uu = 0; uu < 4; uu++) {
            const int u = uu % A;
            const int ja = bd.row0 - 2 + k0 + uu, jc = ja - 2;
            copy(zpc, zpb), copy(zpb, zpa), copy(z3d, z3c), copy(z3c, z3b);
            copy(rc, rb), copy(rb, ra), copy(wdc, wdb), copy(dgc, dgb), copy(ec, eb);
            copy(sd, sc), copy(sc, sb);
            // Row b's records, its south faces from row a's kd (just arrived).
            KRow<V> kr;
            in_krow<PW2>(KB, qk[u], kxb, S0, rio.dt, rio.s2, rio.rs2, kr);
#pragma unroll
            for (int v = 0; v < V; v++) {
                eb[v] = (float)kr.e[v], sb[v] = (float)kr.s[v], dgb[v] = (float)kr.d[v];
                wdb[v] = dgb[v] != 0.0f ? omega / dgb[v] : 0.0f;
            }
            copy(S0, kr.s), copy(KB, qk[u]);
            kxb = qkx[u];
            copy(r64c, r64b), copy(r64b, r64a), copy(r64a, qr64[u]);
#pragma unroll
            for (int v = 0; v < V; v++) ra[v] = (float)r64a[v];
#pragma unroll
            for (int v = 0; v < V; v++) zpa[v] = qz[u][v] + qc[u];
            fetch(min(ja + A, bd.row0 + bd.rows + 1), u);  // past the walk: its last row again
            in_sweep(zpb, zpa, zpc, rb, wdb, dgb, eb, sb, sc, z3b);
            float z4[V];
            in_sweep(z3c, z3b, z3d, rc, wdc, dgc, ec, sc, sd, z4);
            if (own_l && jc >= bd.row0) {
                storev<float, V>(zout + (size_t)jc * w + gi0, z4);
#pragma unroll
                for (int v = 0; v < V; v++) acc += r64c[v] * (double)z4[v];
            }
        }
    }
    return acc;
}

// One band of a level's fused post-smoothing (interior bands of the finest kd-formed level take
// spost_band_in).
template <typename T, int V, int RM, int CB>
__device__ __forceinline__ double spost_any(const Lvl<T> &L, const T *__restrict__ zin,
                                            const T *__restrict__ zc, int wc, T *__restrict__ zout,
                                            int th, const double *__restrict__ dot, int wave,
                                            R64 rio) {
    if constexpr (rm_kd(RM) && V == 2 && CB == 0 && std::is_same<T, float>::value) {
        const Band bd = band_at<V>(L.w, L.h, L.wrap, th, wave);
        if (interior_band<V>(bd, L.h, 3)) {
            if (pow2_exact(rio.s2)) return spost_band_in<1>(L, zin, zc, wc, zout, dot, bd, rio);
            return spost_band_in<0>(L, zin, zc, wc, zout, dot, bd, rio);
        }
    }
    return spost_band<T, V, RM, CB, 0>(L, zin, zc, wc, zout, th, dot, wave, rio);
}

template <typename T, int V, int RM = kRLevel, int CB = 0>
__global__ __launch_bounds__(kBlock) void k_spost(Lvl<T> L, const T *__restrict__ zin,
                                                  const T *__restrict__ zc, int wc,
                                                  T *__restrict__ zout, int th,
                                                  const double *__restrict__ dot,
                                                  double *__restrict__ parts,
                                                  const CgState *__restrict__ st) {
    __shared__ double lds[kBlock / 64];
    if (st && st->done) return;
    R64 rio{};
    if (rm_kd(RM)) rio.kd = st->kd, rio.dt = st->dt, rio.s2 = st->s2, rio.rs2 = 1.0 / st->s2;
    const double acc = spost_any<T, V, RM, CB>(L, zin, zc, wc, zout, th, dot, launch_wave(), rio);
    if (dot) {
        const double sum = block_sum<kBlock>(acc, lds);
        if (threadIdx.x == 0) parts[blockIdx.x] = sum;
    }
}

// Calibration: the pre-smoothing kernel's band walk, loads and store without its stencil work
// (z = r + east), to separate the access pattern's cost from the arithmetic's.
template <int V>
__global__ __launch_bounds__(kBlock) void k_bandcopy(Lvl<float> L, float *__restrict__ zout,
                                                     int th) {
    constexpr int A = 4;
    const int w = L.w, h = L.h;
    const Band bd = band_of<V>(w, h, L.wrap, th);
    const int gi0 = bd.gi0;
    float qr[A][V], qwd[A][V], qdg[A][V], qe[A][V], qs[A][V];
    const int steps = bd.active ? bd.rows + 4 : 0;
#define DTD_FETCH_BC(j, u)                                                                    \
    do {                                                                                      \
        const int j_ = (j);                                                                   \
        if (j_ >= 0 && j_ < h && bd.col_ok) {                                                 \
            const size_t g_ = (size_t)j_ * w + gi0;                                           \
            loadv<float, V>(L.r + g_, qr[u]);                                                 \
            load_coef<float, V>(L.coef, g_, qe[u], qs[u], qdg[u]);                            \
        } else {                                                                              \
            zero(qr[u]), zero(qwd[u]), zero(qdg[u]), zero(qe[u]), zero(qs[u]);                \
        }                                                                                     \
    } while (0)
#pragma unroll
    for (int u = 0; u < A; u++) DTD_FETCH_BC(bd.row0 - 2 + u, u);
    for (int k0 = 0; k0 < steps; k0 += A) {
#pragma unroll
        for (int u = 0; u < A; u++) {
            const int ja = bd.row0 - 2 + k0 + u;
            float out[V];
#pragma unroll
            for (int v = 0; v < V; v++) out[v] = qr[u][v] + qe[u][v] + qs[u][v] + qdg[u][v];
            if (k0 + u + A < steps) DTD_FETCH_BC(ja + A, u);
            if (bd.own && ja >= bd.row0 && ja < bd.row0 + bd.rows && ja < h)
                storev<float, V>(zout + (size_t)ja * w + gi0, out);
        }
    }
#undef DTD_FETCH_BC
}

// ---- the conjugate gradients' direction pass, streamed (MG:633-695) ------------------------
//
// One f64 cell a lane, a band of 62 columns (one halo lane each side) walked row by row with
// four rows of loads in flight.  Per row: the direction p = z + beta p_old (z on the first
// iteration) of the newest row, stored for the owned cells; the pending x += alpha p_old of the
// owned cells; and A p of the row above with its p . A p.  The f64 couplings come packed as
// (east, south) pairs.

// Returns the lane's share of p . A p.
template <typename TZ>
__device__ __forceinline__ double sdir_band(const Grid &g, const double *__restrict__ kd,
                                            double dt, double s2, const TZ *__restrict__ z,
                                            const double *__restrict__ pold,
                                            double *__restrict__ pnew, double *__restrict__ ap,
                                            double *__restrict__ x, int th,
                                            const CgState *__restrict__ st, int wave) {
    constexpr int A = 4;
    const int have_beta = st->have_beta, pending = st->have_pending;
    const double beta = st->beta, alpha = st->alpha;
    const int w = g.w, h = g.h, wrap = g.wrap;
    const Band bd = band_at<1, 1>(w, h, wrap, th, wave);
    const int gi = bd.gi0;
    const bool inner_col = bd.col_ok && gi >= 1 && gi + 1 <= w - 1;
    // Raw loads only: p, and the owned cells' x update, are formed when the row is consumed, so
    // the loads of a row stay in flight for A steps.  The operator's faces are formed from kd
    // with the reference's expression (`face`), the east one from the east lane's kd and the
    // south one of row b when row a (below it) arrives: the same bits as Operator::new's.
    double qk[A], qpo[A], qx[A];
    TZ qz[A];
    const bool need_old = have_beta || pending;
#define DTD_FETCH_DIR(j, u)                                                                   \
    do {                                                                                      \
        const int j_ = (j);                                                                   \
        if (j_ >= 0 && j_ < h && bd.col_ok && j_ <= bd.row0 + bd.rows) {                      \
            const size_t g_ = (size_t)j_ * w + gi;                                            \
            qk[u] = kd[g_], qz[u] = z[g_];                                                    \
            qpo[u] = need_old ? pold[g_] : 0.0;                                               \
            qx[u] = pending && bd.own ? x[g_] : 0.0;                                          \
        } else {                                                                              \
            qk[u] = qpo[u] = qx[u] = 0.0;                                                     \
            qz[u] = (TZ)0;                                                                    \
        }                                                                                     \
    } while (0)
    const bool has_e = gi + 1 < w || wrap;
    double pa = 0, pb = 0, pc = 0, eb = 0, ea = 0, sb = 0, sc = 0, ka = 0;
    double acc = 0.0;
    const int steps = bd.active ? bd.rows + 4 : 0;  // rows row0 - 1 .. row0 + rows (+2 spare)
#pragma unroll
    for (int u = 0; u < A; u++) DTD_FETCH_DIR(bd.row0 - 1 + u, u);
    for (int k0 = 0; k0 < steps; k0 += A) {
#pragma unroll
        for (int u = 0; u < A; u++) {
            const int ja = bd.row0 - 1 + k0 + u, jb = ja - 1;
            pc = pb, pb = pa, eb = ea, sc = sb;
            const double kb = ka;
            ka = qk[u];
            const double ke = from_east(ka);
            ea = has_e ? face(ka, ke, dt, s2) : 0.0;
            sb = ja < h ? face(kb, ka, dt, s2) : 0.0;  // row b's south face
            const double po = qpo[u], xo = qx[u];
            pa = have_beta ? (double)qz[u] + beta * po : (double)qz[u];
            if (k0 + u + A < steps) DTD_FETCH_DIR(ja + A, u);
            if (bd.own && ja >= bd.row0 && ja < bd.row0 + bd.rows && ja < h) {
                const size_t ga = (size_t)ja * w + gi;
                pnew[ga] = pa;
                if (pending) x[ga] = xo + alpha * po;
            }
            // A p of row b.
            const double pe = from_east(pb), pw = from_west(pb), ew = from_west(eb);
            if (bd.own && jb >= bd.row0 && jb < bd.row0 + bd.rows && jb < h) {
                double o;
                if (inner_col && jb > 0 && jb + 1 < h) {
                    const double d = 1.0 + eb + ew + sb + sc;
                    o = d * pb;
                    o -= eb * pe;
                    o -= ew * pw;
                    o -= sb * pa;
                    o -= sc * pc;
                } else {
                    const double wv = (gi > 0 || wrap) ? ew : 0.0;
                    const double nv = jb > 0 ? sc : 0.0;
                    const double d = 1.0 + eb + wv + sb + nv;
                    o = d * pb;
                    if (gi + 1 < w || wrap) o -= eb * pe;
                    if (gi > 0 || wrap) o -= wv * pw;
                    if (jb + 1 < h) o -= sb * pa;
                    if (jb > 0) o -= nv * pc;
                }
                if (ap) ap[(size_t)jb * w + gi] = o;  // null: the pre-smoothing forms A p itself
                acc += pb * o;
            }
        }
    }
#undef DTD_FETCH_DIR
    return acc;
}

template <typename TZ>
__global__ __launch_bounds__(kBlock) void k_sdir(Grid g, const TZ *__restrict__ z,
                                                 const double *__restrict__ pold,
                                                 double *__restrict__ pnew, double *__restrict__ ap,
                                                 double *__restrict__ x, int th,
                                                 double *__restrict__ parts,
                                                 const CgState *__restrict__ st) {
    __shared__ double lds[kBlock / 64];
    if (st->done) return;
    const double acc = sdir_band<TZ>(g, st->kd, st->dt, st->s2, z, pold, pnew, ap, x, th, st,
                                     launch_wave());
    const double sum = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = sum;
}

// ---- the other f64 passes, streamed the same way -------------------------------------------

// The first residual (MG:596-619): r = b - A x, its copy in the cycle's precision, per block
// (r.r, b.b).  x is read from xin; when xout differs (the first guess is b itself) the owned
// cells of x are written there.
// Returns the lane's shares of (r.r, b.b).
template <typename TR>
__device__ __forceinline__ double2 sresid_band(const Grid &g, const double *__restrict__ kd,
                                               double dt, double s2,
                                               const double *__restrict__ b,
                                               const double *__restrict__ xin,
                                               double *__restrict__ xout, double *__restrict__ r,
                                               TR *__restrict__ rc, int th, int wave) {
    constexpr int A = 4;
    const int w = g.w, h = g.h, wrap = g.wrap;
    const Band bd = band_at<1, 1>(w, h, wrap, th, wave);
    const int gi = bd.gi0;
    const bool inner_col = bd.col_ok && gi >= 1 && gi + 1 <= w - 1;
    const bool copy = xout != xin;
    const bool has_e = gi + 1 < w || wrap;
    double qx[A], qk[A], qb[A];
#define DTD_FETCH_RES(j, u)                                                                   \
    do {                                                                                      \
        const int j_ = (j);                                                                   \
        if (j_ >= 0 && j_ < h && bd.col_ok && j_ <= bd.row0 + bd.rows) {                      \
            const size_t g_ = (size_t)j_ * w + gi;                                            \
            qk[u] = kd[g_], qx[u] = xin[g_], qb[u] = b[g_];                                   \
        } else {                                                                              \
            qk[u] = qx[u] = qb[u] = 0.0;                                                      \
        }                                                                                     \
    } while (0)
    double xa = 0, xb = 0, xc = 0, ea = 0, eb = 0, sb = 0, sc = 0, ba = 0, bb = 0, ka = 0;
    double acc_r = 0.0, acc_b = 0.0;
    const int steps = bd.active ? bd.rows + 4 : 0;
#pragma unroll
    for (int u = 0; u < A; u++) DTD_FETCH_RES(bd.row0 - 1 + u, u);
    for (int k0 = 0; k0 < steps; k0 += A) {
#pragma unroll
        for (int u = 0; u < A; u++) {
            const int ja = bd.row0 - 1 + k0 + u, jb = ja - 1;
            xc = xb, xb = xa, eb = ea, sc = sb, bb = ba;
            const double kb = ka;
            ka = qk[u];
            const double ke = from_east(ka);
            xa = qx[u], ba = qb[u];
            ea = has_e ? face(ka, ke, dt, s2) : 0.0;
            sb = ja < h ? face(kb, ka, dt, s2) : 0.0;  // row b's south face
            if (k0 + u + A < steps) DTD_FETCH_RES(ja + A, u);
            const bool rows_a = ja >= bd.row0 && ja < bd.row0 + bd.rows && ja < h;
            if (copy && bd.own && rows_a) xout[(size_t)ja * w + gi] = xa;
            const double xe = from_east(xb), xw = from_west(xb), ew = from_west(eb);
            if (bd.own && jb >= bd.row0 && jb < bd.row0 + bd.rows && jb < h) {
                const double o = lane_apply64(xb, xe, xw, xa, xc, eb, ew, sb, sc, gi, jb, w, h,
                                              wrap, inner_col && jb > 0 && jb + 1 < h);
                const double rv = bb - o;
                const size_t gb = (size_t)jb * w + gi;
                r[gb] = rv;
                if (rc) rc[gb] = (TR)rv;
                acc_r += rv * rv;
                acc_b += bb * bb;
            }
        }
    }
#undef DTD_FETCH_RES
    return make_double2(acc_r, acc_b);
}

// Per block, the pair (sum of v.x, sum of v.y) into parts[2 * slot], parts[2 * slot + 1].
__device__ inline void block_pair(double2 v, double *parts, int slot, double *lds) {
    const double s1 = block_sum<kBlock>(v.x, lds);
    const double s2 = block_sum<kBlock>(v.y, lds);
    if (threadIdx.x == 0) parts[2 * slot] = s1, parts[2 * slot + 1] = s2;
}

template <typename TR>
__global__ __launch_bounds__(kBlock) void k_sresid(Grid g, const double *__restrict__ kd,
                                                   double dt, double s2,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ xin,
                                                   double *__restrict__ xout,
                                                   double *__restrict__ r, TR *__restrict__ rc,
                                                   int th, double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    block_pair(sresid_band<TR>(g, kd, dt, s2, b, xin, xout, r, rc, th, launch_wave()),
               parts, blockIdx.x, lds);
}

// After the loop: the last update x' = x + alpha p (MG:734-736; p of the last iteration, in
// buffer its % 2 ... p0 for odd its), written to xout, and the put-back's sums over x'
// (MG:737-755): per block (sum (b - A x'), sum (diag - 1)).
__device__ __forceinline__ double2 finish1_band(const Grid &g, const double *__restrict__ kd,
                                                double dt, double s2,
                                                const double *__restrict__ b,
                                                const double *__restrict__ x,
                                                const double *__restrict__ p0,
                                                const double *__restrict__ p1,
                                                double *__restrict__ xout, int th,
                                                const CgState *__restrict__ st, int wave) {
    constexpr int A = 4;
    const int pending = st->have_pending;
    const double alpha = st->alpha;
    const double *p = (st->its & 1) ? p0 : p1;
    const int w = g.w, h = g.h, wrap = g.wrap;
    const Band bd = band_at<1, 1>(w, h, wrap, th, wave);
    const int gi = bd.gi0;
    const bool inner_col = bd.col_ok && gi >= 1 && gi + 1 <= w - 1;
    const bool has_e = gi + 1 < w || wrap;
    double qx[A], qp[A], qk[A], qb[A];
#define DTD_FETCH_FIN(j, u)                                                                   \
    do {                                                                                      \
        const int j_ = (j);                                                                   \
        if (j_ >= 0 && j_ < h && bd.col_ok && j_ <= bd.row0 + bd.rows) {                      \
            const size_t g_ = (size_t)j_ * w + gi;                                            \
            qk[u] = kd[g_], qx[u] = x[g_], qb[u] = b[g_];                                     \
            qp[u] = pending ? p[g_] : 0.0;                                                    \
        } else {                                                                              \
            qk[u] = qx[u] = qb[u] = qp[u] = 0.0;                                              \
        }                                                                                     \
    } while (0)
    double xa = 0, xb = 0, xc = 0, ea = 0, eb = 0, sb = 0, sc = 0, ba = 0, bb = 0, ka = 0;
    double acc_l = 0.0, acc_w = 0.0;
    const int steps = bd.active ? bd.rows + 4 : 0;
#pragma unroll
    for (int u = 0; u < A; u++) DTD_FETCH_FIN(bd.row0 - 1 + u, u);
    for (int k0 = 0; k0 < steps; k0 += A) {
#pragma unroll
        for (int u = 0; u < A; u++) {
            const int ja = bd.row0 - 1 + k0 + u, jb = ja - 1;
            xc = xb, xb = xa, eb = ea, sc = sb, bb = ba;
            const double kb = ka;
            ka = qk[u];
            const double ke = from_east(ka);
            xa = pending ? qx[u] + alpha * qp[u] : qx[u];
            ba = qb[u];
            ea = has_e ? face(ka, ke, dt, s2) : 0.0;
            sb = ja < h ? face(kb, ka, dt, s2) : 0.0;  // row b's south face
            if (k0 + u + A < steps) DTD_FETCH_FIN(ja + A, u);
            if (bd.own && ja >= bd.row0 && ja < bd.row0 + bd.rows && ja < h)
                xout[(size_t)ja * w + gi] = xa;
            const double xe = from_east(xb), xw = from_west(xb), ew = from_west(eb);
            if (bd.own && jb >= bd.row0 && jb < bd.row0 + bd.rows && jb < h) {
                const double o = lane_apply64(xb, xe, xw, xa, xc, eb, ew, sb, sc, gi, jb, w, h,
                                              wrap, inner_col && jb > 0 && jb + 1 < h);
                acc_l += bb - o;
                acc_w += lane_diag64(eb, ew, sb, sc, gi, jb, wrap) - 1.0;
            }
        }
    }
#undef DTD_FETCH_FIN
    return make_double2(acc_l, acc_w);
}

__global__ __launch_bounds__(kBlock) void k_finish1(Grid g, const double *__restrict__ kd,
                                                    double dt, double s2,
                                                    const double *__restrict__ b,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ p0,
                                                    const double *__restrict__ p1,
                                                    double *__restrict__ xout, int th,
                                                    double *__restrict__ parts,
                                                    const CgState *__restrict__ st) {
    __shared__ double lds[kBlock / 64];
    block_pair(finish1_band(g, kd, dt, s2, b, x, p0, p1, xout, th, st, launch_wave()),
               parts, blockIdx.x, lds);
}

// The put-back x'' = x' + (lost / weight) (diag - 1) when weight > 0 and lost != 0 (MG:756-761),
// written to xout, and the true residual of x'': per block (sum (b - A x'')^2, sum b^2).
__device__ __forceinline__ double2 finish2_band(const Grid &g, const double *__restrict__ kd,
                                                double dt, double s2,
                                                const double *__restrict__ b,
                                                const double *__restrict__ xp,
                                                double *__restrict__ xout, int th,
                                                const CgState *__restrict__ st, int wave) {
    constexpr int A = 4;
    const double lost = st->lost, weight = st->weight;
    const bool apply = weight > 0.0 && lost != 0.0;
    const double kk = apply ? lost / weight : 0.0;
    const int w = g.w, h = g.h, wrap = g.wrap;
    const Band bd = band_at<1, 1>(w, h, wrap, th, wave);
    const int gi = bd.gi0;
    const bool inner_col = bd.col_ok && gi >= 1 && gi + 1 <= w - 1;
    const bool has_e = gi + 1 < w || wrap;
    double qx[A], qk[A], qb[A];
#define DTD_FETCH_FIN2(j, u)                                                                  \
    do {                                                                                      \
        const int j_ = (j);                                                                   \
        if (j_ >= 0 && j_ < h && bd.col_ok && j_ <= bd.row0 + bd.rows + 1) {                  \
            const size_t g_ = (size_t)j_ * w + gi;                                            \
            qk[u] = kd[g_], qx[u] = xp[g_], qb[u] = b[g_];                                    \
        } else {                                                                              \
            qk[u] = qx[u] = qb[u] = 0.0;                                                      \
        }                                                                                     \
    } while (0)
    // Rows by age: a (just loaded: x' and kd), b (its faces complete now that a is here: x'' of
    // b is formed), c (the residual's row), d.  The window starts two rows above the band, so the
    // x'' of the row above the band has its north face.
    double xpa = 0, xb = 0, xc = 0, xd = 0, ea = 0, eb = 0, ec = 0, sb = 0, sc = 0, sd = 0;
    double ba = 0, bb = 0, bc = 0, ka = 0;
    double acc_r = 0.0, acc_b = 0.0;
    const int steps = bd.active ? bd.rows + 4 : 0;
#pragma unroll
    for (int u = 0; u < A; u++) DTD_FETCH_FIN2(bd.row0 - 2 + u, u);
    for (int k0 = 0; k0 < steps; k0 += A) {
#pragma unroll
        for (int u = 0; u < A; u++) {
            const int ja = bd.row0 - 2 + k0 + u, jb = ja - 1, jc = ja - 2;
            xd = xc, xc = xb, ec = eb, eb = ea, sd = sc, sc = sb, bc = bb, bb = ba;
            const double xpb = xpa, kb = ka;
            ka = qk[u];
            const double ke = from_east(ka);
            xpa = qx[u], ba = qb[u];
            ea = has_e ? face(ka, ke, dt, s2) : 0.0;
            sb = ja < h ? face(kb, ka, dt, s2) : 0.0;  // row b's south face
            if (k0 + u + A < steps) DTD_FETCH_FIN2(ja + A, u);
            // The put-back on row b: x'' = x' + (lost / weight) (diag - 1) (MG:756-761).
            const double ewb = from_west(eb);
            double v = xpb;
            if (apply && jb >= 0 && jb < h && bd.col_ok)
                v = v + kk * (lane_diag64(eb, ewb, sb, sc, gi, jb, wrap) - 1.0);
            xb = v;
            if (bd.own && jb >= bd.row0 && jb < bd.row0 + bd.rows && jb < h)
                xout[(size_t)jb * w + gi] = xb;
            // The true residual of row c.
            const double xe = from_east(xc), xw = from_west(xc), ewc = from_west(ec);
            if (bd.own && jc >= bd.row0 && jc < bd.row0 + bd.rows && jc < h) {
                const double o = lane_apply64(xc, xe, xw, xb, xd, ec, ewc, sc, sd, gi, jc, w, h,
                                              wrap, inner_col && jc > 0 && jc + 1 < h);
                const double rv = bc - o;
                acc_r += rv * rv;
                acc_b += bc * bc;
            }
        }
    }
#undef DTD_FETCH_FIN2
    return make_double2(acc_r, acc_b);
}

__global__ __launch_bounds__(kBlock) void k_finish2(Grid g, const double *__restrict__ kd,
                                                    double dt, double s2,
                                                    const double *__restrict__ b,
                                                    const double *__restrict__ xp,
                                                    double *__restrict__ xout, int th,
                                                    double *__restrict__ parts,
                                                    const CgState *__restrict__ st) {
    __shared__ double lds[kBlock / 64];
    block_pair(finish2_band(g, kd, dt, s2, b, xp, xout, th, st, launch_wave()), parts,
               blockIdx.x, lds);
}

// ---- the first guess (D:370-489) -----------------------------------------------------------

// The x sweep, one thread per row: `cyclic` (D:1050-1071) on rows of a wrapping grid wider than
// two cells, `thomas` (D:1029-1046) otherwise, with the tridiagonal coefficients formed from kd
// as they are used.  Both Thomas solves of `cyclic` share their elimination (they factor the
// same matrix), so they run in one pass; each value is the one the separate solves form.
// cbuf and zbuf are w*h scratch arrays.
// One line of a grid array: element i at p[i * s] (s = 1 along a row of the grid, s = h along a
// row of the transposed grid).
template <typename T>
struct Line {
    T *p;
    int s;
    __device__ T &operator[](int i) const { return p[(size_t)i * s]; }
};

// One thread runs a line's recurrences, so each step would wait on memory: the loads run
// kLodAhead positions ahead, in registers (chunk c + 1 is loaded while chunk c is being used).
// The arithmetic is the recurrences', operation for operation.
constexpr int kLodAhead = 16;
// One wave a block for the one-thread-a-line kernels: their few waves then spread over as many
// CUs (each with its own L1 and memory pipeline) instead of sharing a few.
constexpr int kLineBlock = 64;


// v[q] = L[i0 + D q] (D = 1 or -1) for the positions inside [0, n); 0 elsewhere.
template <int D, typename T>
__device__ __forceinline__ void line_chunk(Line<T> L, int i0, int n, double (&v)[kLodAhead]) {
    const int lo = D > 0 ? i0 : i0 - (kLodAhead - 1), hi = D > 0 ? i0 + kLodAhead - 1 : i0;
    if (lo >= 0 && hi < n) {
        const T *p = L.p + (ptrdiff_t)i0 * L.s;
#pragma unroll
        for (int q = 0; q < kLodAhead; q++) v[q] = p[(ptrdiff_t)(D * q) * L.s];
    } else {
#pragma unroll
        for (int q = 0; q < kLodAhead; q++) {
            const int i = i0 + D * q;
            v[q] = (i >= 0 && i < n) ? (double)L[i] : 0.0;
        }
    }
}

// Runs step(i, q, slot) for i = i0 + D q over [first, end) (D = 1: i from first up to end - 1;
// D = -1: from first down to end + 1), a chunk of kLodAhead at a time, the chunks alternating
// between two register slots: fetch(i0, slot) loads the chunk at i0 into `slot` while the chunk
// before it is used from the other (no copies between the slots, which would wait on the loads).
// The chunk at `first` must be in slot 0 on entry.  Pairs of chunks whose every position lies
// before `whole` (D = 1) or after it (D = -1) run without bounds tests.
template <int D, typename Fetch, typename Step>
__device__ __forceinline__ void line_walk(int first, int end, int whole, Fetch fetch, Step step) {
    constexpr int A = kLodAhead;
    int i0 = first;
    for (; D > 0 ? i0 + 2 * A - 1 < whole : i0 - (2 * A - 1) > whole; i0 += 2 * D * A) {
        fetch(i0 + D * A, 1);
#pragma unroll
        for (int q = 0; q < A; q++) step(i0 + D * q, q, 0);
        fetch(i0 + 2 * D * A, 0);
#pragma unroll
        for (int q = 0; q < A; q++) step(i0 + D * (A + q), q, 1);
    }
    for (; D > 0 ? i0 < end : i0 > end; i0 += 2 * D * A) {
        fetch(i0 + D * A, 1);
#pragma unroll
        for (int q = 0; q < A; q++)
            if (D > 0 ? i0 + q < end : i0 - q > end) step(i0 + D * q, q, 0);
        fetch(i0 + 2 * D * A, 0);
#pragma unroll
        for (int q = 0; q < A; q++)
            if (D > 0 ? i0 + A + q < end : i0 - A - q > end) step(i0 + D * (A + q), q, 1);
    }
}

// `thomas` (D:1029-1046) on a line of n cells that does not wrap: faces from kd between
// neighbours, none past the ends (row q: lo = -face(q - 1), up = -face(q), di = 1 + face(q - 1)
// + face(q)).  b is solved in place; cc is scratch.
__device__ __forceinline__ void lod_thomas(Line<const double> k, Line<double> b, Line<double> cc,
                                           int n, double dt, double s2) {
    const Line<const double> k1{k.p + k.s, k.s};  // k1[i] = k[i + 1]
    double kc = k[0];
    double fprev = n > 1 ? face(kc, k1[0], dt, s2) : 0.0;  // face(i - 1) at step i
    double bt = 1.0 + 0.0 + fprev;
    cc[0] = 0.0;
    double prev = b[0] / bt;
    b[0] = prev;
    if (n > 1) kc = k1[0];
    double kn[2][kLodAhead], bv[2][kLodAhead];
    // Step i (1 .. n - 1) reads k[i + 1] (the faces past the end are 0) and b[i].
    line_chunk<1>(k1, 1, n - 1, kn[0]);
    line_chunk<1>(b, 1, n, bv[0]);
    line_walk<1>(
        1, n, n - 1,
        [&](int i0, int t) {
            line_chunk<1>(k1, i0, n - 1, kn[t]);
            line_chunk<1>(b, i0, n, bv[t]);
        },
        [&](int i, int q, int t) {
            const double fi = i + 1 < n ? face(kc, kn[t][q], dt, s2) : 0.0;
            kc = kn[t][q];
            const double up_prev = -fprev, lo = -fprev;
            const double di = 1.0 + fprev + fi;
            const double ci = up_prev / bt;
            cc[i] = ci;
            bt = di - lo * ci;
            prev = (bv[t][q] - lo * prev) / bt;
            b[i] = prev;
            fprev = fi;
        });
    // Back substitution: b[i] -= cc[i + 1] b[i + 1], with b[i + 1] carried.
    const Line<double> c1{cc.p + cc.s, cc.s};
    double cv[2][kLodAhead];
    line_chunk<-1>(c1, n - 2, n - 1, cv[0]);
    line_chunk<-1>(b, n - 2, n, bv[0]);
    line_walk<-1>(
        n - 2, -1, -1,
        [&](int i0, int t) {
            line_chunk<-1>(c1, i0, n - 1, cv[t]);
            line_chunk<-1>(b, i0, n, bv[t]);
        },
        [&](int i, int q, int t) {
            prev = bv[t][q] - cv[t][q] * prev;
            b[i] = prev;
        });
}

// One row of w cells: kd in k, the right-hand side in b (solved in place), cc and zz scratch.
__device__ __forceinline__ void lod_line_x(Line<const double> k, Line<double> b, Line<double> cc,
                                           Line<double> zz, int w, int wrap_x, double dt,
                                           double s2) {
    const bool ring = wrap_x && w > 1;
    auto fc = [&](int i) -> double {
        if (i + 1 < w) return face(k[i], k[i + 1], dt, s2);
        return ring ? face(k[w - 1], k[0], dt, s2) : 0.0;
    };
    if (wrap_x && w > 2) {
        const double flast = fc(w - 1);
        // lo[i] = -west(i), up[i] = -east(i), di[i] = 1 + west + east; corners alpha = up[w-1],
        // beta = lo[0] (D:386-399).
        const double alpha = -flast, beta = -flast;
        double f0 = fc(0);
        const double di0 = 1.0 + flast + f0;
        const double gamma = -di0;
        // bb[0] = di[0] - gamma, bb[n-1] = di[n-1] - alpha beta / gamma; l[0] = 0, u[n-1] = 0.
        double bt = di0 - gamma;
        cc[0] = 0.0;
        const double x0 = b[0] / bt;
        b[0] = x0;
        double zprev = gamma / bt;
        zz[0] = zprev;
        double xprev = x0;
        double fprev = f0;  // face(i - 1)
        double kc = k[1];   // k[i] at step i
        const Line<const double> k1{k.p + k.s, k.s};  // k1[i] = k[i + 1]
        double kn[2][kLodAhead], bv[2][kLodAhead];
        // Step i (1 .. w - 1) reads k[i + 1] (below w) and b[i].
        line_chunk<1>(k1, 1, w - 1, kn[0]);
        line_chunk<1>(b, 1, w, bv[0]);
        line_walk<1>(
            1, w, w - 1,
            [&](int i0, int t) {
                line_chunk<1>(k1, i0, w - 1, kn[t]);
                line_chunk<1>(b, i0, w, bv[t]);
            },
            [&](int i, int q, int t) {
                const bool last = i == w - 1;
                const double fi = last ? flast : face(kc, kn[t][q], dt, s2);
                kc = kn[t][q];
                const double lo = -fprev, up_prev = -fprev;  // up[i-1] = -face(i-1)
                double di = 1.0 + fprev + fi;
                if (last) di = di - alpha * beta / gamma;
                const double ci = up_prev / bt;
                cc[i] = ci;
                bt = di - lo * ci;
                const double xi = (bv[t][q] - lo * xprev) / bt;
                b[i] = xi;
                xprev = xi;
                const double z0 = last ? alpha : 0.0;
                const double zi = (z0 - lo * zprev) / bt;
                zz[i] = zi;
                zprev = zi;
                fprev = fi;
            });
        // Back substitution of both solves: b[i] -= cc[i + 1] b[i + 1], zz likewise.
        const double blast = xprev, zlast = zprev;  // b[w - 1], zz[w - 1]
        double bn = xprev, zn = zprev;
        const Line<double> c1{cc.p + cc.s, cc.s};
        double cv[2][kLodAhead], zv[2][kLodAhead];
        line_chunk<-1>(c1, w - 2, w - 1, cv[0]);
        line_chunk<-1>(b, w - 2, w, bv[0]);
        line_chunk<-1>(zz, w - 2, w, zv[0]);
        line_walk<-1>(
            w - 2, -1, -1,
            [&](int i0, int t) {
                line_chunk<-1>(c1, i0, w - 1, cv[t]);
                line_chunk<-1>(b, i0, w, bv[t]);
                line_chunk<-1>(zz, i0, w, zv[t]);
            },
            [&](int i, int q, int t) {
                bn = bv[t][q] - cv[t][q] * bn;
                b[i] = bn;
                zn = zv[t][q] - cv[t][q] * zn;
                zz[i] = zn;
            });
        // bn, zn are b[0], zz[0] now.
        const double fact = (bn + beta * blast / gamma) / (1.0 + zn + beta * zlast / gamma);
        line_chunk<1>(b, 0, w, bv[0]);
        line_chunk<1>(zz, 0, w, zv[0]);
        line_walk<1>(
            0, w, w,
            [&](int i0, int t) {
                line_chunk<1>(b, i0, w, bv[t]);
                line_chunk<1>(zz, i0, w, zv[t]);
            },
            [&](int i, int q, int t) { b[i] = bv[t][q] - fact * zv[t][q]; });
        return;
    }
    if (!wrap_x) {
        lod_thomas(k, b, cc, w, dt, s2);
        return;
    }
    // Plain Thomas on one or two wrapping cells: two cells fold their two faces into one; one
    // cell has diagonal 1 (D:400-419).
    double f2 = 0.0;
    if (w == 2) f2 = fc(0) + fc(1);
    auto lo_at = [&](int i) -> double { return i == 0 ? 0.0 : -f2; };
    auto up_at = [&](int i) -> double { return i == w - 1 ? 0.0 : -f2; };
    auto di_at = [&](int i) -> double { return w == 1 ? 1.0 : 1.0 + f2; };
    double bt = di_at(0);
    cc[0] = 0.0;
    b[0] /= bt;
    for (int i = 1; i < w; i++) {
        const double ci = up_at(i - 1) / bt;
        cc[i] = ci;
        const double lo = lo_at(i);
        bt = di_at(i) - lo * ci;
        b[i] = (b[i] - lo * b[i - 1]) / bt;
    }
    for (int i = w - 2; i >= 0; i--) b[i] -= cc[i + 1] * b[i + 1];
}

__device__ __forceinline__ void lod_row(const Grid &g, int wrap_x, const double *kd, double dt,
                                        double s2, double *x, double *cbuf, double *zbuf, int j) {
    const size_t o = (size_t)j * g.w;
    lod_line_x(Line<const double>{kd + o, 1}, Line<double>{x + o, 1}, Line<double>{cbuf + o, 1},
               Line<double>{zbuf + o, 1}, g.w, wrap_x, dt, s2);
}

__global__ __launch_bounds__(kLineBlock) void k_lod_rows(Grid g, int wrap_x, const double *kd, double dt,
                                                     double s2, double *x, double *cbuf,
                                                     double *zbuf) {
    const int j = blockIdx.x * kLineBlock + threadIdx.x;
    if (j < g.h) lod_row(g, wrap_x, kd, dt, s2, x, cbuf, zbuf, j);
}

// The z sweep, one thread per column (coalesced across the wave): `thomas` down the column with
// north and south faces (D:436-488).
__device__ __forceinline__ void lod_col(const Grid &g, const double *kd, double dt, double s2,
                                        double *x, double *cbuf, int i) {
    lod_thomas(Line<const double>{kd + i, g.w}, Line<double>{x + i, g.w},
               Line<double>{cbuf + i, g.w}, g.h, dt, s2);
}

__global__ __launch_bounds__(kLineBlock) void k_lod_cols(Grid g, const double *kd, double dt, double s2,
                                                     double *x, double *cbuf) {
    const int i = blockIdx.x * kLineBlock + threadIdx.x;
    if (i < g.w) lod_col(g, kd, dt, s2, x, cbuf, i);
}

// ---- the first guess, parallel (lines of 16 * 2^k cells) ----------------------------------
//
// The same two sweeps as `lod` (D:370-489): every row's tridiagonal system (cyclic when the grid
// wraps), then every column's, solved by a partition method instead of one Thomas recurrence per
// line (Wang 1981; the Thomas/PCR hybrid of Lászlo, Giles and Appleyard 2016).  A block takes
// one line of n cells, a thread one segment of 16: it eliminates the segment's interior rows
// against the segment's first and last unknowns (forward, then backward, keeping the 14 interior
// rows' coefficients in registers), which leaves for each segment two equations coupling only the
// first and last unknowns of neighbouring segments: a tridiagonal system of 2 n / 16 unknowns
// (cyclic for a wrapping row), solved in LDS by parallel cyclic reduction; the interior unknowns
// follow from it.  The result is the solution of the same systems, equal to the sequential
// recurrences' to rounding (the systems are diagonally dominant).  Columns are solved as the rows
// of the transposed grid.

constexpr int kSeg = 16;

// dst[c * rows + r] = src[r * cols + c], through a 32 x 32 tile in LDS.
__device__ __forceinline__ void transpose_tile(const double *__restrict__ src,
                                               double *__restrict__ dst, int rows, int cols,
                                               int r0, int c0, double (*t)[33]) {
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) t[k][tx] = src[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[(size_t)c * rows + r] = t[tx][k];
    }
}

__global__ void k_transpose(const double *__restrict__ src, double *__restrict__ dst, int rows,
                            int cols) {
    __shared__ double t[32][33];
    transpose_tile(src, dst, rows, cols, blockIdx.y * 32, blockIdx.x * 32, t);
}

// The two reduced equations of a 16-cell segment: F (the segment's first unknown) couples the last
// unknown of the segment before (a[0]) and its own last (c[0]); L couples its own first (a[1])
// and the first of the segment after (c[1]).
struct SegEq {
    double a[2], b[2], c[2], r[2];
};

// One segment of the partition method: kk[0 .. 17] the diffusivity of cells a - 1 .. a + 16
// (has_w / has_e: the faces beyond the segment exist), dts2 = dt / dx^2 (divided once: this
// solver is a first guess and need not repeat the reference's rounding, only its system), xs the
// segment's right-hand side.  Row q of the segment: lo = -f[q], up = -f[q + 1], di = 1 + f[q] +
// f[q + 1].  The interior rows 1 .. 14 are eliminated against x_a and x_b (forward, then
// backward; x_q = de[q] - al[q] x_a - ga[q] x_b), leaving eq.
// seg_reduce with the right-hand side already in d.
__device__ __forceinline__ void seg_reduce_d(const double (&kk)[kSeg + 2], bool has_w,
                                             bool has_e, double dts2, SegEq &eq,
                                             double (&al)[kSeg], double (&ga)[kSeg],
                                             double (&de)[kSeg], const double (&d)[kSeg]) {
    double f[kSeg + 1];
#pragma unroll
    for (int q = 0; q <= kSeg; q++) f[q] = 0.5 * (kk[q] + kk[q + 1]) * dts2;
    if (!has_w) f[0] = 0.0;
    if (!has_e) f[kSeg] = 0.0;
    {
        const double lo = -f[1], up = -f[2], ri = 1.0 / (1.0 + f[1] + f[2]);
        al[1] = lo * ri, ga[1] = up * ri, de[1] = d[1] * ri;
    }
#pragma unroll
    for (int q = 2; q <= kSeg - 2; q++) {
        const double lo = -f[q], up = -f[q + 1], di = 1.0 + f[q] + f[q + 1];
        const double ri = 1.0 / (di - lo * ga[q - 1]);
        al[q] = -lo * al[q - 1] * ri;
        ga[q] = up * ri;
        de[q] = (d[q] - lo * de[q - 1]) * ri;
    }
#pragma unroll
    for (int q = kSeg - 3; q >= 1; q--) {
        de[q] = de[q] - ga[q] * de[q + 1];
        al[q] = al[q] - ga[q] * al[q + 1];
        ga[q] = -ga[q] * ga[q + 1];
    }
    const double lo0 = -f[0], up0 = -f[1], di0 = 1.0 + f[0] + f[1];
    eq.a[0] = lo0;
    eq.b[0] = di0 - up0 * al[1];
    eq.c[0] = -up0 * ga[1];
    eq.r[0] = d[0] - up0 * de[1];
    const double lo15 = -f[kSeg - 1], up15 = -f[kSeg], di15 = 1.0 + f[kSeg - 1] + f[kSeg];
    eq.a[1] = -lo15 * al[kSeg - 2];
    eq.b[1] = di15 - lo15 * ga[kSeg - 2];
    eq.c[1] = up15;
    eq.r[1] = d[kSeg - 1] - lo15 * de[kSeg - 2];
}

__device__ __forceinline__ void seg_reduce(const double (&kk)[kSeg + 2], bool has_w, bool has_e,
                                           double dts2, const double *__restrict__ xs, SegEq &eq,
                                           double (&al)[kSeg], double (&ga)[kSeg],
                                           double (&de)[kSeg], double (&d)[kSeg]) {
#pragma unroll
    for (int q = 0; q < kSeg; q += 2) {
        const double2 v = *reinterpret_cast<const double2 *>(xs + q);
        d[q] = v.x, d[q + 1] = v.y;
    }
    seg_reduce_d(kk, has_w, has_e, dts2, eq, al, ga, de, d);
}

// The segment's 16 values from its end unknowns, into out.
__device__ __forceinline__ void seg_back(double xf, double xl, const double (&al)[kSeg],
                                         const double (&ga)[kSeg], const double (&de)[kSeg],
                                         double *__restrict__ out) {
    double d[kSeg];
    d[0] = xf;
    d[kSeg - 1] = xl;
#pragma unroll
    for (int q = 1; q <= kSeg - 2; q++) d[q] = de[q] - al[q] * xf - ga[q] * xl;
#pragma unroll
    for (int q = 0; q < kSeg; q += 2)
        *reinterpret_cast<double2 *>(out + q) = make_double2(d[q], d[q + 1]);
}

// One line per block (n / 16 threads, a power of two), from src to dst (may be the same array);
// faces from kd along the line, the last one wrapping when `ring`.
template <int MAXT>
__global__ __launch_bounds__(MAXT, 2048 / MAXT) void k_lod_lines(int n, int ring,
                                                    const double *__restrict__ kd,
                                                    const double *__restrict__ src,
                                                    double *__restrict__ dst, double dt,
                                                    double s2) {
    __shared__ double A[2 * MAXT], B[2 * MAXT], Cc[2 * MAXT], R[2 * MAXT], Bi[2 * MAXT];
    const int P = n / kSeg, N = 2 * P, t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * n;
    const double *k = kd + base;
    const double *xs = src + base;
    double *xl = dst + base;
    const int a = t * kSeg, b = a + kSeg - 1;
    // kd[a - 1 .. b + 1] (wrapped for a ring; a missing neighbour gives no face).
    double kk[kSeg + 2];
#pragma unroll
    for (int q = 0; q < kSeg; q += 2) {
        const double2 v = *reinterpret_cast<const double2 *>(k + a + q);
        kk[1 + q] = v.x, kk[2 + q] = v.y;
    }
    const bool has_w = a > 0 || ring, has_e = b + 1 < n || ring;
    kk[0] = has_w ? k[a > 0 ? a - 1 : n - 1] : 0.0;
    kk[kSeg + 1] = has_e ? k[b + 1 < n ? b + 1 : 0] : 0.0;
    double al[kSeg], ga[kSeg], de[kSeg], d[kSeg];
    SegEq eq;
    seg_reduce(kk, has_w, has_e, dt / s2, xs + a, eq, al, ga, de, d);
    A[2 * t] = eq.a[0], B[2 * t] = eq.b[0], Cc[2 * t] = eq.c[0], R[2 * t] = eq.r[0];
    A[2 * t + 1] = eq.a[1], B[2 * t + 1] = eq.b[1], Cc[2 * t + 1] = eq.c[1], R[2 * t + 1] = eq.r[1];
    Bi[2 * t] = 1.0 / B[2 * t];
    Bi[2 * t + 1] = 1.0 / B[2 * t + 1];
    __syncthreads();
    // Parallel cyclic reduction: after the step of stride s every equation couples the unknowns
    // 2s away.  A ring stops at stride N / 2, where both couplings reach the same unknown.
    const int last = ring ? N / 2 : N;
    for (int s = 1; s < last; s <<= 1) {
        double na[2], nb[2], nc[2], nr[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int i = 2 * t + e;
            int im = i - s, ip = i + s;
            double a_m = 0, bi_m = 1, c_m = 0, r_m = 0, a_p = 0, bi_p = 1, c_p = 0, r_p = 0;
            if (ring) {
                im = (im + N) % N, ip = ip % N;
            }
            if (im >= 0) a_m = A[im], bi_m = Bi[im], c_m = Cc[im], r_m = R[im];
            if (ip < N) a_p = A[ip], bi_p = Bi[ip], c_p = Cc[ip], r_p = R[ip];
            const double k1 = A[i] * bi_m, k2 = Cc[i] * bi_p;
            na[e] = -a_m * k1;
            nc[e] = -c_p * k2;
            nb[e] = B[i] - c_m * k1 - a_p * k2;
            nr[e] = R[i] - r_m * k1 - r_p * k2;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int i = 2 * t + e;
            A[i] = na[e], B[i] = nb[e], Cc[i] = nc[e], R[i] = nr[e], Bi[i] = 1.0 / nb[e];
        }
        __syncthreads();
    }
    double xf, xl_;
    {
        double sol[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int i = 2 * t + e;
            if (ring && N > 1) {
                // b_i x_i + (a_i + c_i) x_j = r_i and its partner equation j = i + N/2.
                const int j = (i + N / 2) % N;
                const double g1 = A[i] + Cc[i], g2 = A[j] + Cc[j];
                sol[e] = (R[i] * B[j] - g1 * R[j]) / (B[i] * B[j] - g1 * g2);
            } else {
                sol[e] = R[i] / B[i];
            }
        }
        xf = sol[0], xl_ = sol[1];
    }
    seg_back(xf, xl_, al, ga, de, xl + a);
}

// The wave's LDS writes before this point are seen by its reads after it: one wave's LDS
// accesses take effect in program order; the fence keeps the compiler from moving them across.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One line of n cells (a multiple of 16, 32 <= n <= 1024) by the partition method in one wave,
// for lines of any such length: lane t takes the segment [16 t, 16 t + 16) and reduces it
// (seg_reduce); the N = 2 n / 16 reduced equations go to LDS and are solved by parallel cyclic
// reduction (two equations a lane, ceil(log2 N) steps, coefficients past either end taken as the
// identity), a ring (periodic row) with the Sherman-Morrison correction (Press et al., Numerical
// Recipes, 2nd ed., section 2.7: the corners folded into the first and last diagonals, a second
// right-hand side u = (gamma, 0, .., 0, alpha) reduced alongside); the lanes then recover their
// interiors.  The same system as the reference's row recurrences, solved to rounding.  sm: 6 x 128
// doubles of LDS for this wave alone; every lane of the wave calls it (`on`: the wave has a line),
// as it synchronises the wave.
__device__ __forceinline__ void lod_line_wave(int n, int ring, bool on, const double *__restrict__ k,
                                              const double *__restrict__ src,
                                              double *__restrict__ dst, double dts2,
                                              double *sm) {
    const int lane = threadIdx.x & 63, P = n / kSeg, N = 2 * P;
    double *A = sm, *B = sm + 128, *C = sm + 256, *R = sm + 384, *U = sm + 512, *Bi = sm + 640;
    const bool mine = on && lane < P;
    const int a = lane * kSeg, b = a + kSeg - 1;
    double al[kSeg], ga[kSeg], de[kSeg], d[kSeg];
    if (mine) {
        double kk[kSeg + 2];
#pragma unroll
        for (int q = 0; q < kSeg; q += 2) {
            const double2 v = *reinterpret_cast<const double2 *>(k + a + q);
            kk[1 + q] = v.x, kk[2 + q] = v.y;
        }
        const bool has_w = a > 0 || ring, has_e = b + 1 < n || ring;
        kk[0] = has_w ? k[a > 0 ? a - 1 : n - 1] : 0.0;
        kk[kSeg + 1] = has_e ? k[b + 1 < n ? b + 1 : 0] : 0.0;
        SegEq eq;
        seg_reduce(kk, has_w, has_e, dts2, src + a, eq, al, ga, de, d);
#pragma unroll
        for (int e = 0; e < 2; e++) {
            A[2 * lane + e] = eq.a[e], B[2 * lane + e] = eq.b[e], C[2 * lane + e] = eq.c[e];
            R[2 * lane + e] = eq.r[e];
        }
    }
    wave_lds_sync();
    // Row 0 couples x[N - 1] by beta = A[0], row N - 1 couples x[0] by alpha = C[N - 1] (a ring;
    // dropped for a closed line).
    double beta = 0.0, alpha = 0.0, gamma = 1.0;
    if (on && ring) beta = A[0], alpha = C[N - 1], gamma = -B[0];
    wave_lds_sync();
    if (mine) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int i = 2 * lane + e;
            double ui = 0.0;
            if (i == 0) {
                if (ring) B[0] -= gamma, ui = gamma;
                A[0] = 0.0;
            }
            if (i == N - 1) {
                if (ring) B[N - 1] -= alpha * beta / gamma, ui = alpha;
                C[N - 1] = 0.0;
            }
            U[i] = ui;
            Bi[i] = 1.0 / B[i];
        }
    }
    wave_lds_sync();
    // After the step of stride s every equation couples the unknowns 2 s away; past the ends
    // nothing.
    for (int s = 1; s < N; s <<= 1) {
        double na[2], nb[2], nc[2], nr[2], nu[2];
        if (mine) {
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int i = 2 * lane + e, im = i - s, ip = i + s;
                double a_m = 0.0, bi_m = 1.0, c_m = 0.0, r_m = 0.0, u_m = 0.0;
                double a_p = 0.0, bi_p = 1.0, c_p = 0.0, r_p = 0.0, u_p = 0.0;
                if (im >= 0) a_m = A[im], bi_m = Bi[im], c_m = C[im], r_m = R[im], u_m = U[im];
                if (ip < N) a_p = A[ip], bi_p = Bi[ip], c_p = C[ip], r_p = R[ip], u_p = U[ip];
                const double k1 = A[i] * bi_m, k2 = C[i] * bi_p;
                na[e] = -a_m * k1;
                nc[e] = -c_p * k2;
                nb[e] = B[i] - c_m * k1 - a_p * k2;
                nr[e] = R[i] - r_m * k1 - r_p * k2;
                nu[e] = U[i] - u_m * k1 - u_p * k2;
            }
        }
        wave_lds_sync();
        if (mine) {
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int i = 2 * lane + e;
                A[i] = na[e], B[i] = nb[e], C[i] = nc[e], R[i] = nr[e], U[i] = nu[e];
                Bi[i] = 1.0 / nb[e];
            }
        }
        wave_lds_sync();
    }
    // Every equation stands alone: x = R / B, z = U / B; the ring's correction x - fact z.
    double x[2] = {0.0, 0.0};
    if (mine) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int i = 2 * lane + e;
            x[e] = R[i] * Bi[i];
            if (ring) A[i] = x[e], C[i] = U[i] * Bi[i];
        }
    }
    if (ring) {
        wave_lds_sync();
        if (mine) {
            const double fact = (A[0] + beta * A[N - 1] / gamma) /
                                (1.0 + C[0] + beta * C[N - 1] / gamma);
#pragma unroll
            for (int e = 0; e < 2; e++) x[e] -= fact * C[2 * lane + e];
        }
    }
    if (mine) seg_back(x[0], x[1], al, ga, de, dst + a);
}

// ---- the first guess's column sweep by the partition method, in place --------------------
//
// The columns' systems (D:436-488) solved like the rows' (segments of 16 rows, each reduced to two
// equations coupling its first and last unknowns), with no transposes: a wave takes 64 adjacent
// columns of one segment, so every load and store of a row is one coalesced access.  Three passes:
// col_reduce writes each segment's two equations (k = 2 s, 2 s + 1 of column i) to red, as four
// N x w arrays a, b, c, r (N = 2 h / 16; entry k w + i); col_thomas solves each column's N
// equations (a closed line: Thomas's recurrences, one thread a column), leaving the unknowns in
// r; col_back repeats the segment's elimination (the same arithmetic, so the same coefficients)
// and writes the segment's 16 values.  h is a multiple of 16 of at least 32.

// Column i's segment s: kd of rows 16 s - 1 .. 16 s + 16 (zero beyond the ends) and x of its rows.
__device__ __forceinline__ void col_seg_load(int w, int h, const double *__restrict__ kd,
                                             const double *__restrict__ x, int i, int s,
                                             double (&kk)[kSeg + 2], double (&d)[kSeg]) {
    const size_t c0 = (size_t)s * kSeg * w + i;
#pragma unroll
    for (int q = 0; q < kSeg; q++) kk[1 + q] = gld(kd + c0 + (size_t)q * w);
#pragma unroll
    for (int q = 0; q < kSeg; q++) d[q] = gld(x + c0 + (size_t)q * w);
    kk[0] = s > 0 ? gld(kd + c0 - w) : 0.0;
    kk[kSeg + 1] = (s + 1) * kSeg < h ? gld(kd + c0 + (size_t)kSeg * w) : 0.0;
}

__device__ __forceinline__ void col_reduce(int w, int h, const double *__restrict__ kd,
                                           const double *__restrict__ x, double dts2,
                                           double *__restrict__ red, int i, int s) {
    const int N = 2 * (h / kSeg);
    const size_t Nw = (size_t)N * w;
    double kk[kSeg + 2], d[kSeg], al[kSeg], ga[kSeg], de[kSeg];
    col_seg_load(w, h, kd, x, i, s, kk, d);
    SegEq eq;
    seg_reduce_d(kk, s > 0, (s + 1) * kSeg < h, dts2, eq, al, ga, de, d);
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const size_t o = (size_t)(2 * s + e) * w + i;
        gst(red + o, eq.a[e]);
        gst(red + Nw + o, eq.b[e]);
        gst(red + 2 * Nw + o, eq.c[e]);
        gst(red + 3 * Nw + o, eq.r[e]);
    }
}

// Thomas's recurrences on column i of N tridiagonal equations stored as four N x w arrays a, b,
// c, r (entry k w + i; a closed line), the unknowns left in r.  The loads of 16 steps go out
// together ahead of the recurrence's chain (a column's equations lie w apart: each load of a
// step is a separate access).
__device__ __forceinline__ void tri_thomas(int w, int N, double *__restrict__ red, int i) {
    constexpr int U = 16;
    const size_t Nw = (size_t)N * w;
    const double *A = red, *B = red + Nw;
    double *C = red + 2 * Nw, *R = red + 3 * Nw;
    double cp = 0.0, rp = 0.0;
    for (int k0 = 0; k0 < N; k0 += U) {
        double a[U], b[U], c[U], r[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + u;
            const size_t o = (size_t)(k < N ? k : N - 1) * w + i;
            a[u] = k > 0 ? gld(A + o) : 0.0, b[u] = gld(B + o), c[u] = gld(C + o),
            r[u] = gld(R + o);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + u;
            if (k >= N) break;
            const double inv = 1.0 / (b[u] - a[u] * cp);
            cp = c[u] * inv;
            rp = (r[u] - a[u] * rp) * inv;
            const size_t o = (size_t)k * w + i;
            gst(C + o, cp), gst(R + o, rp);
        }
    }
    double xn = rp;
    for (int k1 = N - 2; k1 >= 0; k1 -= U) {
        double c[U], r[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k1 - u;
            const size_t o = (size_t)(k >= 0 ? k : 0) * w + i;
            c[u] = gld(C + o), r[u] = gld(R + o);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k1 - u;
            if (k < 0) break;
            xn = r[u] - c[u] * xn;
            gst(R + (size_t)k * w + i, xn);
        }
    }
}

__device__ __forceinline__ void col_thomas(int w, int h, double *__restrict__ red, int i) {
    tri_thomas(w, 2 * (h / kSeg), red, i);
}

// A second partition for long columns (N = 2 h / 16 a multiple of 16 of at least 32): the
// reduced system's equations in groups of 16, each group's interior eliminated against its first
// and last unknowns (the same recurrences as seg_reduce with general coefficients: row q reads
// a[q] x[q - 1] + b[q] x[q] + c[q] x[q + 1] = r[q]), leaving 2 N / 16 equations (red2, the same
// four-array layout), solved by tri_thomas; the groups' interiors follow.
__device__ __forceinline__ void tri_group_load(int w, int N, const double *__restrict__ red, int i,
                                               int grp, double (&a)[kSeg], double (&b)[kSeg],
                                               double (&c)[kSeg], double (&r)[kSeg]) {
    const size_t Nw = (size_t)N * w, o0 = (size_t)grp * kSeg * w + i;
#pragma unroll
    for (int q = 0; q < kSeg; q++) {
        const size_t o = o0 + (size_t)q * w;
        a[q] = gld(red + o), b[q] = gld(red + Nw + o), c[q] = gld(red + 2 * Nw + o),
        r[q] = gld(red + 3 * Nw + o);
    }
}
__device__ __forceinline__ void tri_group_reduce(const double (&a)[kSeg], const double (&b)[kSeg],
                                                 const double (&c)[kSeg], const double (&r)[kSeg],
                                                 SegEq &eq, double (&al)[kSeg],
                                                 double (&ga)[kSeg], double (&de)[kSeg]) {
    {
        const double ri = 1.0 / b[1];
        al[1] = a[1] * ri, ga[1] = c[1] * ri, de[1] = r[1] * ri;
    }
#pragma unroll
    for (int q = 2; q <= kSeg - 2; q++) {
        const double ri = 1.0 / (b[q] - a[q] * ga[q - 1]);
        al[q] = -a[q] * al[q - 1] * ri;
        ga[q] = c[q] * ri;
        de[q] = (r[q] - a[q] * de[q - 1]) * ri;
    }
#pragma unroll
    for (int q = kSeg - 3; q >= 1; q--) {
        de[q] = de[q] - ga[q] * de[q + 1];
        al[q] = al[q] - ga[q] * al[q + 1];
        ga[q] = -ga[q] * ga[q + 1];
    }
    eq.a[0] = a[0];
    eq.b[0] = b[0] - c[0] * al[1];
    eq.c[0] = -c[0] * ga[1];
    eq.r[0] = r[0] - c[0] * de[1];
    eq.a[1] = -a[kSeg - 1] * al[kSeg - 2];
    eq.b[1] = b[kSeg - 1] - a[kSeg - 1] * ga[kSeg - 2];
    eq.c[1] = c[kSeg - 1];
    eq.r[1] = r[kSeg - 1] - a[kSeg - 1] * de[kSeg - 2];
}
__device__ __forceinline__ void tri_group_fwd(int w, int N, const double *__restrict__ red,
                                              double *__restrict__ red2, int i, int grp) {
    double a[kSeg], b[kSeg], c[kSeg], r[kSeg], al[kSeg], ga[kSeg], de[kSeg];
    tri_group_load(w, N, red, i, grp, a, b, c, r);
    SegEq eq;
    tri_group_reduce(a, b, c, r, eq, al, ga, de);
    const size_t N2w = (size_t)(2 * (N / kSeg)) * w;
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const size_t o = (size_t)(2 * grp + e) * w + i;
        gst(red2 + o, eq.a[e]);
        gst(red2 + N2w + o, eq.b[e]);
        gst(red2 + 2 * N2w + o, eq.c[e]);
        gst(red2 + 3 * N2w + o, eq.r[e]);
    }
}
__device__ __forceinline__ void tri_group_back(int w, int N, double *__restrict__ red,
                                               const double *__restrict__ red2, int i, int grp) {
    double a[kSeg], b[kSeg], c[kSeg], r[kSeg], al[kSeg], ga[kSeg], de[kSeg];
    tri_group_load(w, N, red, i, grp, a, b, c, r);
    SegEq eq;
    tri_group_reduce(a, b, c, r, eq, al, ga, de);
    const size_t N2w = (size_t)(2 * (N / kSeg)) * w;
    const double *R2 = red2 + 3 * N2w;
    const double xf = R2[(size_t)(2 * grp) * w + i], xl = R2[(size_t)(2 * grp + 1) * w + i];
    double *R = red + 3 * (size_t)N * w;
    const size_t o0 = (size_t)grp * kSeg * w + i;
    gst(R + o0, xf);
#pragma unroll
    for (int q = 1; q <= kSeg - 2; q++) gst(R + o0 + (size_t)q * w, de[q] - al[q] * xf - ga[q] * xl);
    gst(R + o0 + (size_t)(kSeg - 1) * w, xl);
}

__device__ __forceinline__ void col_back(int w, int h, const double *__restrict__ kd,
                                         double *__restrict__ x, double dts2,
                                         const double *__restrict__ red, int i, int s) {
    const int N = 2 * (h / kSeg);
    const double *R = red + 3 * (size_t)N * w;
    double kk[kSeg + 2], d[kSeg], al[kSeg], ga[kSeg], de[kSeg];
    col_seg_load(w, h, kd, x, i, s, kk, d);
    SegEq eq;
    seg_reduce_d(kk, s > 0, (s + 1) * kSeg < h, dts2, eq, al, ga, de, d);
    const double xf = R[(size_t)(2 * s) * w + i], xl = R[(size_t)(2 * s + 1) * w + i];
    const size_t c0 = (size_t)s * kSeg * w + i;
    gst(x + c0, xf);
#pragma unroll
    for (int q = 1; q <= kSeg - 2; q++) gst(x + c0 + (size_t)q * w, de[q] - al[q] * xf - ga[q] * xl);
    gst(x + c0 + (size_t)(kSeg - 1) * w, xl);
}

// The single grid's passes: blocks of 64 columns x 4 segments (grid ceil(w / 64) x ceil(P / 4)),
// and one thread a column.
__global__ __launch_bounds__(kBlock) void k_col_reduce(int w, int h, const double *__restrict__ kd,
                                                       const double *__restrict__ x, double dts2,
                                                       double *__restrict__ red) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), s = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < w && s < h / kSeg) col_reduce(w, h, kd, x, dts2, red, i, s);
}
__global__ __launch_bounds__(64) void k_col_thomas(int w, int h, double *__restrict__ red) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < w) col_thomas(w, h, red, i);
}
// The second partition's passes (mode 0: the groups' equations into red2; 1: the groups'
// values into red's r), blocks of 64 columns x 4 groups; and Thomas on red2, one thread a column.
__global__ __launch_bounds__(kBlock) void k_tri_group(int w, int N, double *__restrict__ red,
                                                      double *__restrict__ red2, int mode) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), grp = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= w || grp >= N / kSeg) return;
    if (mode == 0)
        tri_group_fwd(w, N, red, red2, i, grp);
    else
        tri_group_back(w, N, red, red2, i, grp);
}
__global__ __launch_bounds__(64) void k_tri_thomas(int w, int N, double *__restrict__ red) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < w) tri_thomas(w, N, red, i);
}
__global__ __launch_bounds__(kBlock) void k_col_back(int w, int h, const double *__restrict__ kd,
                                                     double *__restrict__ x, double dts2,
                                                     const double *__restrict__ red) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), s = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < w && s < h / kSeg) col_back(w, h, kd, x, dts2, red, i, s);
}

// ---- sums in the reference's order (DTD_SUMS_STRIPS) ----------------------------------------
//
// The reference adds each pass's products per strip of STRIP_ROWS rows, in row-major order from
// -0.0 (`sum_zero`, D:507-510), then the strip sums in strip order (MG:596-619, 652-694, 703-723;
// `Level::sweep`'s r.z, MG:314-318), and the put-back's two sums per row, then the rows in order
// (MG:740-755).  These kernels keep that order exactly: a strip's (a row's) sum is one lane's chain
// of adds.  They serve the mode that reproduces the reference's iterates bit for bit; the fast
// path sums per block instead.

constexpr int kStripRows = 8;  // STRIP_ROWS (D:502)

// out[s] = the strip's sum of x[c] * y[c] (y widened to f64), c over the strip's cells in order.
template <typename TY>
__global__ void k_strip_sums(const double *__restrict__ x, const TY *__restrict__ y, long n,
                             long strip_cells, int nstrips, double *__restrict__ out,
                             const CgState *__restrict__ st) {
    if (st && st->done) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstrips) return;
    const long c0 = (long)s * strip_cells, c1 = min(c0 + strip_cells, n);
    constexpr int U = 8;
    double a = -0.0;
    long c = c0;
    for (; c + U <= c1; c += U) {
        double xv[U], yv[U];
#pragma unroll
        for (int k = 0; k < U; k++) xv[k] = x[c + k], yv[k] = (double)y[c + k];
#pragma unroll
        for (int k = 0; k < U; k++) a += xv[k] * yv[k];
    }
    for (; c < c1; c++) a += x[c] * (double)y[c];
    out[s] = a;
}

// The strip sums added in order from -0.0 (and b.b's for the first residual), then the step;
// the stop flag into host memory when given.
template <int STEP>
__global__ void k_strips_step(const double *__restrict__ sums, int nstrips,
                              const double *__restrict__ bb_sums, CgState *st, int *flag) {
    if (threadIdx.x != 0) return;
    if (STEP == kStepRR0 || !st->done) {
        double t = -0.0;
        for (int k = 0; k < nstrips; k++) t += sums[k];
        double bb = 0.0;
        if (STEP == kStepRR0 && bb_sums) {
            bb = -0.0;
            for (int k = 0; k < nstrips; k++) bb += bb_sums[k];
        }
        cg_step_apply<STEP>(t, bb, st);
    }
    if (flag) *reinterpret_cast<volatile int *>(flag) = st->done;
}

// r = b - A x in f64 (the stored faces, `apply_row`'s order) and its f32 copy (MG:596-619, less
// the sum); x takes xin first when they differ.
__global__ void k_resid_es(Grid g, const double2 *__restrict__ es, const double *__restrict__ b,
                           const double *__restrict__ x, double *__restrict__ r,
                           float *__restrict__ r32) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const int j = c / g.w, i = c - j * g.w;
    const double rv = b[c] - apply_es(g, es, c, i, j, [&](int k) { return x[k]; });
    r[c] = rv;
    r32[c] = (float)rv;
}

// The direction pass less A p (MG:633-660): the pending x += alpha p_old, p = z + beta p_old (z
// on the first iteration).
__global__ void k_dir_cells(int n, const float *__restrict__ z, const double *__restrict__ pold,
                            double *__restrict__ pnew, double *__restrict__ x,
                            const CgState *__restrict__ st) {
    if (st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    if (st->have_pending) x[c] += st->alpha * pold[c];
    pnew[c] = st->have_beta ? (double)z[c] + st->beta * pold[c] : (double)z[c];
}

// After the loop: x' = x + alpha p when an update is pending (MG:734-736), p the direction of the
// last iteration (p0 after an odd count), into xout.
__global__ void k_last_update(int n, const double *__restrict__ x, const double *__restrict__ p0,
                              const double *__restrict__ p1, double *__restrict__ xout,
                              const CgState *__restrict__ st) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const double *p = (st->its & 1) ? p0 : p1;
    xout[c] = st->have_pending ? x[c] + st->alpha * p[c] : x[c];
}

// The put-back's row sums (MG:740-751): per row, sum (b - A x) and sum (diag - 1) in column
// order from -0.0; one lane a row.
__global__ void k_putback_rows(Grid g, const double2 *__restrict__ es, const double *__restrict__ b,
                               const double *__restrict__ ax, double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.h) return;
    double l = -0.0, d1 = -0.0;
    for (int i = 0; i < g.w; i++) {
        const int c = j * g.w + i;
        l += b[c] - ax[c];
        d1 += diag_es(g, es, c, i, j) - 1.0;
    }
    out[2 * (size_t)j] = l;
    out[2 * (size_t)j + 1] = d1;
}

// The rows folded in order from (-0.0, -0.0) (MG:752-755): the put-back's lost and weight.
__global__ void k_putback_fold(const double *__restrict__ rows, int h, CgState *st) {
    if (threadIdx.x != 0) return;
    double l = -0.0, d1 = -0.0;
    for (int j = 0; j < h; j++) {
        l = l + rows[2 * (size_t)j];
        d1 = d1 + rows[2 * (size_t)j + 1];
    }
    st->lost = l;
    st->weight = d1;
}

}  // namespace dtd
