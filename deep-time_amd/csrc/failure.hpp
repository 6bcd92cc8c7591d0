// The submarine slope failure (`marine::slope_failure`) on the device: the caller's step after
// the diffusion solve, on the same resident heights.  Included once, by dtdiff.hip.
//
// Reference files (relative to the deep-time repository root):
//   MAR  crates/deeptime-core/src/surface/marine.rs
//   GRID crates/deeptime-core/src/surface/grid.rs
//   PROJ crates/deeptime-format/src/projection.rs
//
// The reference's CPU path (MAR:292-580) keeps an active set of the cells near the last
// iteration's movers; it equals the full sweep over every cell (MAR:595-655) bit for bit.  The
// device runs the full sweep: each iteration is a share pass and a change pass over every cell,
// each cell's value formed by the reference's expression in its order (D8 neighbours in the fixed
// order, dry or missing ones skipped), in f64, so the heights and the sediment are the
// reference's bit for bit.  Only the metres moved, a diagnostic sum over the movers, is added in
// another order (per block, then the blocks in order; the reference adds cell by cell), which
// makes it equal to rounding.  An iteration that moves nothing ends the loop, as in the
// reference: its step marks the state done and every later launch returns at once, so the host
// launches all the iterations without waiting for any of them.
#pragma once

#include "context.hpp"
#include "kernels.hpp"

namespace dtd {

struct FailState {
    double moved;  // metres moved so far (sum |dz| / 2)
    int done;      // an iteration moved nothing
    int its;       // iterations run, the one that moved nothing included
    int result;    // the buffer holding the heights: 0 the caller's, 1 the scratch copy
    int pad;
};

struct FGrid {
    int w, h, wrap, n;
    double crit_side, crit_diag;  // S_c * dx and S_c * (dx * sqrt 2) (`step_len`, GRID:65-71)
    double rate, sea_level;
};

// The neighbour of (i, j) at D8 offset k (E, NE, N, NW, W, SW, S, SE; "north" is row j - 1,
// GRID:21-31) by the projection's rule (PROJ:350-368): rows clamp, columns wrap when the grid
// does, and the cell itself is no neighbour (GRID:46-60).  -1 for none.
__device__ __forceinline__ int d8_neighbour(const FGrid &g, int c, int i, int j, int k) {
    constexpr int di[8] = {1, 1, 0, -1, -1, -1, 0, 1};
    constexpr int dj[8] = {0, -1, -1, -1, 0, 1, 1, 1};
    const int nj = j + dj[k];
    if (nj < 0 || nj >= g.h) return -1;
    int ni = i + di[k];
    if (ni < 0 || ni >= g.w) {
        if (!g.wrap) return -1;
        ni = ni < 0 ? ni + g.w : ni - g.w;  // |di| = 1: one turn (a 1-wide grid: the cell)
    }
    const int m = nj * g.w + ni;
    return m == c ? -1 : m;
}

__device__ __forceinline__ bool fail_wet(const FGrid &g, const unsigned char *sea,
                                         const double *h, int c) {
    return sea[c] && h[c] < g.sea_level;
}

__device__ __forceinline__ double fail_crit(const FGrid &g, int k) {
    return (k & 1) ? g.crit_diag : g.crit_side;
}

// Each cell's share of what it would give (the `share` closure, MAR:346-366):
// min(avail / give, 1) with give = sum over wet neighbours of rate * max(h_c - h_m - S_c d, 0);
// 0 for a dry cell, one without sediment, or one that gives nothing.
__global__ __launch_bounds__(kBlock) void k_fail_share(FGrid g, const unsigned char *__restrict__ sea,
                                                      const double *__restrict__ h,
                                                      const double *__restrict__ avail,
                                                      double *__restrict__ scale,
                                                      const FailState *__restrict__ st) {
    if (st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const int j = c / g.w, i = c - j * g.w;
    double s = 0.0;
    const double hc = h[c];
    if (sea[c] && hc < g.sea_level && avail[c] > 0.0) {
        double give = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int m = d8_neighbour(g, c, i, j, k);
            if (m >= 0 && fail_wet(g, sea, h, m))
                give += g.rate * fmax(hc - h[m] - fail_crit(g, k), 0.0);
        }
        if (give > 0.0) s = fmin(avail[c] / give, 1.0);
    }
    scale[c] = s;
}

// Each cell's change (the `change` closure, MAR:367-383): sum over wet neighbours of
// rate * (in * share_m - out * share_c); a mover (dz != 0) takes h += dz and
// avail = max(avail + dz, 0) (MAR:640-648).  The heights go from hin to hout (every cell, so the
// buffers alternate); parts gets the block's (sum |dz| / 2, movers).
__global__ __launch_bounds__(kBlock) void k_fail_change(FGrid g,
                                                       const unsigned char *__restrict__ sea,
                                                       const double *__restrict__ hin,
                                                       double *__restrict__ hout,
                                                       double *__restrict__ avail,
                                                       const double *__restrict__ scale,
                                                       double *__restrict__ parts,
                                                       const FailState *__restrict__ st) {
    __shared__ double lds[kBlock / 64];
    if (st->done) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double half = 0.0, mover = 0.0;
    if (c < g.n) {
        const int j = c / g.w, i = c - j * g.w;
        const double hc = hin[c];
        double dz = 0.0;
        if (sea[c] && hc < g.sea_level) {
            const double sc = scale[c];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int m = d8_neighbour(g, c, i, j, k);
                if (m < 0 || !fail_wet(g, sea, hin, m)) continue;
                const double hm = hin[m], ck = fail_crit(g, k);
                const double out = fmax(hc - hm - ck, 0.0);
                const double inn = fmax(hm - hc - ck, 0.0);
                dz += g.rate * (inn * scale[m] - out * sc);
            }
        }
        if (dz != 0.0) {
            hout[c] = hc + dz;
            avail[c] = fmax(avail[c] + dz, 0.0);
            half = fabs(dz) * 0.5;
            mover = 1.0;
        } else {
            hout[c] = hc;
        }
    }
    block_pair(make_double2(half, mover), parts, blockIdx.x, lds);
}

// The iteration's sums (the blocks in order): the metres moved, and the stop when nothing moved
// (MAR:649-652).
__global__ void k_fail_step(const double *__restrict__ parts, int nparts, FailState *st) {
    __shared__ double lds[kReduceBlock / 64];
    if (st->done) return;
    const double moved = reduce_parts(parts, nparts, 2, 0, lds);
    const double movers = reduce_parts(parts, nparts, 2, 1, lds);
    if (threadIdx.x == 0) {
        st->moved += moved;
        st->its += 1;
        st->result ^= 1;
        if (movers == 0.0) st->done = 1;
    }
}

__global__ void k_fail_init(FailState *st) {
    if (threadIdx.x == 0) *st = FailState{0.0, 0, 0, 0, 0};
}

// The heights back into the caller's array when the last iteration left them in the scratch.
__global__ __launch_bounds__(kBlock) void k_fail_final(int n, const double *__restrict__ scratch,
                                                      double *__restrict__ h,
                                                      const FailState *__restrict__ st) {
    if (st->result == 0) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) h[c] = scratch[c];
}

// ---- host side -------------------------------------------------------------------------------

struct FailWs {
    size_t cap = 0;  // cells
    double *h2 = nullptr, *scale = nullptr, *parts = nullptr;
    FailState *st = nullptr;
    FailState *h_st = nullptr;  // pinned
    // Device copies of a host call's arrays (dtd_slope_failure), made on its first use.
    double *hd = nullptr, *ad = nullptr;
    unsigned char *sd = nullptr;
};

void free_fail(dtd_ctx *ctx) {
    FailWs *f = ctx ? ctx->fail : nullptr;
    if (!f) return;
    (void)hipFree(f->h2);
    (void)hipFree(f->scale);
    (void)hipFree(f->parts);
    (void)hipFree(f->st);
    (void)hipFree(f->hd);
    (void)hipFree(f->ad);
    (void)hipFree(f->sd);
    if (f->h_st) (void)hipHostFree(f->h_st);
    delete f;
    ctx->fail = nullptr;
}

static FailWs &fit_fail(dtd_ctx *ctx, size_t n, bool host) {
    if (ctx->fail && ctx->fail->cap >= n && (!host || ctx->fail->hd)) return *ctx->fail;
    free_fail(ctx);
    FailWs *f = new FailWs();
    ctx->fail = f;
    f->cap = n;
    HIPCHK(hipMalloc(&f->h2, n * sizeof(double)));
    HIPCHK(hipMalloc(&f->scale, n * sizeof(double)));
    HIPCHK(hipMalloc(&f->parts, 2 * (size_t)nblocks((long)n) * sizeof(double)));
    HIPCHK(hipMalloc(&f->st, sizeof(FailState)));
    HIPCHK(hipHostMalloc((void **)&f->h_st, sizeof(FailState), hipHostMallocDefault));
    if (host) {
        HIPCHK(hipMalloc(&f->hd, n * sizeof(double)));
        HIPCHK(hipMalloc(&f->ad, n * sizeof(double)));
        HIPCHK(hipMalloc(&f->sd, n));
    }
    return *f;
}

// The slope failure on device arrays: h (in/out), avail (in/out), sea (0/1 a cell).  Returns the
// metres moved; *its gets the iterations run.
static double slope_failure(dtd_ctx *ctx, FailWs &f, size_t w, size_t h, double dx, int wrap_x,
                            const unsigned char *sea, double sea_level, double critical_slope,
                            double rate, unsigned iterations, double *hgt, double *avail,
                            unsigned *its) {
    const size_t n = w * h;
    hipStream_t s = ctx->stream;
    // The reference's critical drops: S_c * step_len(k), step_len = dx or dx * SQRT_2.
    const double sqrt2 = 1.4142135623730951;
    const FGrid g{(int)w, (int)h, wrap_x ? 1 : 0, (int)n, critical_slope * dx,
                  critical_slope * (dx * sqrt2), rate, sea_level};
    const int nb = nblocks((long)n);
    k_fail_init<<<1, 64, 0, s>>>(f.st);
    LAUNCH_CHECK();
    for (unsigned it = 0; it < iterations; it++) {
        double *hin = it % 2 == 0 ? hgt : f.h2, *hout = it % 2 == 0 ? f.h2 : hgt;
        k_fail_share<<<nb, kBlock, 0, s>>>(g, sea, hin, avail, f.scale, f.st);
        LAUNCH_CHECK();
        k_fail_change<<<nb, kBlock, 0, s>>>(g, sea, hin, hout, avail, f.scale, f.parts, f.st);
        LAUNCH_CHECK();
        k_fail_step<<<1, kReduceBlock, 0, s>>>(f.parts, nb, f.st);
        LAUNCH_CHECK();
    }
    k_fail_final<<<nb, kBlock, 0, s>>>((int)n, f.h2, hgt, f.st);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(f.h_st, f.st, sizeof(FailState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (its) *its = (unsigned)f.h_st->its;
    return f.h_st->moved;
}

}  // namespace dtd
