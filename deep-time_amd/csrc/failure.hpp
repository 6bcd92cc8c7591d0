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
// device runs the full sweep, one tiled pass an iteration (k_fail_sweep), each cell's value formed
// by the reference's expression in its order (D8 neighbours in the fixed order, dry or missing
// ones skipped), in f64, so the heights and the sediment are the reference's bit for bit.  Only
// the metres moved, a diagnostic sum over the movers, is added in another order (per block, then
// the blocks in order; the reference adds cell by cell), which makes it equal to rounding.  An
// iteration that moves nothing ends the loop, as in the reference: its step marks the state done
// and every later launch returns at once, so the host launches all the iterations without waiting
// for any of them.
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

__device__ __forceinline__ double fail_crit(const FGrid &g, int k) {
    return (k & 1) ? g.crit_diag : g.crit_side;
}

// One iteration in one pass over 64 x 16 tiles.  A block loads the heights and the wet flags of
// its tile with a 2-cell halo into LDS (a position where the grid has no cell is dry; columns wrap
// where the grid wraps), forms the shares of the tile and a 1-cell ring (the
// `share` closure, MAR:346-366: min(avail / give, 1), give = sum over wet neighbours of
// rate * max(h_c - h_m - S_c d, 0)), then each tile cell's change (the `change` closure,
// MAR:367-383: sum over wet neighbours of rate * (in * share_m - out * share_c)).  A ring cell's
// share is formed again by every block that needs it, from the same inputs, so it has the same
// bits.  Heights and sediment go from (hin, ain) to (hout, aout), every cell (the buffers
// alternate, so no block reads a value another block of the same pass wrote); a mover
// (dz != 0) takes h + dz and max(avail + dz, 0) (MAR:640-648).  parts gets the block's
// (sum |dz| / 2, movers).
constexpr int kFailTX = 64, kFailTY = 16;
constexpr int kFailRX = kFailTX + 4, kFailRY = kFailTY + 4;  // heights: 2-cell halo
constexpr int kFailSX = kFailTX + 2, kFailSY = kFailTY + 2;  // shares: 1-cell ring

__global__ __launch_bounds__(kBlock) void k_fail_sweep(FGrid g,
                                                      const unsigned char *__restrict__ sea,
                                                      const double *__restrict__ hin,
                                                      double *__restrict__ hout,
                                                      const double *__restrict__ ain,
                                                      double *__restrict__ aout,
                                                      double *__restrict__ parts,
                                                      const FailState *__restrict__ st) {
    __shared__ double hs[kFailRY][kFailRX];
    __shared__ unsigned char wet[kFailRY][kFailRX];  // 0 also where the grid has no cell
    __shared__ double sc[kFailSY][kFailSX];
    __shared__ double as[kFailSY][kFailSX];  // sediment of the share region
    __shared__ double lds[kBlock / 64];
    constexpr int di[8] = {1, 1, 0, -1, -1, -1, 0, 1};
    constexpr int dj[8] = {0, -1, -1, -1, 0, 1, 1, 1};
    if (st->done) return;
    const int ntx = (g.w + kFailTX - 1) / kFailTX;
    const int i0 = (blockIdx.x % ntx) * kFailTX, j0 = (blockIdx.x / ntx) * kFailTY;
    // The region: rows clamp, columns wrap when the grid does (PROJ:350-368).  Every position's
    // loads are issued before any is used.
    constexpr int kPer = (kFailRY * kFailRX + kBlock - 1) / kBlock;
    int ms[kPer];
    double hv[kPer], av[kPer];
    unsigned char sv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const int q = threadIdx.x + u * kBlock;
        const int ly = q / kFailRX, lx = q - ly * kFailRX;
        const int gj = j0 + ly - 2;
        int gi = i0 + lx - 2, m = -1;
        if (q < kFailRY * kFailRX && gj >= 0 && gj < g.h) {
            if (gi >= 0 && gi < g.w) {
                m = gj * g.w + gi;
            } else if (g.wrap) {
                gi %= g.w;
                if (gi < 0) gi += g.w;
                m = gj * g.w + gi;
            }
        }
        ms[u] = m;
        const bool ring = ly >= 1 && ly <= kFailSY && lx >= 1 && lx <= kFailSX;
        hv[u] = m >= 0 ? hin[m] : 0.0;
        sv[u] = m >= 0 ? sea[m] : 0;
        av[u] = m >= 0 && ring ? ain[m] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const int q = threadIdx.x + u * kBlock;
        if (q >= kFailRY * kFailRX) break;
        const int ly = q / kFailRX, lx = q - ly * kFailRX;
        hs[ly][lx] = hv[u];
        wet[ly][lx] = ms[u] >= 0 && sv[u] && hv[u] < g.sea_level;
        if (ly >= 1 && ly <= kFailSY && lx >= 1 && lx <= kFailSX) as[ly - 1][lx - 1] = av[u];
    }
    __syncthreads();
    // A neighbour exists when the grid has it and it is not the cell itself (GRID:46-60): only on
    // a wrapping grid one cell wide are the east and west neighbours the cell itself.
    const bool self_ew = g.wrap && g.w == 1;
    for (int q = threadIdx.x; q < kFailSY * kFailSX; q += kBlock) {
        const int sy = q / kFailSX, sx = q - sy * kFailSX;
        const int ly = sy + 1, lx = sx + 1;
        double s = 0.0;
        if (wet[ly][lx]) {
            const double av = as[sy][sx];
            if (av > 0.0) {
                const double hc = hs[ly][lx];
                double give = 0.0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int ny = ly + dj[k], nx = lx + di[k];
                    if ((k & 3) == 0 && self_ew) continue;
                    if (wet[ny][nx])
                        give += g.rate * fmax(hc - hs[ny][nx] - fail_crit(g, k), 0.0);
                }
                if (give > 0.0) s = fmin(av / give, 1.0);
            }
        }
        sc[sy][sx] = s;
    }
    __syncthreads();
    double half = 0.0, movers = 0.0;
    for (int q = threadIdx.x; q < kFailTX * kFailTY; q += kBlock) {
        const int ty = q / kFailTX, tx = q - ty * kFailTX;
        if (i0 + tx >= g.w || j0 + ty >= g.h) continue;
        const int ly = ty + 2, lx = tx + 2, c = (j0 + ty) * g.w + i0 + tx;
        const double hc = hs[ly][lx];
        double dz = 0.0;
        if (wet[ly][lx]) {
            const double s0 = sc[ly - 1][lx - 1];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int ny = ly + dj[k], nx = lx + di[k];
                if (((k & 3) == 0 && self_ew) || !wet[ny][nx]) continue;
                const double hm = hs[ny][nx], ck = fail_crit(g, k);
                const double out = fmax(hc - hm - ck, 0.0);
                const double inn = fmax(hm - hc - ck, 0.0);
                dz += g.rate * (inn * sc[ny - 1][nx - 1] - out * s0);
            }
        }
        const double av = as[ly - 1][lx - 1];
        if (dz != 0.0) {
            hout[c] = hc + dz;
            aout[c] = fmax(av + dz, 0.0);
            half += fabs(dz) * 0.5;
            movers += 1.0;
        } else {
            hout[c] = hc;
            aout[c] = av;
        }
    }
    block_pair(make_double2(half, movers), parts, blockIdx.x, lds);
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

// The heights and the sediment back into the caller's arrays when the last iteration left them
// in the scratch copies.
__global__ __launch_bounds__(kBlock) void k_fail_final(int n, const double *__restrict__ h2,
                                                      const double *__restrict__ a2,
                                                      double *__restrict__ h,
                                                      double *__restrict__ avail,
                                                      const FailState *__restrict__ st) {
    if (st->result == 0) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) h[c] = h2[c], avail[c] = a2[c];
}

// ---- host side -------------------------------------------------------------------------------

struct FailWs {
    size_t cap = 0;  // cells
    long tiles = 0;  // partial sums
    double *h2 = nullptr, *a2 = nullptr, *parts = nullptr;
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
    (void)hipFree(f->a2);
    (void)hipFree(f->parts);
    (void)hipFree(f->st);
    (void)hipFree(f->hd);
    (void)hipFree(f->ad);
    (void)hipFree(f->sd);
    if (f->h_st) (void)hipHostFree(f->h_st);
    delete f;
    ctx->fail = nullptr;
}

static long fail_tiles(size_t w, size_t h) {
    return (long)((w + kFailTX - 1) / kFailTX) * (long)((h + kFailTY - 1) / kFailTY);
}

static FailWs &fit_fail(dtd_ctx *ctx, size_t w, size_t h, bool host) {
    const size_t n = w * h;
    if (ctx->fail && ctx->fail->cap >= n && ctx->fail->tiles >= fail_tiles(w, h) &&
        (!host || ctx->fail->hd))
        return *ctx->fail;
    free_fail(ctx);
    FailWs *f = new FailWs();
    ctx->fail = f;
    f->cap = n;
    f->tiles = fail_tiles(w, h);
    HIPCHK(hipMalloc(&f->h2, n * sizeof(double)));
    HIPCHK(hipMalloc(&f->a2, n * sizeof(double)));
    HIPCHK(hipMalloc(&f->parts, 2 * (size_t)fail_tiles(w, h) * sizeof(double)));
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
    const int nt = (int)fail_tiles(w, h);
    k_fail_init<<<1, 64, 0, s>>>(f.st);
    LAUNCH_CHECK();
    for (unsigned it = 0; it < iterations; it++) {
        const bool even = it % 2 == 0;
        k_fail_sweep<<<nt, kBlock, 0, s>>>(g, sea, even ? hgt : f.h2, even ? f.h2 : hgt,
                                           even ? avail : f.a2, even ? f.a2 : avail, f.parts, f.st);
        LAUNCH_CHECK();
        k_fail_step<<<1, kReduceBlock, 0, s>>>(f.parts, nt, f.st);
        LAUNCH_CHECK();
    }
    k_fail_final<<<nblocks((long)n), kBlock, 0, s>>>((int)n, f.h2, f.a2, hgt, avail, f.st);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(f.h_st, f.st, sizeof(FailState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (its) *its = (unsigned)f.h_st->its;
    return f.h_st->moved;
}

}  // namespace dtd
