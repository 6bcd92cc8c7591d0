// The learned first guess's network on the device: a 4-level U-Net of 3x3 convolutions over NHWC
// bf16 tensors (periodic in x, zero-padded in y, as the planet grid), its backward passes and its
// optimiser.  Included once, by dtdiff.hip.  The solve never depends on it: a guess it gives only
// changes where conjugate gradients start, and the same stopping test ends the solve.
//
// Tensors: n images of h x w pixels with c channels, element (img, y, x, ch) at
// ((img h + y) w + x) c + ch, bf16 (the upper half of an f32).  Channel counts are multiples of 16
// and w of 16, so that a wave's 16-pixel tiles stay in one row.  Weights of a convolution with cin
// inputs and cout outputs: wt[co][tap][ci] (tap = 3 (dy + 1) + dx + 1), bf16, with f32 masters.
//
// The convolution is an implicit GEMM on the matrix cores: pixels x (9 cin) times (9 cin) x cout,
// v_mfma_f32_16x16x32_bf16 (16x16x16 for 16-channel inputs) with f32 accumulation, one wave a
// tile of up to 64 pixels and 64 outputs.  Its backward pass for the input is the same kernel with the taps mirrored and the
// weights' channel roles swapped (the adjoint of "read x at p + d, wrapping in x, zero past the
// rows" is "read at p - d" under the same rules).  The weights' gradient is a GEMM over pixels:
// per chunk of pixels one wave a (16 outputs, 16 inputs, tap) tile on channel-major copies of the
// layer's input and output gradient, then the chunks added in order, so every sum has a fixed
// order and the training is deterministic.
#pragma once

#include "context.hpp"

namespace dtd {
namespace nn {

typedef short s4_t __attribute__((ext_vector_type(4)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// f32 -> bf16, round to nearest even (NaN kept a NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct Shape {
    int n, h, w;  // images, rows, columns
};

__device__ __forceinline__ long pixels(const Shape &s) { return (long)s.n * s.h * s.w; }

// y[p][co] = act(bias[co] + sum over taps and ci of x[p + d(tap)][ci] wt[co][tap][ci]); act is
// max(0, .) when relu.  y32 (optional) takes the first y32c outputs in f32 as well.  One wave a
// tile of MT x 16 consecutive pixels of one row and NT x 16 outputs (blockIdx.y picks the
// outputs' group): each weight fragment it loads feeds MT matrix instructions and each input
// fragment NT.  CIN > 0 fixes the input channels at compile time, so that the tap and channel
// loops unroll and their loads issue ahead of the matrix instructions; CIN = 0 reads cin.  KW = 32
// (cin a multiple of 32) takes 32 channels a step with v_mfma_f32_16x16x32_bf16 (16-byte
// fragments, lane l holding channels 8 (l / 16) .. + 7 of its row), KW = 16 takes 16 with
// v_mfma_f32_16x16x16_bf16 (8-byte fragments, channels 4 (l / 16) .. + 3).
template <int NT, int MT, int CIN, int KW>
__global__ __launch_bounds__(256) void k_conv3x3(Shape s, int cin_, int cout,
                                                 const uint16_t *__restrict__ x,
                                                 const uint16_t *__restrict__ wt,
                                                 const float *__restrict__ bias, int relu,
                                                 uint16_t *__restrict__ y, float *__restrict__ y32,
                                                 int y32c) {
    constexpr int TP = 16 * MT;  // pixels a wave
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long tile = (long)blockIdx.x * 4 + wave;
    if (tile * TP >= pixels(s)) return;
    const int co0 = blockIdx.y * 16 * NT;
    const long p0 = tile * TP;
    const int col0 = (int)(p0 % s.w);
    const long rowi = p0 / s.w;  // image row index over all images
    const int row = (int)(rowi % s.h);
    const long img = rowi / s.h;
    const int m = lane & 15, kq = (lane >> 4) * 4;
    const int cin = CIN > 0 ? CIN : cin_;
    constexpr int U = CIN > 0 && CIN <= 128 ? (CIN / KW > 0 ? CIN / KW : 1) : 4;
    f4_t acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[i][t] = f4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const int yy = row + dy;
        const bool inside = yy >= 0 && yy < s.h;
        const uint16_t *xp[MT];
#pragma unroll
        for (int i = 0; i < MT; i++) {
            int xx = col0 + 16 * i + m + dx;
            xx = xx < 0 ? xx + s.w : (xx >= s.w ? xx - s.w : xx);
            xp[i] = inside ? x + ((img * s.h + yy) * s.w + xx) * (long)cin : nullptr;
        }
        const int ko = KW == 32 ? 2 * kq : kq;  // this lane's first channel in a step
        const uint16_t *wp = wt + ((size_t)(co0 + m) * 9 + tap) * cin + ko;
#pragma unroll U
        for (int c0 = 0; c0 < cin; c0 += KW) {
            if constexpr (KW == 32) {
                b8_t a[MT], b[NT];
#pragma unroll
                for (int i = 0; i < MT; i++) {
                    a[i] = b8_t{};
                    if (inside) a[i] = *reinterpret_cast<const b8_t *>(xp[i] + c0 + ko);
                }
#pragma unroll
                for (int t = 0; t < NT; t++)
                    b[t] = *reinterpret_cast<const b8_t *>(wp + (size_t)16 * t * 9 * cin + c0);
#pragma unroll
                for (int i = 0; i < MT; i++)
#pragma unroll
                    for (int t = 0; t < NT; t++)
                        acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[t], acc[i][t],
                                                                            0, 0, 0);
            } else {
                s4_t a[MT], b[NT];
#pragma unroll
                for (int i = 0; i < MT; i++) {
                    a[i] = s4_t{0, 0, 0, 0};
                    if (inside) a[i] = *reinterpret_cast<const s4_t *>(xp[i] + c0 + ko);
                }
#pragma unroll
                for (int t = 0; t < NT; t++)
                    b[t] = *reinterpret_cast<const s4_t *>(wp + (size_t)16 * t * 9 * cin + c0);
#pragma unroll
                for (int i = 0; i < MT; i++)
#pragma unroll
                    for (int t = 0; t < NT; t++)
                        acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[t],
                                                                              acc[i][t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = co0 + 16 * t + m;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < MT; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long p = p0 + 16 * i + kq + q;
                float v = acc[i][t][q] + bv;
                if (relu) v = fmaxf(v, 0.0f);
                if (y) y[p * cout + co] = f2bf(v);
                if (y32 && co < y32c) y32[p * y32c + co] = v;
            }
    }
}

// The same convolution with its operands staged in LDS, for w a multiple of 64 MT, cin of KC,
// cout of 16 NTB: one block a segment of 64 MT pixels of a row and 16 NTB outputs.  For each row
// offset dy and chunk of KC input channels the block stages the segment's input row (with the
// wrapped halo pixel at each end) and the three taps' weights of that dy; then the three column
// offsets dx read shifted windows of the staged row.  Wave v takes pixels 16 MT v .. 16 MT
// (v + 1) - 1 of the segment and all the block's outputs.  Rows past the poles contribute nothing
// and are skipped.  KC = 32 steps with v_mfma_f32_16x16x32_bf16, KC = 16 (16-channel inputs) with
// v_mfma_f32_16x16x16_bf16.
template <int NTB, int MT, int KC = 32>
__global__ __launch_bounds__(256) void k_conv3x3_lds(Shape s, int cin, int cout,
                                                     const uint16_t *__restrict__ x,
                                                     const uint16_t *__restrict__ xu, int cu,
                                                     const uint16_t *__restrict__ wt,
                                                     const float *__restrict__ bias, int relu,
                                                     uint16_t *__restrict__ y,
                                                     float *__restrict__ y32, int y32c) {
    constexpr int BP = 64 * MT, AS = KC + 8, BCO = 16 * NTB, CPP = KC / 8;  // 16-byte chunks a row
    __shared__ __align__(16) uint16_t la[(BP + 2) * AS];
    __shared__ __align__(16) uint16_t lb[3 * BCO * AS];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const long p0 = (long)blockIdx.x * BP;
    const int co0 = blockIdx.y * BCO;
    const int col0 = (int)(p0 % s.w);
    const long rowi = p0 / s.w;
    const int row = (int)(rowi % s.h);
    const long img = rowi / s.h;
    const int m = lane & 15, k8 = (lane >> 4) * (KC / 4);
    f4_t acc[MT][NTB];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int n = 0; n < NTB; n++) acc[i][n] = f4_t{0.0f, 0.0f, 0.0f, 0.0f};
    // The stages: the row offsets inside the grid (the same for the whole block) times the
    // 32-channel chunks.  Each thread holds its share of the next stage's operands in registers
    // while the block computes the current one.
    const int nch = cin / KC;
    const int dy0 = row > 0 ? -1 : 0, dy1 = row + 1 < s.h ? 1 : 0;  // the rows inside the grid
    const int nst = (dy1 - dy0 + 1) * nch;
    constexpr int NA = (BP + 2) * CPP, NB = 3 * BCO * CPP;  // 16-byte chunks of a stage
    constexpr int UA = (NA + 255) / 256, UB = (NB + 255) / 256;
    u4_t ra[UA], rb[UB];
    // With xu, input channels [0, cu) are the nearest upsampling of xu (h/2 x w/2, cu channels)
    // and [cu, cin) are x (cin - cu channels): the concatenation read in place.
    auto fetch = [&](int k) {
        const int dy = dy0 + k / nch, c0 = (k % nch) * KC;
        const bool up = xu && c0 < cu;
        const int cx = xu ? cin - cu : cin, cc = xu && !up ? c0 - cu : c0;
        const uint16_t *xrow = up ? xu + (img * (s.h / 2) + (row + dy) / 2) * (long)(s.w / 2) * cu
                                  : x + (img * s.h + row + dy) * (long)s.w * cx;
#pragma unroll
        for (int u = 0; u < UA; u++) {
            const int e = tid + 256 * u;
            if (e < NA) {
                const int pp = e / CPP, ch = (e % CPP) * 8;
                int xx = col0 - 1 + pp;
                xx = xx < 0 ? xx + s.w : (xx >= s.w ? xx - s.w : xx);
                ra[u] = up ? *reinterpret_cast<const u4_t *>(xrow + (long)(xx / 2) * cu + c0 + ch)
                           : *reinterpret_cast<const u4_t *>(xrow + (long)xx * cx + cc + ch);
            }
        }
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int e = tid + 256 * u;
            if (e < NB) {
                const int t = e / (BCO * CPP), r = e % (BCO * CPP), co = r / CPP,
                          ch = (r % CPP) * 8;
                const int tap = (dy + 1) * 3 + t;
                rb[u] = *reinterpret_cast<const u4_t *>(wt + ((size_t)(co0 + co) * 9 + tap) * cin +
                                                        c0 + ch);
            }
        }
    };
    if (nst > 0) fetch(0);
    for (int k = 0; k < nst; k++) {
        __syncthreads();  // the last stage's reads are done
#pragma unroll
        for (int u = 0; u < UA; u++) {
            const int e = tid + 256 * u;
            if (e < NA) *reinterpret_cast<u4_t *>(&la[(e / CPP) * AS + (e % CPP) * 8]) = ra[u];
        }
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int e = tid + 256 * u;
            if (e < NB) {
                const int t = e / (BCO * CPP), r = e % (BCO * CPP);
                *reinterpret_cast<u4_t *>(&lb[(t * BCO + r / CPP) * AS + (r % CPP) * 8]) = rb[u];
            }
        }
        __syncthreads();
        if (k + 1 < nst) fetch(k + 1);
#pragma unroll
        for (int t = 0; t < 3; t++) {  // dx = t - 1: pixel q reads staged pixel q + t
            if constexpr (KC == 32) {
                b8_t a[MT], b[NTB];
#pragma unroll
                for (int i = 0; i < MT; i++)
                    a[i] = *reinterpret_cast<const b8_t *>(
                        &la[(wave * 16 * MT + 16 * i + m + t) * AS + k8]);
#pragma unroll
                for (int n = 0; n < NTB; n++)
                    b[n] = *reinterpret_cast<const b8_t *>(&lb[(t * BCO + 16 * n + m) * AS + k8]);
#pragma unroll
                for (int i = 0; i < MT; i++)
#pragma unroll
                    for (int n = 0; n < NTB; n++)
                        acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[n], acc[i][n],
                                                                            0, 0, 0);
            } else {
                s4_t a[MT], b[NTB];
#pragma unroll
                for (int i = 0; i < MT; i++)
                    a[i] = *reinterpret_cast<const s4_t *>(
                        &la[(wave * 16 * MT + 16 * i + m + t) * AS + k8]);
#pragma unroll
                for (int n = 0; n < NTB; n++)
                    b[n] = *reinterpret_cast<const s4_t *>(&lb[(t * BCO + 16 * n + m) * AS + k8]);
#pragma unroll
                for (int i = 0; i < MT; i++)
#pragma unroll
                    for (int n = 0; n < NTB; n++)
                        acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[n],
                                                                              acc[i][n], 0, 0, 0);
            }
        }
    }
    const int kq = (lane >> 4) * 4;
#pragma unroll
    for (int n = 0; n < NTB; n++) {
        const int co = co0 + 16 * n + m;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < MT; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long p = p0 + wave * 16 * MT + 16 * i + kq + q;
                float v = acc[i][n][q] + bv;
                if (relu) v = fmaxf(v, 0.0f);
                if (y) y[p * cout + co] = f2bf(v);
                if (y32 && co < y32c) y32[p * y32c + co] = v;
            }
    }
}

// wt_t[ci][tap][co] = wt[co][8 - tap][ci]: the weights of the input-gradient convolution.
__global__ __launch_bounds__(256) void k_conv_flip(int cin, int cout, const uint16_t *__restrict__ wt,
                                                   uint16_t *__restrict__ wt_t) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)cin * 9 * cout) return;
    const int co = (int)(i % cout), tap = (int)((i / cout) % 9), ci = (int)(i / (9L * cout));
    wt_t[i] = wt[((size_t)co * 9 + (8 - tap)) * cin + ci];
}

// Eight bf16 values in 16 bytes (element 2k in the low half of word k), for the elementwise
// kernels: one thread a group of eight channels.

__device__ __forceinline__ void unpack8(u4_t v, float *f) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        f[2 * k] = __uint_as_float(v[k] << 16);
        f[2 * k + 1] = __uint_as_float(v[k] & 0xFFFF0000u);
    }
}

__device__ __forceinline__ u4_t pack8(const float *f) {
    u4_t v;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
    return v;
}

__device__ __forceinline__ u4_t ld8(const uint16_t *p) { return *reinterpret_cast<const u4_t *>(p); }
__device__ __forceinline__ void st8(uint16_t *p, u4_t v) { *reinterpret_cast<u4_t *>(p) = v; }

// The parent (h/2 x w/2 grid) of full-size pixel p.
__device__ __forceinline__ long parent_of(const Shape &s, long p) {
    const int col = (int)(p % s.w);
    const long r = p / s.w;
    const int row = (int)(r % s.h);
    const long img = r / s.h;
    return (img * (s.h / 2) + row / 2) * (s.w / 2) + col / 2;
}

// The first of the four children (full-size pixel) of half-size pixel pc.
__device__ __forceinline__ long child_of(const Shape &s, long pc) {
    const int hh = s.h / 2, ww = s.w / 2;
    const int xc = (int)(pc % ww);
    const long r = pc / ww;
    const int yc = (int)(r % hh);
    const long img = r / hh;
    return (img * s.h + 2 * yc) * s.w + 2 * xc;
}

// dz = dy where y > 0, else 0 (the ReLU's backward pass on its output); eight a thread.
__global__ __launch_bounds__(256) void k_relu_bwd(long n, const uint16_t *__restrict__ y,
                                                  const uint16_t *__restrict__ dy,
                                                  uint16_t *__restrict__ dz) {
    const long i = ((long)blockIdx.x * 256 + threadIdx.x) * 8;
    if (i >= n) return;
    if (i + 8 <= n) {
        const u4_t yv = ld8(y + i), dv = ld8(dy + i);
        u4_t o;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t lo = __uint_as_float(yv[k] << 16) > 0.0f ? 0x0000FFFFu : 0u;
            const uint32_t hi = __uint_as_float(yv[k] & 0xFFFF0000u) > 0.0f ? 0xFFFF0000u : 0u;
            o[k] = dv[k] & (lo | hi);
        }
        st8(dz + i, o);
    } else {
        for (long j = i; j < n; j++) dz[j] = bf2f(y[j]) > 0.0f ? dy[j] : (uint16_t)0;
    }
}

// xt[ch][p] = x[p'][ch] for all pixels p, where p' is p moved dx columns along its row (wrapping):
// channel-major copies, the column shift applied once here so that the weights' gradient reads
// every shifted operand as one aligned run.  Tiles of 64 pixels x 16 channels.
__global__ __launch_bounds__(256) void k_to_planes(Shape s, int c, int dx,
                                                   const uint16_t *__restrict__ x,
                                                   uint16_t *__restrict__ xt) {
    __shared__ uint16_t t[16][65];
    const long np = pixels(s);
    const long p0 = (long)blockIdx.x * 64;
    const int c0 = blockIdx.y * 16;
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
        const int pp = e / 16, cc = e % 16;
        const long p = p0 + pp;
        uint16_t v = 0;
        if (p < np) {
            const int col = (int)(p % s.w);
            int xx = col + dx;
            xx = xx < 0 ? xx + s.w : (xx >= s.w ? xx - s.w : xx);
            v = x[(p - col + xx) * c + c0 + cc];
        }
        t[cc][pp] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
        const int cc = e / 64, pp = e % 64;
        if (p0 + pp < np) xt[(long)(c0 + cc) * np + p0 + pp] = t[cc][pp];
    }
}

// Per chunk of `chunk` pixels, part[k][co][tap][ci] = sum over the chunk's pixels p of
// dz[p][co] x[p + d(tap)][ci], from channel-major copies dzt[co][p] and xs[dx + 1][ci][p] (x with
// its columns moved by dx).  One wave a (16 outputs, 16 inputs, tap) tile; blockIdx.x the chunk.
// KW = 32: 32 pixels a step (v_mfma_f32_16x16x32_bf16, each lane 8 consecutive pixels of one row,
// np a multiple of 32); KW = 16: 16 pixels a step (16x16x16, 4 pixels a lane).
template <int KW>
__global__ __launch_bounds__(256) void k_conv_dw(Shape s, int cin, int cout,
                                                 const uint16_t *__restrict__ xs,
                                                 const uint16_t *__restrict__ dzt, int chunk,
                                                 float *__restrict__ part) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int combo = blockIdx.y * 4 + wave;
    const int ncit = cin / 16, ncot = cout / 16;
    if (combo >= ncot * ncit * 9) return;
    const int tap = combo % 9, cit = (combo / 9) % ncit, cot = combo / (9 * ncit);
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const long np = pixels(s);
    const long pa = (long)blockIdx.x * chunk, pb = min(np, pa + chunk);
    const int m = lane & 15, kq = (lane >> 4) * 4, ko = (lane >> 4) * (KW / 4);
    const uint16_t *dzp = dzt + (long)(cot * 16 + m) * np;
    const uint16_t *xp = xs + ((long)(dx + 1) * cin + cit * 16 + m) * np;
    const long shift = (long)dy * s.w;  // the row above or below, read only inside the grid
    f4_t acc = f4_t{0.0f, 0.0f, 0.0f, 0.0f};
    // Unrolled so that several steps' loads are in flight before their matrix instructions.
#pragma unroll 4
    for (long p = pa; p < pb; p += KW) {
        const long q0 = p + ko;  // this lane's pixels q0 .. q0 + KW / 4 - 1, in one row
        const int row = (int)((q0 / s.w) % s.h), yy = row + dy;
        const bool in = yy >= 0 && yy < s.h;
        if constexpr (KW == 32) {
            const b8_t a = *reinterpret_cast<const b8_t *>(dzp + q0);
            b8_t b = b8_t{};
            if (in) b = *reinterpret_cast<const b8_t *>(xp + (q0 + shift));
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        } else {
            const s4_t a = *reinterpret_cast<const s4_t *>(dzp + q0);
            s4_t b = s4_t{0, 0, 0, 0};
            if (in) b = *reinterpret_cast<const s4_t *>(xp + (q0 + shift));
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    // D[co][ci]: lane holds rows kq + q (outputs), column m (inputs).
    float *out = part + (size_t)blockIdx.x * cout * 9 * cin;
#pragma unroll
    for (int q = 0; q < 4; q++)
        out[((size_t)(cot * 16 + kq + q) * 9 + tap) * cin + cit * 16 + m] = acc[q];
}

// dw[i] = the chunks' parts added in chunk order; db[co] = sum over pixels of dzt[co][.]
// (a fixed tree over the block's threads).
__global__ __launch_bounds__(256) void k_dw_reduce(long nw, int chunks,
                                                   const float *__restrict__ part,
                                                   float *__restrict__ dw) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= nw) return;
    float s = 0.0f;
    for (int k = 0; k < chunks; k++) s += part[(size_t)k * nw + i];
    dw[i] = s;
}

__global__ __launch_bounds__(256) void k_db(long np, const uint16_t *__restrict__ dzt,
                                            float *__restrict__ db) {
    __shared__ float t[256];
    const uint16_t *r = dzt + (long)blockIdx.x * np;
    float s = 0.0f;
    for (long p = threadIdx.x; p < np; p += 256) s += bf2f(r[p]);
    t[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) t[threadIdx.x] += t[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[blockIdx.x] = t[0];
}

// 2 x 2 mean: y (h/2 x w/2) from x (h x w), c channels; eight a thread.
__global__ __launch_bounds__(256) void k_pool(Shape s, int c, const uint16_t *__restrict__ x,
                                              uint16_t *__restrict__ y) {
    const int g = c / 8;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)s.n * (s.h / 2) * (s.w / 2) * g) return;
    const int ch = (int)(i % g) * 8;
    const long pc = i / g;
    const long base = child_of(s, pc) * c + ch, down = (long)s.w * c;
    float a[8], b[8], d[8], e[8], o[8];
    unpack8(ld8(x + base), a), unpack8(ld8(x + base + c), b);
    unpack8(ld8(x + base + down), d), unpack8(ld8(x + base + down + c), e);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = 0.25f * ((a[k] + b[k]) + (d[k] + e[k]));
    st8(y + pc * c + ch, pack8(o));
}

// cat[p][0 .. cu) = u[parent(p)] (nearest, from the half-size grid), cat[p][cu .. cu + cs) =
// skip[p]; s is the full-size shape; eight channels a thread.
__global__ __launch_bounds__(256) void k_up_cat(Shape s, int cu, int cs, const uint16_t *__restrict__ u,
                                                const uint16_t *__restrict__ skip,
                                                uint16_t *__restrict__ cat) {
    const int c = cu + cs, g = c / 8;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= pixels(s) * g) return;
    const int ch = (int)(i % g) * 8;
    const long p = i / g;
    st8(cat + p * c + ch, ch >= cu ? ld8(skip + p * cs + (ch - cu))
                                   : ld8(u + parent_of(s, p) * cu + ch));
}

// The concat's backward pass: du[q][ch] = sum of dcat[p][ch] over q's four children (ch < cu),
// and dskip[p][ch] = dcat[p][cu + ch] + 0.25 dpool[parent(p)][ch] (the skip's tensor also fed the
// next level's pooling; dpool null: none).  s is the full-size shape; eight channels a thread.
__global__ __launch_bounds__(256) void k_up_cat_bwd_u(Shape s, int cu, int cs,
                                                      const uint16_t *__restrict__ dcat,
                                                      uint16_t *__restrict__ du) {
    const int c = cu + cs, g = cu / 8;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)s.n * (s.h / 2) * (s.w / 2) * g) return;
    const int ch = (int)(i % g) * 8;
    const long pc = i / g;
    const long base = child_of(s, pc) * c + ch, down = (long)s.w * c;
    float a[8], b[8], d[8], e[8], o[8];
    unpack8(ld8(dcat + base), a), unpack8(ld8(dcat + base + c), b);
    unpack8(ld8(dcat + base + down), d), unpack8(ld8(dcat + base + down + c), e);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = (a[k] + b[k]) + (d[k] + e[k]);
    st8(du + pc * cu + ch, pack8(o));
}

__global__ __launch_bounds__(256) void k_skip_bwd(Shape s, int cu, int cs,
                                                  const uint16_t *__restrict__ dcat,
                                                  const uint16_t *__restrict__ dpool,
                                                  uint16_t *__restrict__ dskip) {
    const int c = cu + cs, g = cs / 8;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= pixels(s) * g) return;
    const int ch = (int)(i % g) * 8;
    const long p = i / g;
    float v[8];
    unpack8(ld8(dcat + p * c + cu + ch), v);
    if (dpool) {
        float q[8];
        unpack8(ld8(dpool + parent_of(s, p) * cs + ch), q);
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] += 0.25f * q[k];
    }
    st8(dskip + p * cs + ch, pack8(v));
}

// The network's input (16 channels, four used): h / scale, ln kd, ln(dt kd / dx^2) / 4 and the
// base guess / scale (0 without one).
__global__ __launch_bounds__(256) void k_nn_input(long np, const double *__restrict__ h,
                                                  const double *__restrict__ kd,
                                                  const double *__restrict__ base, double dt,
                                                  double dx2, double scale,
                                                  uint16_t *__restrict__ x) {
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const double k = fmax(kd[p], 1e-30);
    const unsigned c0 = f2bf((float)(h[p] / scale));
    const unsigned c1 = f2bf((float)log(k));
    const unsigned c2 = f2bf((float)(0.25 * log(dt * k / dx2)));
    const unsigned c3 = base ? f2bf((float)(base[p] / scale)) : 0u;
    // The pixel's 16 channels (32 bytes, 16-byte aligned) in two stores.
    u4_t *o = reinterpret_cast<u4_t *>(x + p * 16);
    o[0] = u4_t{c0 | (c1 << 16), c2 | (c3 << 16), 0u, 0u};
    o[1] = u4_t{0u, 0u, 0u, 0u};
}

// The guess: g = base + scale out[p][0] (the network predicts the change from the base guess).
__global__ __launch_bounds__(256) void k_nn_guess(long np, const double *__restrict__ base,
                                                  const float *__restrict__ out, int outc,
                                                  double scale, double *__restrict__ g) {
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p < np) g[p] = base[p] + scale * (double)out[p * outc];
}

// ---- the learned coarse correction: the network at 1/f of the grid's resolution -------------

// Block means: out[J (w/f) + I] = the mean over the f x f fine cells of block (I, J) of in, the
// cells added in row-major order from 0 and the sum divided by f^2; with log_mean, exp of the mean
// of log(max(in, 1e-30)) (for kd: the block's geometric mean).  One thread a coarse cell.
__global__ __launch_bounds__(256) void k_block_mean(int w, int h, int f, int log_mean,
                                                    const double *__restrict__ in,
                                                    double *__restrict__ out) {
    const int wc = w / f, hc = h / f;
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k >= (long)wc * hc) return;
    const int J = (int)(k / wc), I = (int)(k - (long)J * wc);
    double acc = 0.0;
    for (int j = J * f; j < J * f + f; j++) {
        const double *row = in + (size_t)j * w + (size_t)I * f;
        for (int i = 0; i < f; i++) acc += log_mean ? log(fmax(row[i], 1e-30)) : row[i];
    }
    const double m = acc / (double)(f * f);
    out[k] = log_mean ? exp(m) : m;
}

// g = base + scale (P out[:, 0]): P interpolates bilinearly between the coarse cells' centres
// (coarse cell I centred on fine position f I + (f - 1) / 2), periodic in x, constant past the
// first and last rows' centres.
__global__ __launch_bounds__(256) void k_nn_guess_coarse(int w, int h, int f,
                                                         const double *__restrict__ base,
                                                         const float *__restrict__ out, int outc,
                                                         double scale, double *__restrict__ g) {
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= (long)w * h) return;
    const int wc = w / f, hc = h / f;
    const int j = (int)(p / w), i = (int)(p - (long)j * w);
    const float X = ((float)i + 0.5f) / (float)f - 0.5f, Y = ((float)j + 0.5f) / (float)f - 0.5f;
    const float fx = floorf(X), fy = floorf(Y);
    const float tx = X - fx, ty = Y - fy;
    const int i0 = (((int)fx % wc) + wc) % wc, i1 = (i0 + 1) % wc;
    const int j0 = min(max((int)fy, 0), hc - 1), j1 = min(max((int)fy + 1, 0), hc - 1);
    auto o = [&](int jj, int ii) { return out[((long)jj * wc + ii) * outc]; };
    const float a = (1.0f - tx) * o(j0, i0) + tx * o(j0, i1);
    const float b = (1.0f - tx) * o(j1, i0) + tx * o(j1, i1);
    g[p] = base[p] + scale * (double)((1.0f - ty) * a + ty * b);
}

// Loss against the target t: the network's out[p][0] against (t - base) / scale (formed in f64);
// parts[block] = sum of the squared errors over the block (a fixed tree), dy[p][0] =
// 2 (out - target) / np, the other 15 channels 0.
__global__ __launch_bounds__(256) void k_nn_loss(long np, const float *__restrict__ out, int outc,
                                                 const double *__restrict__ t,
                                                 const double *__restrict__ base, double scale,
                                                 float *__restrict__ parts,
                                                 uint16_t *__restrict__ dy) {
    __shared__ float red[256];
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    float e = 0.0f;
    if (p < np) {
        const float d = out[p * outc] - (float)((t[p] - base[p]) / scale);
        e = d * d;
        uint16_t *o = dy + p * 16;
        o[0] = f2bf(2.0f * d / (float)np);
#pragma unroll
        for (int c = 1; c < 16; c++) o[c] = 0;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on f32 masters w with moments m, v; wb gets bf16(w).  bc1, bc2 are the
// bias corrections 1 - beta^t.
__global__ __launch_bounds__(256) void k_adam(long n, float *__restrict__ w, float *__restrict__ m,
                                              float *__restrict__ v, const float *__restrict__ g,
                                              float lr, float b1, float b2, float eps, float bc1,
                                              float bc2, uint16_t *__restrict__ wb) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi, v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

__global__ __launch_bounds__(256) void k_to_bf16(long n, const float *__restrict__ a,
                                                 uint16_t *__restrict__ b) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

// ---- host side ---------------------------------------------------------------------------

constexpr int kChunk = 4096;  // pixels a chunk of the weights' gradient

static void check_shape(int n, int h, int w) {
    if (n < 1 || h < 1 || w < 16 || w % 16)
        throw Error{DTD_EINVAL, "a network tensor needs n, h >= 1 and w a multiple of 16"};
    if ((long)n * h * w > (long)INT32_MAX * 16) throw Error{DTD_ETOOBIG, "too many pixels"};
}

static void check_channels(int c) {
    if (c < 16 || c % 16) throw Error{DTD_EINVAL, "channel counts must be multiples of 16"};
}

static unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

// The kernels read bf16 tensors eight or four values at a time.
static void check_aligned(const void *p, size_t a) {
    if (p && (reinterpret_cast<uintptr_t>(p) % a))
        throw Error{DTD_EINVAL, a == 16 ? "network tensors must be 16-byte aligned"
                                        : "network tensors must be 8-byte aligned"};
}

template <int NT, int CIN>
static void conv_launch_c(hipStream_t s, Shape sh, int cin, int cout, const uint16_t *x,
                          const uint16_t *wt, const float *bias, int relu, uint16_t *y,
                          float *y32, int y32c) {
    const long np = (long)sh.n * sh.h * sh.w;
    const bool k32 = cin % 32 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(wt) % 16 == 0;
    if (k32 && sh.w % 128 == 0) {           // operands staged in LDS
        const bool wide = sh.w % 256 == 0;  // 256-pixel segments, four tiles a wave
        const unsigned bx = (unsigned)(np / (wide ? 256 : 128));
        if (cout % 32 != 0 && wide)
            k_conv3x3_lds<1, 4><<<dim3(bx, (unsigned)(cout / 16)), 256, 0, s>>>(
                sh, cin, cout, x, nullptr, 0, wt, bias, relu, y, y32, y32c);
        else if (cout % 32 != 0)
            k_conv3x3_lds<1, 2><<<dim3(bx, (unsigned)(cout / 16)), 256, 0, s>>>(
                sh, cin, cout, x, nullptr, 0, wt, bias, relu, y, y32, y32c);
        else if (cout % 64 == 0 && wide)
            k_conv3x3_lds<4, 4><<<dim3(bx, (unsigned)(cout / 64)), 256, 0, s>>>(
                sh, cin, cout, x, nullptr, 0, wt, bias, relu, y, y32, y32c);
        else if (cout % 64 == 0)
            k_conv3x3_lds<4, 2><<<dim3(bx, (unsigned)(cout / 64)), 256, 0, s>>>(
                sh, cin, cout, x, nullptr, 0, wt, bias, relu, y, y32, y32c);
        else if (wide)
            k_conv3x3_lds<2, 4><<<dim3(bx, (unsigned)(cout / 32)), 256, 0, s>>>(
                sh, cin, cout, x, nullptr, 0, wt, bias, relu, y, y32, y32c);
        else
            k_conv3x3_lds<2, 2><<<dim3(bx, (unsigned)(cout / 32)), 256, 0, s>>>(
                sh, cin, cout, x, nullptr, 0, wt, bias, relu, y, y32, y32c);
        LAUNCH_CHECK();
        return;
    }
    const bool k16 = cin % 16 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(wt) % 16 == 0;
    if (k16 && sh.w % 128 == 0) {  // 16-channel stages in LDS
        const bool wide = sh.w % 256 == 0;
        const dim3 grid((unsigned)(np / (wide ? 256 : 128)), 1);
        auto go = [&](auto kern, int ntb) {
            kern<<<dim3(grid.x, (unsigned)(cout / (16 * ntb))), 256, 0, s>>>(
                sh, cin, cout, x, nullptr, 0, wt, bias, relu, y, y32, y32c);
        };
        if (cout % 32 != 0)
            wide ? go(k_conv3x3_lds<1, 4, 16>, 1) : go(k_conv3x3_lds<1, 2, 16>, 1);
        else if (cout % 64 == 0)
            wide ? go(k_conv3x3_lds<4, 4, 16>, 4) : go(k_conv3x3_lds<4, 2, 16>, 4);
        else
            wide ? go(k_conv3x3_lds<2, 4, 16>, 2) : go(k_conv3x3_lds<2, 2, 16>, 2);
        LAUNCH_CHECK();
        return;
    }
    const bool m4 = sh.w % 64 == 0;  // four 16-pixel tiles a wave, in one row
    dim3 grid((unsigned)((np / (m4 ? 64 : 16) + 3) / 4), (unsigned)(cout / (16 * NT)));
    if (m4 && k32)
        k_conv3x3<NT, 4, CIN, 32><<<grid, 256, 0, s>>>(sh, cin, cout, x, wt, bias, relu, y, y32,
                                                       y32c);
    else if (m4)
        k_conv3x3<NT, 4, CIN, 16><<<grid, 256, 0, s>>>(sh, cin, cout, x, wt, bias, relu, y, y32,
                                                       y32c);
    else if (k32)
        k_conv3x3<NT, 1, CIN, 32><<<grid, 256, 0, s>>>(sh, cin, cout, x, wt, bias, relu, y, y32,
                                                       y32c);
    else
        k_conv3x3<NT, 1, CIN, 16><<<grid, 256, 0, s>>>(sh, cin, cout, x, wt, bias, relu, y, y32,
                                                       y32c);
    LAUNCH_CHECK();
}

// The convolution of [nearest upsampling of u (h/2 x w/2, cu channels), skip (cs channels)] without
// forming the concatenation: the LDS-staged kernel reads each channel chunk from its source.
static bool conv_cat_ok(Shape sh, int cu, int cs, const void *u, const void *skip, const void *wt) {
    return sh.w % 128 == 0 && sh.h % 2 == 0 && cu % 32 == 0 && cs % 32 == 0 &&
           reinterpret_cast<uintptr_t>(u) % 16 == 0 && reinterpret_cast<uintptr_t>(skip) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(wt) % 16 == 0;
}

static void conv_cat(hipStream_t s, Shape sh, int cu, int cs, int cout, const uint16_t *u,
                     const uint16_t *skip, const uint16_t *wt, const float *bias, int relu,
                     uint16_t *y, float *y32, int y32c) {
    const long np = (long)sh.n * sh.h * sh.w;
    const int cin = cu + cs;
    const bool wide = sh.w % 256 == 0;
    const unsigned bx = (unsigned)(np / (wide ? 256 : 128));
    if (cout % 32 != 0 && wide)
        k_conv3x3_lds<1, 4><<<dim3(bx, (unsigned)(cout / 16)), 256, 0, s>>>(
            sh, cin, cout, skip, u, cu, wt, bias, relu, y, y32, y32c);
    else if (cout % 32 != 0)
        k_conv3x3_lds<1, 2><<<dim3(bx, (unsigned)(cout / 16)), 256, 0, s>>>(
            sh, cin, cout, skip, u, cu, wt, bias, relu, y, y32, y32c);
    else if (cout % 64 == 0 && wide)
        k_conv3x3_lds<4, 4><<<dim3(bx, (unsigned)(cout / 64)), 256, 0, s>>>(
            sh, cin, cout, skip, u, cu, wt, bias, relu, y, y32, y32c);
    else if (cout % 64 == 0)
        k_conv3x3_lds<4, 2><<<dim3(bx, (unsigned)(cout / 64)), 256, 0, s>>>(
            sh, cin, cout, skip, u, cu, wt, bias, relu, y, y32, y32c);
    else if (wide)
        k_conv3x3_lds<2, 4><<<dim3(bx, (unsigned)(cout / 32)), 256, 0, s>>>(
            sh, cin, cout, skip, u, cu, wt, bias, relu, y, y32, y32c);
    else
        k_conv3x3_lds<2, 2><<<dim3(bx, (unsigned)(cout / 32)), 256, 0, s>>>(
            sh, cin, cout, skip, u, cu, wt, bias, relu, y, y32, y32c);
    LAUNCH_CHECK();
}

// The network's input channel counts get their own unrolled kernels; others take the general one.
template <int NT>
static void conv_launch(hipStream_t s, Shape sh, int cin, int cout, const uint16_t *x,
                        const uint16_t *wt, const float *bias, int relu, uint16_t *y, float *y32,
                        int y32c) {
    switch (cin) {
    case 16: return conv_launch_c<NT, 16>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    case 32: return conv_launch_c<NT, 32>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    case 64: return conv_launch_c<NT, 64>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    case 96: return conv_launch_c<NT, 96>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    case 128: return conv_launch_c<NT, 128>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    case 192: return conv_launch_c<NT, 192>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    case 256: return conv_launch_c<NT, 256>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    case 384: return conv_launch_c<NT, 384>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    default: return conv_launch_c<NT, 0>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    }
}

static void conv(hipStream_t s, Shape sh, int cin, int cout, const uint16_t *x,
                    const uint16_t *wt, const float *bias, int relu, uint16_t *y, float *y32,
                    int y32c) {
    if (cout % 64 == 0)
        conv_launch<4>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    else if (cout % 32 == 0)
        conv_launch<2>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
    else
        conv_launch<1>(s, sh, cin, cout, x, wt, bias, relu, y, y32, y32c);
}

}  // namespace nn
}  // namespace dtd
