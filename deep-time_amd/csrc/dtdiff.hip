// dtdiff: the C ABI of include/dtdiff.h over the HIP kernels of kernels.hpp (gfx950).
//
// Reference files (relative to the deep-time repository root):
//   D     crates/deeptime-core/src/surface/diffusion.rs
//   MG    crates/deeptime-core/src/surface/diffusion/multigrid.rs
//   MGGPU crates/deeptime-core/src/surface/diffusion/multigrid/gpu.rs
//
// Design: the whole solve stays on the device.  The operator, the cycle's levels, the first
// guess, the f64 conjugate-gradient vectors and their scalars (CgState) are resident in HBM, so
// an iteration needs no host transfer (the reference's device cycle moved r and z across the bus
// every iteration, MGGPU:669-700).  One iteration (direction pass, p.Ap, update pass, r.r, the
// V-cycle, r.z) is captured once per grid shape as a hipGraph, in two copies whose direction
// buffers alternate; the host launches them in turn and reads the device's stop flag one
// iteration behind, so the loop runs ahead of the flag by at most one (skipped) iteration.
#include "../../include/dtdiff.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "context.hpp"
#include "kernels.hpp"

namespace dtd {

void free_ws(Workspace &ws) {
    for (int k = 0; k < 2; k++) {
        if (ws.graph[k]) (void)hipGraphExecDestroy(ws.graph[k]);
        if (ws.jgraph[k]) (void)hipGraphExecDestroy(ws.jgraph[k]);
    }
    if (ws.jdiag) (void)hipFree(ws.jdiag);
    if (ws.jrows) (void)hipFree(ws.jrows);
    if (ws.jst) (void)hipFree(ws.jst);
    if (ws.jzero) (void)hipFree(ws.jzero);
    if (ws.jlive) (void)hipFree(ws.jlive);
    if (ws.jflags) (void)hipFree(ws.jflags);
    if (ws.jlist) (void)hipFree(ws.jlist);
    for (void *p : ws.owned) (void)hipFree(p);
    ws = Workspace();
}

// Lays out the workspace of a w x h grid for the cycle's precision: the f64 vectors, the partial
// sums and the state, and each level of the cycle (filled by build_operator).  take(bytes) hands
// out device memory (hipMalloc in fit; a carve of the fixed scratch array in the
// allocation-free entry point).  lean: without b, kd and the f64 operator record, which only the
// host-array entry points, the Jacobi solve and the single-piece entry points use.
template <class Take>
static void lay_out(Workspace &ws, const dtd_opts &o, int w, int h, int wrap, bool lean,
                    Take take) {
    const int prec = o.cycle_precision;
    const long n = (long)w * h;
    ws.w = w;
    ws.h = h;
    ws.wrap = wrap;
    ws.prec = prec;
    ws.n = n;
    auto f64 = [&](size_t cnt) { return (double *)take((cnt ? cnt : 1) * sizeof(double)); };
    if (!lean) ws.b = f64(n);
    ws.x = f64(n);
    ws.r = f64(n);
    if (prec != DTD_CYCLE_F64) ws.r2 = f64(n);
    ws.p0 = f64(n);
    ws.p1 = f64(n);
    ws.ap = f64(n);
    if (!lean) ws.kd = f64(n);
    ws.nparts = parts_cap(w, h);
    ws.parts = f64(2 * (size_t)ws.nparts + 2);
    ws.pair = f64(2);
    ws.st = (CgState *)take(sizeof(CgState));
    if (!lean) ws.es = (double2 *)take((n ? n : 1) * sizeof(double2));
    if (!lean && o.sums == DTD_SUMS_STRIPS)
        ws.sums = f64(2 * (size_t)((h + kStripRows - 1) / kStripRows) + 2 * (size_t)h);
    auto sh = shapes(w, h, wrap);
    auto own = [&](size_t cnt, size_t elt) { return take((cnt ? cnt : 1) * elt); };
    for (size_t l = 0; l < sh.size(); l++) {
        const long nl = (long)sh[l].w * sh[l].h;
        if (prec == DTD_CYCLE_F64) {
            Lvl<double> L{};
            L.w = sh[l].w, L.h = sh[l].h, L.wrap = sh[l].wrap, L.n = (int)nl;
            L.omega = (double)o.omega;
            L.coef = (double *)own(kRec * nl, 8);
            L.mass = l == 0 ? nullptr : (double *)own(nl, 8);
            L.r = l == 0 ? ws.r : (double *)own(nl, 8);  // the f64 cycle reads the CG residual
            L.z[0] = (double *)own(nl, 8);
            L.z[1] = (double *)own(nl, 8);
            L.bil = o.transfer == DTD_TRANSFER_BILINEAR && l + 1 < sh.size();
            if (o.coarsest == DTD_COARSEST_EXACT && l + 1 == sh.size())
                L.chol = (double *)own((size_t)nl * nl, 8);
            ws.ld.push_back(L);
        } else {
            Lvl<float> L{};
            L.w = sh[l].w, L.h = sh[l].h, L.wrap = sh[l].wrap, L.n = (int)nl;
            L.omega = o.omega;
            L.coef = (float *)own(kRec * nl, 4);
            L.mass = l == 0 ? nullptr : (float *)own(nl, 4);
            L.r = (float *)own(nl, 4);
            L.z[0] = (float *)own(nl, 4);
            L.z[1] = (float *)own(nl, 4);
            // The bf16 variant: the two finest levels (not a coarsest) keep bf16 records too.
            if (prec == DTD_CYCLE_BF16 && l < 2 && l + 1 < sh.size())
                L.coefh = (unsigned short *)own(4 * nl, 2);
            L.bil = o.transfer == DTD_TRANSFER_BILINEAR && l + 1 < sh.size();
            if (o.coarsest == DTD_COARSEST_EXACT && l + 1 == sh.size())
                L.chol = (double *)own((size_t)nl * nl, 8);
            ws.lf.push_back(L);
        }
    }
    if (!ws.lf.empty()) ws.d_lf = (Lvl<float> *)own(ws.lf.size() * sizeof(Lvl<float>), 1);
    if (!ws.ld.empty()) ws.d_ld = (Lvl<double> *)own(ws.ld.size() * sizeof(Lvl<double>), 1);
}

// The level table the one-block bottom of the cycle reads, copied to the device on stream s (the
// host copy lives in ws, as long as the device one).
static void upload_levels(Workspace &ws, hipStream_t s) {
    if (ws.d_lf)
        HIPCHK(hipMemcpyAsync(ws.d_lf, ws.lf.data(), ws.lf.size() * sizeof(Lvl<float>),
                              hipMemcpyHostToDevice, s));
    if (ws.d_ld)
        HIPCHK(hipMemcpyAsync(ws.d_ld, ws.ld.data(), ws.ld.size() * sizeof(Lvl<double>),
                              hipMemcpyHostToDevice, s));
}

// Puts a tuned plan's band heights and bottom size into the workspace (0 / -1 where it keeps the
// built-in rules).
static void apply_plan(Workspace &ws, const Plan &p) {
    for (size_t l = 0; l < ws.lf.size(); l++) {
        ws.lf[l].th_pre = l < p.lvl.size() ? p.lvl[l][0] : 0;
        ws.lf[l].th_post = l < p.lvl.size() ? p.lvl[l][1] : 0;
    }
    for (size_t l = 0; l < ws.ld.size(); l++) {
        ws.ld[l].th_pre = l < p.lvl.size() ? p.lvl[l][0] : 0;
        ws.ld[l].th_post = l < p.lvl.size() ? p.lvl[l][1] : 0;
    }
    ws.th_dir = p.th_dir, ws.th_resid = p.th_resid, ws.th_fin1 = p.th_fin1;
    ws.th_fin2 = p.th_fin2, ws.bottom = p.bottom;
}

// Sizes the context's workspace for a w x h grid and the cycle's precision (kept while the shape
// stays).
static void fit(dtd_ctx *ctx, int w, int h, int wrap) {
    Workspace &ws = ctx->ws;
    if (ws.w == w && ws.h == h && ws.wrap == wrap && ws.prec == ctx->opts.cycle_precision) return;
    free_ws(ws);
    ctx->cycle_levels = 0;  // a cycle built by dtd_cycle_new belonged to the old shape
    auto take = [&](size_t bytes) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, bytes));
        ws.owned.push_back(p);
        return p;
    };
    lay_out(ws, ctx->opts, w, h, wrap, false, take);
    auto it = ctx->plans.find(PlanKey{w, h, wrap, ctx->opts.cycle_precision});
    if (it != ctx->plans.end()) apply_plan(ws, it->second);
    upload_levels(ws, ctx->stream);
}

static Grid grid_of(const Workspace &ws) { return Grid{ws.w, ws.h, ws.wrap, (int)ws.n}; }

// ---- device memory for callers that cannot allocate (no reference counterpart) ---------------
//
// dtd_solve_device_noalloc lays its workspace out in one fixed array per GPU instead of
// hipMalloc'ed buffers.  The array exists on every GPU the library is loaded on, from then until
// the process ends.  In host code its name stands for no device address: the address on each
// device comes from hipGetSymbolAddress, looked up once per device by dtd_create (with that
// device current) and kept in scratch_base.  Every call on one GPU shares that GPU's array.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static void *scratch_base[kMaxDevices];

// Looks up the array's address on the current device `dev` (once).
static void scratch_init(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw Error{DTD_EINVAL, "device ordinal out of range"};
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_base[dev]) return;
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error{DTD_EDEVICE, "the scratch array has no device address", true};
    scratch_base[dev] = p;
}

static void *scratch_of(int dev) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    return dev >= 0 && dev < kMaxDevices ? scratch_base[dev] : nullptr;
}

// Carves consecutive 256-byte aligned pieces from [base, base + cap); with base null it only
// counts the bytes (the size check before any device work).
struct Carver {
    unsigned char *base;
    size_t cap, used = 0;
    void *operator()(size_t bytes) {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + bytes;
        if (base && used > cap)
            throw Error{DTD_ETOOBIG, "the grid's workspace exceeds the fixed scratch array"};
        return base ? base + off : nullptr;
    }
};

// Puts a scratch-backed workspace in the context's place for one call (direct launches, no
// graphs) and gives the context its own workspace back afterwards, also when the call throws.
struct ScratchSwap {
    dtd_ctx *ctx;
    Workspace saved;
    bool graph;
    ScratchSwap(dtd_ctx *c, Workspace &&sw)
        : ctx(c), saved(std::move(c->ws)), graph(c->use_graph) {
        c->ws = std::move(sw);
        c->use_graph = false;
    }
    ~ScratchSwap() {
        ctx->ws = std::move(saved);
        ctx->use_graph = graph;
    }
};

__global__ void k_sum1(const double *__restrict__ parts, int nparts, double *out) {
    __shared__ double lds[kReduceBlock / 64];
    const double v = reduce_parts(parts, nparts, 1, 0, lds);
    if (threadIdx.x == 0) out[0] = v;
}

// A fresh state for a solve: tol2 (or, with rel2 > 0, tol2 = rel2 * b.b from the first residual),
// the iteration cap, and the fields the operator's faces come from.
__global__ void k_init_state(CgState *st, double tol2, double rel2, int max_iter, const double *kd,
                             double dt, double s2) {
    if (threadIdx.x != 0) return;
    CgState s{};
    s.tol2 = tol2;
    s.rel2 = rel2;
    s.max_iter = max_iter;
    s.kd = kd;
    s.dt = dt;
    s.s2 = s2;
    *st = s;
}

// A state that never stops and takes every branch of the passes (beta and the pending update on,
// both zero), for timing the loop's kernels on resident data without changing it (the faces from
// the workspace's own kd buffer).
__global__ void k_fill64(double *a, long n, double v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_time_state(CgState *st, const double *kd) {
    if (threadIdx.x != 0) return;
    CgState s{};
    s.max_iter = 1 << 30;
    s.have_beta = 1;
    s.have_pending = 1;
    s.kd = kd;
    s.dt = 1.0;
    s.s2 = 1.0;
    *st = s;
}

// The coarse levels from the finest (`Hierarchy::new`, MG:481-487; device `COARSEN`/`CDIAG`,
// MGGPU:62-140): one launch a level.
template <typename T>
static void build_coarse(std::vector<Lvl<T>> &lv, T scale, hipStream_t s, size_t from = 1) {
    for (size_t l = from; l < lv.size(); l++) {
        k_coarsen<T><<<nblocks(lv[l].n), kBlock, 0, s>>>(lv[l - 1], lv[l], scale);
        LAUNCH_CHECK();
    }
    // The exact coarsest solve's factor (an option), from the coarsest level's records.
    const Lvl<T> &C = lv.back();
    if (C.chol) {
        if (C.n > 64) throw Error{DTD_ESTATE, "the coarsest level exceeds one wave"};
        k_chol<T><<<1, 64, 0, s>>>(C);
        LAUNCH_CHECK();
    }
}

// The bf16 copies of the records of the levels that keep them (after coarsening, which reads the
// f32 records).
static void pack_bf16(Workspace &ws, hipStream_t s) {
    for (auto &L : ws.lf)
        if (L.coefh) {
            k_pack_bf16<<<nblocks(L.n), 256, 0, s>>>(L.n, L.coef, L.omega, L.coefh);
            LAUNCH_CHECK();
        }
}

static bool kd_records(const dtd_ctx *ctx);

// `Operator::new` on the device from kd (already on the device), with the cycle's levels.
static void build_operator(dtd_ctx *ctx, const double *kd, double dx, double dt) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    ctx->cycle_levels = 0;  // the levels now hold this solve's operator, not dtd_cycle_new's
    const bool f64 = ws.prec == DTD_CYCLE_F64;
    if (!f64 && kd_records(ctx)) {
        // The fused finest smoothers form the finest records from kd, and nothing else of the
        // solve reads them: level 1 straight from kd, no finest records stored.
        const dim3 cg((ws.lf[1].w + 63) / 64,
                      (ws.lf[1].h + (kBlock / 64) * kCoarsenRows - 1) / ((kBlock / 64) * kCoarsenRows));
        k_coarsen_kd<<<cg, kBlock, 0, s>>>(ws.lf[0], ws.lf[1], ctx->opts.scale,
                                                            kd, dt, dx * dx);
        LAUNCH_CHECK();
        build_coarse<float>(ws.lf, ctx->opts.scale, s, 2);
        pack_bf16(ws, s);
        return;
    }
    // The f64 passes form the faces from kd themselves; only the cycle's finest records are kept
    // (and, for the reference's summation order, the f64 faces its unfused passes read).
    k_operator<<<nblocks(ws.n), kBlock, 0, s>>>(grid_of(ws), kd, dt, dx * dx,
                                                ctx->opts.sums == DTD_SUMS_STRIPS ? ws.es : nullptr,
                                                f64 ? nullptr : ws.lf[0].coef,
                                                f64 ? ws.ld[0].coef : nullptr);
    LAUNCH_CHECK();
    if (f64)
        build_coarse<double>(ws.ld, (double)ctx->opts.scale, s);
    else
        build_coarse<float>(ws.lf, ctx->opts.scale, s);
    pack_bf16(ws, s);
}

// Returns the block count (with RM = kRUpdate, the r.r partial sums left in parts).  A level
// with bf16 records is smoothed from them (CB = 1).
template <typename T, int V, int RM, int CB>
static int launch_pre_cb(Lvl<T> &L, Lvl<T> &C, T *zout, const CgState *st, hipStream_t s,
                         R64 rio, double *parts, long cap) {
    const int th = L.th_pre > 0 ? L.th_pre : band_height<V>((const void *)k_spre<T, V, RM, CB>, L.w, L.h);
    const int nb = stream_grid<V, LaneGeom<V>::halo>(L.w, L.h, th);
    if (rm_update(RM)) need_parts(nb, cap);
    k_spre<T, V, RM, CB><<<nb, kBlock, 0, s>>>(L, zout, C, th, st, rio, parts);
    LAUNCH_CHECK();
    return nb;
}
template <typename T, int V, int RM = kRLevel>
static int launch_pre(Lvl<T> &L, Lvl<T> &C, T *zout, const CgState *st, hipStream_t s,
                      R64 rio = R64{}, double *parts = nullptr, long cap = 0) {
    if constexpr (std::is_same<T, float>::value)
        if (L.coefh) return launch_pre_cb<T, V, RM, 1>(L, C, zout, st, s, rio, parts, cap);
    return launch_pre_cb<T, V, RM, 0>(L, C, zout, st, s, rio, parts, cap);
}

template <typename T, int V, int RM, int CB>
static int launch_post_cb(Lvl<T> &L, const T *zin, const T *zc, int wc, T *zout,
                          const double *dot, double *parts, long cap, const CgState *st,
                          hipStream_t s) {
    const int th =
        L.th_post > 0 ? L.th_post : band_height<V>((const void *)k_spost<T, V, RM, CB>, L.w, L.h);
    const int nb = stream_grid<V, LaneGeom<V>::halo>(L.w, L.h, th);
    if (dot) need_parts(nb, cap);
    k_spost<T, V, RM, CB><<<nb, kBlock, 0, s>>>(L, zin, zc, wc, zout, th, dot, parts, st);
    LAUNCH_CHECK();
    return nb;
}
template <typename T, int V, int RM = kRLevel>
static int launch_post(Lvl<T> &L, const T *zin, const T *zc, int wc, T *zout, const double *dot,
                       double *parts, long cap, const CgState *st, hipStream_t s) {
    if constexpr (std::is_same<T, float>::value)
        if (L.coefh)
            return launch_post_cb<T, V, RM, 1>(L, zin, zc, wc, zout, dot, parts, cap, st, s);
    return launch_post_cb<T, V, RM, 0>(L, zin, zc, wc, zout, dot, parts, cap, st, s);
}

// The streamed direction pass; returns its block count (the p.Ap partial sums).
template <typename TZ>
static int launch_dir(Workspace &ws, const TZ *z, const double *pold, double *pnew,
                      hipStream_t s, double *ap) {
    const int th = ws.th_dir > 0 ? ws.th_dir : band_height<1, 1>((const void *)k_sdir<TZ>, ws.w, ws.h);
    const int nb = stream_grid<1, 1>(ws.w, ws.h, th);
    need_parts(nb, ws.nparts);
    k_sdir<TZ><<<nb, kBlock, 0, s>>>(grid_of(ws), z, pold, pnew, ap, ws.x, th, ws.parts, ws.st);
    LAUNCH_CHECK();
    return nb;
}

// The first residual r = b - A xin (xout = xin's copy when they differ); returns the block count
// (the (r.r, b.b) partial pairs).
static int launch_resid(Workspace &ws, const double *kd, double dt, double s2, const double *b,
                        const double *xin, hipStream_t s, bool r32 = true) {
    const Grid g = grid_of(ws);
    if (ws.prec == DTD_CYCLE_F64) {
        const int th = ws.th_resid > 0 ? ws.th_resid
                                       : band_height<1, 1>((const void *)k_sresid<double>, ws.w, ws.h);
        const int nb = stream_grid<1, 1>(ws.w, ws.h, th);
        need_parts(nb, ws.nparts);
        k_sresid<double><<<nb, kBlock, 0, s>>>(g, kd, dt, s2, b, xin, ws.x, ws.r, (double *)nullptr,
                                               th, ws.parts);
        LAUNCH_CHECK();
        return nb;
    }
    const int th = ws.th_resid > 0 ? ws.th_resid
                                   : band_height<1, 1>((const void *)k_sresid<float>, ws.w, ws.h);
    const int nb = stream_grid<1, 1>(ws.w, ws.h, th);
    need_parts(nb, ws.nparts);
    k_sresid<float><<<nb, kBlock, 0, s>>>(g, kd, dt, s2, b, xin, ws.x, ws.r, r32 ? ws.lf[0].r : nullptr,
                                          th, ws.parts);
    LAUNCH_CHECK();
    return nb;
}

// Levels of at most this many cells (below the finest) are cycled by one block, k_bottom
// (DTD_BOTTOM_CELLS overrides it; 0 turns it off).
static int bottom_cells() {
    static const int v = [] {
        const char *e = getenv("DTD_BOTTOM_CELLS");
        return e ? atoi(e) : 2048;
    }();
    return v;
}

// The fused smoothing kernels serve every level but the coarsest when nu = 2 (DTD_NO_FUSE=1
// turns them off, for comparison).
static bool fused_enabled() {
    static const bool on = [] {
        const char *v = getenv("DTD_NO_FUSE");
        return !(v && *v && *v != '0');
    }();
    return on;
}

// `vcycle` (MG:500-523) from level l, enqueued on s.  The first sweep writes z[0] and the
// sweeps alternate after it; with 2 nu - 1 sweeps after the first on every level but the
// coarsest (and COARSE_SWEEPS - 1 = 15 there), each level's result lands in z[1].  With dot,
// the last sweep of level l leaves its per-block sums dot * z in parts (block count returned).
// rm (f32 levels, fused smoothing, finest level only): where the finest level's r comes from
// (kRLevel: its own r; kR64: the f64 residual dot; kRUpdate: rio.in - alpha A p, written to
// rio.out = dot, with the r.r step of the conjugate gradients launched after the pre-smoothing).
// parts holds the partial sums of at most cap blocks.  Levels of at most `bottom` cells (< 0:
// the built-in size) go to the one-block bottom.
template <typename T>
static int vcycle(std::vector<Lvl<T>> &lv, size_t l, int nu, const double *dot, double *parts,
                  long cap, CgState *st, hipStream_t s, int rm = kRLevel, R64 rio = R64{},
                  const Lvl<T> *dlv = nullptr, int bottom = -1) {
    Lvl<T> &L = lv[l];
    size_t lds = 0;
    for (size_t k = l; k < lv.size(); k++) lds += (kRec + 3) * (size_t)lv[k].n * sizeof(T);
    if (bottom < 0) bottom = bottom_cells();
    if (dlv && l >= 1 && nu == 2 && fused_enabled() && L.n <= bottom && !L.coefh &&
        lds <= kBottomLds && lv.size() - l <= (size_t)kBottomLevels) {
        // This level and every coarser one in one block, in LDS.
        k_bottom<T><<<1, kBottomBlock, lds, s>>>(dlv, (int)l, (int)lv.size(), st);
        LAUNCH_CHECK();
        return 1;
    }
    if (l + 1 == lv.size()) {
        if (L.n > 64) throw Error{DTD_ESTATE, "the coarsest level exceeds one wave"};
        k_coarsest<T><<<1, 64, 0, s>>>(L, L.z[1], 16, dot, parts, st);
        LAUNCH_CHECK();
        return 1;
    }
    const int nb = nblocks(L.n);
    if (dot) need_parts(nb, cap);
    Lvl<T> &C0 = lv[l + 1];
    if (nu == 2 && fused_enabled() && !L.bil) {
        // Fused smoothing (kernels.hpp): pre-smoothing and restriction in one launch, the coarse
        // cycle, then both post-smoothing sweeps in one launch; 8-byte lanes when the width
        // allows.
        constexpr int VW = 8 / sizeof(T);
        const bool wide = L.w % VW == 0;
        if constexpr (std::is_same<T, float>::value) {
            if (rm != kRLevel) {
                if (l != 0) throw Error{DTD_ESTATE, "the f64 residual feeds the finest level only"};
                if (rm_kd(rm) && L.coefh)
                    throw Error{DTD_ESTATE, "kd-formed records with bf16 records"};
                if (rm_update(rm)) {
                    int nbp;
                    if (rm == kRUpdateK)
                        nbp = wide ? launch_pre<T, VW, kRUpdateK>(L, C0, L.z[0], st, s, rio, parts, cap)
                                   : launch_pre<T, 1, kRUpdateK>(L, C0, L.z[0], st, s, rio, parts, cap);
                    else
                        nbp = wide ? launch_pre<T, VW, kRUpdate>(L, C0, L.z[0], st, s, rio, parts, cap)
                                   : launch_pre<T, 1, kRUpdate>(L, C0, L.z[0], st, s, rio, parts, cap);
                    k_cg_step<kStepRR><<<1, kReduceBlock, 0, s>>>(parts, nbp, st);
                    LAUNCH_CHECK();
                } else if (rm == kR64K) {
                    wide ? launch_pre<T, VW, kR64K>(L, C0, L.z[0], st, s, rio)
                         : launch_pre<T, 1, kR64K>(L, C0, L.z[0], st, s, rio);
                } else if (wide) {
                    launch_pre<T, VW, kR64>(L, C0, L.z[0], st, s, rio);
                } else {
                    launch_pre<T, 1, kR64>(L, C0, L.z[0], st, s, rio);
                }
                vcycle<T>(lv, l + 1, nu, nullptr, parts, cap, st, s, kRLevel, R64{}, dlv, bottom);
                if (rm_kd(rm))
                    return wide ? launch_post<T, VW, kR64K>(L, L.z[0], C0.z[1], C0.w, L.z[1], dot,
                                                             parts, cap, st, s)
                                : launch_post<T, 1, kR64K>(L, L.z[0], C0.z[1], C0.w, L.z[1], dot,
                                                            parts, cap, st, s);
                return wide ? launch_post<T, VW, kR64>(L, L.z[0], C0.z[1], C0.w, L.z[1], dot, parts,
                                                        cap, st, s)
                            : launch_post<T, 1, kR64>(L, L.z[0], C0.z[1], C0.w, L.z[1], dot, parts,
                                                       cap, st, s);
            }
        }
        if (wide)
            launch_pre<T, VW>(L, C0, L.z[0], st, s);
        else
            launch_pre<T, 1>(L, C0, L.z[0], st, s);
        vcycle<T>(lv, l + 1, nu, nullptr, parts, cap, st, s, kRLevel, R64{}, dlv, bottom);
        return wide ? launch_post<T, VW>(L, L.z[0], C0.z[1], C0.w, L.z[1], dot, parts, cap, st, s)
                    : launch_post<T, 1>(L, L.z[0], C0.z[1], C0.w, L.z[1], dot, parts, cap, st, s);
    }
    int cur = 0;
    k_first<T><<<nb, kBlock, 0, s>>>(L, L.z[0], st);
    LAUNCH_CHECK();
    for (int k = 1; k < nu; k++) {
        k_sweep<T><<<nb, kBlock, 0, s>>>(L, L.z[cur], L.z[cur ^ 1], nullptr, 0, nullptr, nullptr,
                                         st);
        LAUNCH_CHECK();
        cur ^= 1;
    }
    k_restrict<T><<<nblocks(C0.n), kBlock, 0, s>>>(L, L.z[cur], C0, st);
    LAUNCH_CHECK();
    vcycle<T>(lv, l + 1, nu, nullptr, parts, cap, st, s, kRLevel, R64{}, dlv, bottom);
    {
        const double *d = nu == 1 ? dot : nullptr;
        if (L.bil) {
            // z' = z + P zc in place, then the sweep from z'.
            k_add_prolong<T><<<nb, kBlock, 0, s>>>(L, C0, L.z[cur], st);
            LAUNCH_CHECK();
            k_sweep<T><<<nb, kBlock, 0, s>>>(L, L.z[cur], L.z[cur ^ 1], nullptr, 0, d, parts, st);
        } else {
            k_sweep<T><<<nb, kBlock, 0, s>>>(L, L.z[cur], L.z[cur ^ 1], C0.z[1], C0.w, d, parts,
                                             st);
        }
        LAUNCH_CHECK();
        cur ^= 1;
    }
    for (int k = 1; k < nu; k++) {
        const double *d = k + 1 == nu ? dot : nullptr;
        k_sweep<T><<<nb, kBlock, 0, s>>>(L, L.z[cur], L.z[cur ^ 1], nullptr, 0, d, parts, st);
        LAUNCH_CHECK();
        cur ^= 1;
    }
    if (cur != 1) throw Error{DTD_ESTATE, "the cycle's result is not in z[1]"};
    return nb;
}

static int run_vcycle(dtd_ctx *ctx, const double *dot, CgState *st, int rm = kRLevel,
                      R64 rio = R64{}) {
    Workspace &ws = ctx->ws;
    const int nu = (int)ctx->opts.nu;
    if (ws.prec == DTD_CYCLE_F64)
        return vcycle<double>(ws.ld, 0, nu, dot, ws.parts, ws.nparts, st, ctx->stream, kRLevel,
                              R64{}, ws.d_ld, ws.bottom);
    return vcycle<float>(ws.lf, 0, nu, dot, ws.parts, ws.nparts, st, ctx->stream, rm, rio,
                         ws.d_lf, ws.bottom);
}

// Whether the f32 cycle's finest level takes the f64 residual straight from the conjugate
// gradients (fused smoothing; the residual's update rides on the pre-smoothing pass).
static bool fused_update(const Workspace &ws, const dtd_opts &o) {
    return ws.prec != DTD_CYCLE_F64 && o.nu == 2 && fused_enabled() && ws.lf.size() > 1 &&
           o.transfer == DTD_TRANSFER_AGGREGATE;
}

// Whether the fused finest level forms its records (and the pre-smoothing A p) from kd and p:
// the f32 cycle without bf16 records on the finest level.
static bool kd_records(const dtd_ctx *ctx) {
    return ctx->use_kd && ctx->opts.sums != DTD_SUMS_STRIPS && fused_update(ctx->ws, ctx->opts) &&
           !ctx->ws.lf[0].coefh;
}

// One conjugate-gradient iteration (MG:630-733): parity 0 reads the direction from p1 and writes
// p0 (odd iterations), parity 1 the other way.
static void enqueue_iteration(dtd_ctx *ctx, int parity) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    double *pold = parity == 0 ? ws.p1 : ws.p0;
    double *pnew = parity == 0 ? ws.p0 : ws.p1;
    const bool kd = kd_records(ctx);
    const int nbd = ws.prec == DTD_CYCLE_F64
                        ? launch_dir<double>(ws, ws.ld[0].z[1], pold, pnew, s, ws.ap)
                        : launch_dir<float>(ws, ws.lf[0].z[1], pold, pnew, s, kd ? nullptr : ws.ap);
    k_cg_step<kStepPAP><<<1, kReduceBlock, 0, s>>>(ws.parts, nbd, ws.st);
    LAUNCH_CHECK();
    if (fused_update(ws, ctx->opts)) {
        // r_new = r_old - alpha A p inside the finest pre-smoothing (r.r and the stop test after
        // it); the residual's buffers alternate with the parity: parity 0 reads r, writes r2.
        // With kd-formed records the pre-smoothing forms A p from p itself.
        double *rin = parity == 0 ? ws.r : ws.r2, *rout = parity == 0 ? ws.r2 : ws.r;
        const int nbz = kd ? run_vcycle(ctx, rout, ws.st, kRUpdateK,
                                        R64{rin, nullptr, rout, 0.0, pnew, nullptr, 0.0, 0.0})
                           : run_vcycle(ctx, rout, ws.st, kRUpdate, R64{rin, ws.ap, rout, 0.0});
        k_cg_step<kStepRZ><<<1, kReduceBlock, 0, s>>>(ws.parts, nbz, ws.st,
                                                       ctx->dev_flag + parity);
        LAUNCH_CHECK();
        return;
    }
    const int nbu = nblocks(ws.n, kCgCells);
    if (ws.prec == DTD_CYCLE_F64)
        k_update<double><<<nbu, kBlock, 0, s>>>((int)ws.n, ws.ap, ws.r, nullptr, ws.parts, ws.st);
    else
        k_update<float><<<nbu, kBlock, 0, s>>>((int)ws.n, ws.ap, ws.r, ws.lf[0].r, ws.parts,
                                               ws.st);
    LAUNCH_CHECK();
    k_cg_step<kStepRR><<<1, kReduceBlock, 0, s>>>(ws.parts, nbu, ws.st);
    LAUNCH_CHECK();
    const int nbz = run_vcycle(ctx, ws.r, ws.st);
    k_cg_step<kStepRZ><<<1, kReduceBlock, 0, s>>>(ws.parts, nbz, ws.st, ctx->dev_flag + parity);
    LAUNCH_CHECK();
}

static void capture_graphs(dtd_ctx *ctx) {
    Workspace &ws = ctx->ws;
    for (int k = 0; k < 2; k++) {
        if (ws.graph[k]) continue;
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_iteration(ctx, k);
        } catch (...) {
            (void)hipStreamEndCapture(ctx->stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        HIPCHK(hipStreamEndCapture(ctx->stream, &gr));
        hipError_t e = hipGraphInstantiate(&ws.graph[k], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHK(e);
    }
}

// Whether the parallel line solver takes lines of n cells: 16 * 2^k cells, 4 to 1024 segments.
static bool parallel_line(int n) {
    if (n % kSeg) return false;
    const int p = n / kSeg;
    return p >= 4 && p <= 1024 && (p & (p - 1)) == 0;
}

static bool parallel_lod_enabled() {
    static const bool on = [] {
        const char *v = getenv("DTD_SEQ_LOD");
        return !(v && *v && *v != '0');
    }();
    return on;
}

// `lod` (D:370-489): the splitting's first guess of b into x (x may be b).
static void lod(dtd_ctx *ctx, const double *b, double *x, const double *kd, double dx, double dt,
                int wrap_x) {
    Workspace &ws = ctx->ws;
    const Grid g = grid_of(ws);
    const double s2 = dx * dx;
    hipStream_t s = ctx->stream;
    if (parallel_lod_enabled() && ctx->opts.sums != DTD_SUMS_STRIPS && parallel_line(ws.w) &&
        parallel_line(ws.h)) {
        // Rows from b into x; then the columns in place by the segments' reduced equations (in
        // p0, n / 2 values).
        auto lines = [&](int nlines, int n, int ring, const double *k, const double *src,
                         double *dst) {
            if (n / kSeg <= 512)
                k_lod_lines<512><<<nlines, n / kSeg, 0, s>>>(n, ring, k, src, dst, dt, s2);
            else
                k_lod_lines<1024><<<nlines, n / kSeg, 0, s>>>(n, ring, k, src, dst, dt, s2);
            LAUNCH_CHECK();
        };
        lines(ws.h, ws.w, wrap_x && ws.w > 2, kd, b, x);
        const dim3 cg((ws.w + 63) / 64, (ws.h / kSeg + 3) / 4);
        k_col_reduce<<<cg, kBlock, 0, s>>>(ws.w, ws.h, kd, x, dt / s2, ws.p0);
        LAUNCH_CHECK();
        const int N = 2 * (ws.h / kSeg);
        if (N % kSeg == 0 && N >= 2 * kSeg) {
            // Long columns: the reduced systems by a second partition (its equations after red,
            // at p0 + n / 2).
            double *red2 = ws.p0 + ws.n / 2;
            const dim3 gg((ws.w + 63) / 64, (N / kSeg + 3) / 4);
            k_tri_group<<<gg, kBlock, 0, s>>>(ws.w, N, ws.p0, red2, 0);
            LAUNCH_CHECK();
            k_tri_thomas<<<(ws.w + 63) / 64, 64, 0, s>>>(ws.w, 2 * (N / kSeg), red2);
            LAUNCH_CHECK();
            k_tri_group<<<gg, kBlock, 0, s>>>(ws.w, N, ws.p0, red2, 1);
            LAUNCH_CHECK();
        } else {
            k_col_thomas<<<(ws.w + 63) / 64, 64, 0, s>>>(ws.w, ws.h, ws.p0);
            LAUNCH_CHECK();
        }
        k_col_back<<<cg, kBlock, 0, s>>>(ws.w, ws.h, kd, x, dt / s2, ws.p0);
        LAUNCH_CHECK();
        return;
    }
    if (x != b)
        HIPCHK(hipMemcpyAsync(x, b, ws.n * sizeof(double), hipMemcpyDeviceToDevice, s));
    k_lod_rows<<<(ws.h + kLineBlock - 1) / kLineBlock, kLineBlock, 0, s>>>(g, wrap_x, kd, dt, s2, x,
                                                                           ws.p0, ws.p1);
    LAUNCH_CHECK();
    k_lod_cols<<<(ws.w + kLineBlock - 1) / kLineBlock, kLineBlock, 0, s>>>(g, kd, dt, s2, x, ws.p0);
    LAUNCH_CHECK();
}

enum Guess { kGuessLod, kGuessRhs, kGuessKeep };

// The device entry points' last step: once the solve has finished without error, its answer
// (ws.x) is copied into the caller's h_dev.  A failure before this point leaves h_dev as it was.
static void copy_answer(dtd_ctx *ctx, double *h_dev) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpyAsync(h_dev, ctx->ws.x, ctx->ws.n * sizeof(double), hipMemcpyDeviceToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

// The last update, the put-back and the true residual, as solve_core ends (k_finish2 from x' in
// ws.ap with the state's lost and weight); the pair of sums (r.r, b.b) into ws.pair.
static void finish_put_back(dtd_ctx *ctx, const double *b, const double *kd, double dx, double dt) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    const int th2 = ws.th_fin2 > 0 ? ws.th_fin2
                                   : band_height<1, 1>((const void *)k_finish2, ws.w, ws.h);
    const int nb2 = stream_grid<1, 1>(ws.w, ws.h, th2);
    need_parts(nb2, ws.nparts);
    k_finish2<<<nb2, kBlock, 0, s>>>(grid_of(ws), kd, dt, dx * dx, b, ws.ap, ws.x, th2, ws.parts,
                                     ws.st);
    LAUNCH_CHECK();
    k_pair_sum<<<1, kReduceBlock, 0, s>>>(ws.parts, nb2, ws.pair);
    LAUNCH_CHECK();
}

// solve_core with the reference's summation order (DTD_SUMS_STRIPS): the same solve with every
// sum of the loop taken per strip of 8 rows by one lane, the strips in order (k_strip_sums,
// k_strips_step), and the put-back's per row (k_putback_rows, k_putback_fold), so the iterates,
// the iteration count and h1 are the reference's bit for bit.  The passes run unfused on the
// stored f64 faces (the cycle on the stored f32 records), the first guess is the splitting's own
// recurrences, and the host reads the stop flag after each iteration (no graphs).
static size_t solve_core_strips(dtd_ctx *ctx, const double *b, const double *kd, double dx,
                                double dt, int wrap_x, double tol_m, double rel2, bool setup,
                                Guess guess, dtd_stats *stats, double *rel, double *bnorm) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    const Grid g = grid_of(ws);
    if (!ws.sums || ws.lf.empty()) throw Error{DTD_ESTATE, "the workspace has no strip sums"};
    const int n = (int)ws.n, nb = nblocks(ws.n);
    const int ns = (ws.h + kStripRows - 1) / kStripRows, sb = (ns + 63) / 64;
    const long strip_cells = (long)kStripRows * ws.w;
    double *sums = ws.sums, *bbs = ws.sums + ns, *rows = ws.sums + 2 * ns;
    float *r32 = ws.lf[0].r, *z32 = ws.lf[0].z[1];
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    if (setup) build_operator(ctx, kd, dx, dt);
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    if (guess == kGuessLod && ctx->opts.guess == DTD_GUESS_LOD)
        lod(ctx, b, ws.x, kd, dx, dt, wrap_x);
    else if (guess != kGuessKeep)
        HIPCHK(hipMemcpyAsync(ws.x, b, ws.n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipEventRecord(ctx->ev[2], s));
    k_init_state<<<1, 64, 0, s>>>(ws.st, tol_m * tol_m * (double)ws.n, rel2,
                                  (int)ctx->opts.max_iter, kd, dt, dx * dx);
    LAUNCH_CHECK();
    // r = b - A x, r.r (MG:596-626); the cycle's z and r.z (MG:627).
    k_resid_es<<<nb, kBlock, 0, s>>>(g, ws.es, b, ws.x, ws.r, r32);
    LAUNCH_CHECK();
    k_strip_sums<double><<<sb, 64, 0, s>>>(ws.r, ws.r, ws.n, strip_cells, ns, sums, nullptr);
    LAUNCH_CHECK();
    if (rel2 > 0.0) {
        k_strip_sums<double><<<sb, 64, 0, s>>>(b, b, ws.n, strip_cells, ns, bbs, nullptr);
        LAUNCH_CHECK();
    }
    k_strips_step<kStepRR0><<<1, 64, 0, s>>>(sums, ns, rel2 > 0.0 ? bbs : nullptr, ws.st, nullptr);
    LAUNCH_CHECK();
    run_vcycle(ctx, nullptr, ws.st);
    k_strip_sums<float><<<sb, 64, 0, s>>>(ws.r, z32, ws.n, strip_cells, ns, sums, ws.st);
    LAUNCH_CHECK();
    k_strips_step<kStepRZ0><<<1, 64, 0, s>>>(sums, ns, nullptr, ws.st, ctx->dev_flag);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(s));
    // The loop (MG:630-733): iteration k (1-based) writes its direction to p0 when k is odd.
    const int max_iter = (int)ctx->opts.max_iter;
    for (int k = 0; k < max_iter && !ctx->host_flag[0]; k++) {
        const double *pold = (k & 1) ? ws.p0 : ws.p1;
        double *pnew = (k & 1) ? ws.p1 : ws.p0;
        k_dir_cells<<<nb, kBlock, 0, s>>>(n, z32, pold, pnew, ws.x, ws.st);
        LAUNCH_CHECK();
        k_apply<<<nb, kBlock, 0, s>>>(g, ws.es, pnew, ws.ap);
        LAUNCH_CHECK();
        k_strip_sums<double><<<sb, 64, 0, s>>>(pnew, ws.ap, ws.n, strip_cells, ns, sums, ws.st);
        LAUNCH_CHECK();
        k_strips_step<kStepPAP><<<1, 64, 0, s>>>(sums, ns, nullptr, ws.st, nullptr);
        LAUNCH_CHECK();
        k_update<float><<<nblocks(ws.n, kCgCells), kBlock, 0, s>>>(n, ws.ap, ws.r, r32, ws.parts,
                                                                   ws.st);
        LAUNCH_CHECK();
        k_strip_sums<double><<<sb, 64, 0, s>>>(ws.r, ws.r, ws.n, strip_cells, ns, sums, ws.st);
        LAUNCH_CHECK();
        k_strips_step<kStepRR><<<1, 64, 0, s>>>(sums, ns, nullptr, ws.st, nullptr);
        LAUNCH_CHECK();
        run_vcycle(ctx, nullptr, ws.st);
        k_strip_sums<float><<<sb, 64, 0, s>>>(ws.r, z32, ws.n, strip_cells, ns, sums, ws.st);
        LAUNCH_CHECK();
        k_strips_step<kStepRZ><<<1, 64, 0, s>>>(sums, ns, nullptr, ws.st, ctx->dev_flag);
        LAUNCH_CHECK();
        HIPCHK(hipStreamSynchronize(s));
    }
    // x' = x + alpha p (the last direction) into ap; A x' into r; the put-back's sums per row,
    // the rows in order; then the put-back and the true residual.
    k_last_update<<<nb, kBlock, 0, s>>>(n, ws.x, ws.p0, ws.p1, ws.ap, ws.st);
    LAUNCH_CHECK();
    k_apply<<<nb, kBlock, 0, s>>>(g, ws.es, ws.ap, ws.r);
    LAUNCH_CHECK();
    k_putback_rows<<<(ws.h + 63) / 64, 64, 0, s>>>(g, ws.es, b, ws.r, rows);
    LAUNCH_CHECK();
    k_putback_fold<<<1, 64, 0, s>>>(rows, ws.h, ws.st);
    LAUNCH_CHECK();
    finish_put_back(ctx, b, kd, dx, dt);
    HIPCHK(hipEventRecord(ctx->ev[3], s));
    HIPCHK(hipMemcpyAsync(ctx->host_st, ws.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ctx->host_pair, ws.pair, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double rr = ctx->host_pair[0], bb = ctx->host_pair[1];
    if (rel) *rel = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    if (bnorm) *bnorm = std::sqrt(bb);
    if (stats) {
        stats->iterations += (size_t)ctx->host_st->its;
        stats->ms_setup += elapsed(ctx->ev[0], ctx->ev[1]);
        stats->ms_guess += elapsed(ctx->ev[1], ctx->ev[2]);
        stats->ms_cg += elapsed(ctx->ev[2], ctx->ev[3]);
        stats->ms_total += elapsed(ctx->ev[0], ctx->ev[3]);
    }
    return (size_t)ctx->host_st->its;
}

// One solve on the workspace (`multigrid_in` + `cg_on`, D:101-116, MG:562-763).  b is the
// right-hand side on the device (the caller's array or ws.b; only read); kd likewise.  The first
// guess is the splitting of b (or b itself when the context says so), or (kGuessKeep) the x
// already in ws.x.  With rel2 > 0 the tolerance is tol2 = rel2 * b.b, else tol_m^2 * n.  Leaves
// h1 in ws.x and returns the iterations; *rel gets ||b - A h1|| / ||b|| and *bnorm ||b||.
static size_t solve_core(dtd_ctx *ctx, const double *b, const double *kd, double dx, double dt,
                         int wrap_x, double tol_m, double rel2, bool setup, Guess guess,
                         dtd_stats *stats, double *rel, double *bnorm) {
    if (ctx->opts.sums == DTD_SUMS_STRIPS)
        return solve_core_strips(ctx, b, kd, dx, dt, wrap_x, tol_m, rel2, setup, guess, stats, rel,
                                 bnorm);
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    const Grid g = grid_of(ws);
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    if (setup) build_operator(ctx, kd, dx, dt);
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    const double *xin = ws.x;
    if (guess == kGuessLod && ctx->opts.guess == DTD_GUESS_LOD) {
        lod(ctx, b, ws.x, kd, dx, dt, wrap_x);
    } else if (guess != kGuessKeep) {
        xin = b;  // the first residual copies b into x
    }
    HIPCHK(hipEventRecord(ctx->ev[2], s));
    k_init_state<<<1, 64, 0, s>>>(ws.st, tol_m * tol_m * (double)ws.n, rel2,
                                  (int)ctx->opts.max_iter, kd, dt, dx * dx);
    LAUNCH_CHECK();
    const bool fused = fused_update(ws, ctx->opts);
    const int nbr = launch_resid(ws, kd, dt, dx * dx, b, xin, s, !fused);
    k_cg_step<kStepRR0><<<1, kReduceBlock, 0, s>>>(ws.parts, nbr, ws.st);
    LAUNCH_CHECK();
    const int nbz = fused ? run_vcycle(ctx, ws.r, ws.st, kd_records(ctx) ? kR64K : kR64,
                                       R64{ws.r, nullptr, nullptr, 0.0})
                          : run_vcycle(ctx, ws.r, ws.st);
    k_cg_step<kStepRZ0><<<1, kReduceBlock, 0, s>>>(ws.parts, nbz, ws.st);
    LAUNCH_CHECK();
    if (ctx->use_graph) capture_graphs(ctx);
    // The loop: iteration k (1-based) runs with parity (k - 1) % 2; the stop flag of iteration
    // k is read while iteration k + 1 runs.
    const int max_iter = (int)ctx->opts.max_iter;
    for (int k = 0; k < max_iter; k++) {
        const int par = k & 1;
        if (ctx->use_graph)
            HIPCHK(hipGraphLaunch(ws.graph[par], s));
        else
            enqueue_iteration(ctx, par);
        HIPCHK(hipEventRecord(ctx->flag_ev[par], s));
        if (k >= 1) {
            HIPCHK(hipEventSynchronize(ctx->flag_ev[par ^ 1]));
            if (ctx->host_flag[par ^ 1]) break;
        }
    }
    // The last update and the put-back's sums (x' into ap), the put-back and the true residual
    // (x'' into x).
    {
        const int th = ws.th_fin1 > 0 ? ws.th_fin1
                                      : band_height<1, 1>((const void *)k_finish1, ws.w, ws.h);
        const int nb = stream_grid<1, 1>(ws.w, ws.h, th);
        need_parts(nb, ws.nparts);
        k_finish1<<<nb, kBlock, 0, s>>>(g, kd, dt, dx * dx, b, ws.x, ws.p0, ws.p1, ws.ap, th, ws.parts,
                                        ws.st);
        LAUNCH_CHECK();
        k_putback_step<<<1, kReduceBlock, 0, s>>>(ws.parts, nb, ws.st);
        LAUNCH_CHECK();
        const int th2 = ws.th_fin2 > 0 ? ws.th_fin2
                                       : band_height<1, 1>((const void *)k_finish2, ws.w, ws.h);
        const int nb2 = stream_grid<1, 1>(ws.w, ws.h, th2);
        need_parts(nb2, ws.nparts);
        k_finish2<<<nb2, kBlock, 0, s>>>(g, kd, dt, dx * dx, b, ws.ap, ws.x, th2, ws.parts, ws.st);
        LAUNCH_CHECK();
        k_pair_sum<<<1, kReduceBlock, 0, s>>>(ws.parts, nb2, ws.pair);
        LAUNCH_CHECK();
    }
    HIPCHK(hipEventRecord(ctx->ev[3], s));
    HIPCHK(hipMemcpyAsync(ctx->host_st, ws.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ctx->host_pair, ws.pair, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double rr = ctx->host_pair[0], bb = ctx->host_pair[1];
    if (rel) *rel = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    if (bnorm) *bnorm = std::sqrt(bb);
    if (stats) {
        stats->iterations += (size_t)ctx->host_st->its;
        stats->ms_setup += elapsed(ctx->ev[0], ctx->ev[1]);
        stats->ms_guess += elapsed(ctx->ev[1], ctx->ev[2]);
        stats->ms_cg += elapsed(ctx->ev[2], ctx->ev[3]);
        stats->ms_total += elapsed(ctx->ev[0], ctx->ev[3]);
    }
    return (size_t)ctx->host_st->its;
}

// ---- the tuner (dtd_tune; no reference counterpart) -----------------------------------------
//
// The reference tunes its cycle by a hand-run sweep of settings (`tune_the_cycle`,
// multigrid/tests.rs:303-360).  Here the cycle's settings stay the reference's; what is tuned is
// how the work is cut into launches: the band height of every streamed launch (the fused pre- and
// post-smoothing of each level, the finest level's f64 passes) and the level size from which the
// one-block bottom of the cycle takes over.  Band heights are searched kernel by kernel, each
// candidate launched back to back on the resident data of a solve and timed with HIP events; the
// bottom size and the whole plan are then judged by timing whole solves, and the built-in rules
// stay unless the plan is faster.  A band height changes which cells a block's partial sums
// cover, so a tuned solve's sums round differently; the cycle's correction is unchanged.

// Band heights tried: short bands in many rounds of waves, up to tall ones that cover a finest
// level in about one round (the halo rows re-read then cost least).
static constexpr int kTuneTh[] = {4, 8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 72, 96, 128};
static constexpr int kTuneBottom[] = {256, 512, 1024, 2048};

// Average duration (ms) of `reps` back-to-back launches enqueued by f, after one untimed.
template <class F>
static double time_launches(dtd_ctx *ctx, int reps, F f) {
    f();
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    for (int k = 0; k < reps; k++) f();
    HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev[5]));
    return elapsed(ctx->ev[4], ctx->ev[5]) / reps;
}

// Median device time (ms) of `reps` whole solves (after one untimed), setup and first guess
// included, to the relative residual 1e-10; drops the captured graphs first (their launch shapes
// follow the plan).
static double time_solves(dtd_ctx *ctx, const double *b, const double *kd, double dx, double dt,
                          int wrap_x, int reps) {
    Workspace &ws = ctx->ws;
    for (auto &g : ws.graph)
        if (g) (void)hipGraphExecDestroy(g), g = nullptr;
    std::vector<double> t;
    for (int k = 0; k <= reps; k++) {
        dtd_stats st{};
        solve_core(ctx, b, kd, dx, dt, wrap_x, 0.0, 1e-20, true, kGuessLod, &st, nullptr, nullptr);
        if (k) t.push_back(st.ms_total);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

// The smallest-time candidate band height for one launch kind; set(th) applies it.
template <class Set, class Launch>
static int best_th(dtd_ctx *ctx, int rows, int reps, Set set, Launch launch) {
    int best = 0;
    double tbest = 0.0;
    for (int th : kTuneTh) {
        if (th > ((rows + 3) / 4) * 4 && th != kTuneTh[0]) break;
        set(th);
        const double t = time_launches(ctx, reps, launch);
        if (!best || t < tbest) best = th, tbest = t;
    }
    return best;
}

template <typename T>
static void tune_levels(dtd_ctx *ctx, std::vector<Lvl<T>> &lv, int reps, int min_cells) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    constexpr int VW = 8 / sizeof(T);
    for (size_t l = 0; l + 1 < lv.size(); l++) {
        Lvl<T> &L = lv[l], &C = lv[l + 1];
        if (l > 0 && L.n <= min_cells) break;  // from here the bottom block may take the levels
        const bool wide = L.w % VW == 0;
        const bool top = l == 0 && ws.prec != DTD_CYCLE_F64;
        const bool kd = top && kd_records(ctx);
        auto pre = [&] {
            if constexpr (std::is_same<T, float>::value) {
                if (kd) {
                    const R64 rio{ws.r, nullptr, ws.r2, 0.0, ws.p0, nullptr, 0.0, 0.0};
                    wide ? launch_pre<T, VW, kRUpdateK>(L, C, L.z[0], ws.st, s, rio, ws.parts, ws.nparts)
                         : launch_pre<T, 1, kRUpdateK>(L, C, L.z[0], ws.st, s, rio, ws.parts, ws.nparts);
                    return;
                }
                if (top) {
                    const R64 rio{ws.r, ws.ap, ws.r2, 0.0};
                    wide ? launch_pre<T, VW, kRUpdate>(L, C, L.z[0], ws.st, s, rio, ws.parts, ws.nparts)
                         : launch_pre<T, 1, kRUpdate>(L, C, L.z[0], ws.st, s, rio, ws.parts, ws.nparts);
                    return;
                }
            }
            wide ? launch_pre<T, VW>(L, C, L.z[0], ws.st, s) : launch_pre<T, 1>(L, C, L.z[0], ws.st, s);
        };
        auto post = [&] {
            if constexpr (std::is_same<T, float>::value) {
                if (kd) {
                    wide ? launch_post<T, VW, kR64K>(L, L.z[0], C.z[1], C.w, L.z[1], ws.r, ws.parts,
                                                     ws.nparts, ws.st, s)
                         : launch_post<T, 1, kR64K>(L, L.z[0], C.z[1], C.w, L.z[1], ws.r, ws.parts,
                                                    ws.nparts, ws.st, s);
                    return;
                }
                if (top) {
                    wide ? launch_post<T, VW, kR64>(L, L.z[0], C.z[1], C.w, L.z[1], ws.r, ws.parts,
                                                    ws.nparts, ws.st, s)
                         : launch_post<T, 1, kR64>(L, L.z[0], C.z[1], C.w, L.z[1], ws.r, ws.parts,
                                                   ws.nparts, ws.st, s);
                    return;
                }
            }
            wide ? launch_post<T, VW>(L, L.z[0], C.z[1], C.w, L.z[1], nullptr, nullptr, 0, ws.st, s)
                 : launch_post<T, 1>(L, L.z[0], C.z[1], C.w, L.z[1], nullptr, nullptr, 0, ws.st, s);
        };
        L.th_pre = best_th(ctx, L.h, reps, [&](int th) { L.th_pre = th; }, pre);
        L.th_post = best_th(ctx, L.h, reps, [&](int th) { L.th_post = th; }, post);
    }
}

// Tunes the plan of the workspace's shape for the solve of (b, kd, dt) and keeps it in the context.
static void tune(dtd_ctx *ctx, const double *b, const double *kd, double dx, double dt, int wrap_x,
                 int reps, Plan &plan) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    apply_plan(ws, Plan{});
    const double ms_rule = time_solves(ctx, b, kd, dx, dt, wrap_x, reps);
    const bool fused = ctx->opts.nu == 2 && fused_enabled() && ws.lf.size() + ws.ld.size() > 1 &&
                       ctx->opts.transfer == DTD_TRANSFER_AGGREGATE;
    // Kernel by kernel on the resident data, with a state that never stops.
    k_time_state<<<1, 64, 0, s>>>(ws.st, kd);
    LAUNCH_CHECK();
    const int kreps = std::max(3, reps);
    if (fused) {
        if (ws.prec == DTD_CYCLE_F64)
            tune_levels<double>(ctx, ws.ld, kreps, kTuneBottom[0]);
        else
            tune_levels<float>(ctx, ws.lf, kreps, kTuneBottom[0]);
    }
    const Grid g = grid_of(ws);
    const double s2 = dx * dx;
    const bool f64 = ws.prec == DTD_CYCLE_F64;
    ws.th_dir = best_th(ctx, ws.h, kreps, [&](int th) { ws.th_dir = th; }, [&] {
        f64 ? launch_dir<double>(ws, ws.ld[0].z[1], ws.p1, ws.p0, s, ws.ap)
            : launch_dir<float>(ws, ws.lf[0].z[1], ws.p1, ws.p0, s,
                                kd_records(ctx) ? nullptr : ws.ap);
    });
    ws.th_resid = best_th(ctx, ws.h, kreps, [&](int th) { ws.th_resid = th; },
                          [&] { launch_resid(ws, kd, dt, s2, b, ws.x, s, !fused_update(ws, ctx->opts)); });
    auto fin = [&](int which, int th) {
        const int nb = stream_grid<1, 1>(ws.w, ws.h, th);
        need_parts(nb, ws.nparts);
        if (which == 1)
            k_finish1<<<nb, kBlock, 0, s>>>(g, kd, dt, s2, b, ws.x, ws.p0, ws.p1, ws.ap, th, ws.parts,
                                            ws.st);
        else
            k_finish2<<<nb, kBlock, 0, s>>>(g, kd, dt, s2, b, ws.ap, ws.x, th, ws.parts, ws.st);
        LAUNCH_CHECK();
    };
    ws.th_fin1 = best_th(ctx, ws.h, kreps, [&](int th) { ws.th_fin1 = th; }, [&] { fin(1, ws.th_fin1); });
    ws.th_fin2 = best_th(ctx, ws.h, kreps, [&](int th) { ws.th_fin2 = th; }, [&] { fin(2, ws.th_fin2); });
    // The bottom's size, by whole solves with the band heights found.
    double t_best = 0.0;
    int b_best = -1;
    if (fused) {
        for (int bc : kTuneBottom) {
            ws.bottom = bc;
            const double t = time_solves(ctx, b, kd, dx, dt, wrap_x, reps);
            if (b_best < 0 || t < t_best) b_best = bc, t_best = t;
        }
        ws.bottom = b_best;
    } else {
        t_best = time_solves(ctx, b, kd, dx, dt, wrap_x, reps);
    }
    // Keep the plan only if the whole solve is faster than with the built-in rules.
    Plan p;
    for (size_t l = 0; l < std::max(ws.lf.size(), ws.ld.size()); l++)
        p.lvl.push_back(ws.prec == DTD_CYCLE_F64 ? std::array<int, 2>{ws.ld[l].th_pre, ws.ld[l].th_post}
                                                 : std::array<int, 2>{ws.lf[l].th_pre, ws.lf[l].th_post});
    p.th_dir = ws.th_dir, p.th_resid = ws.th_resid, p.th_fin1 = ws.th_fin1, p.th_fin2 = ws.th_fin2;
    p.bottom = ws.bottom;
    const double ms_plan = time_solves(ctx, b, kd, dx, dt, wrap_x, reps);
    plan = ms_plan < ms_rule ? p : Plan{};
    plan.ms_rule = ms_rule;
    plan.ms_tuned = std::min(ms_plan, ms_rule);
    apply_plan(ws, plan);
    for (auto &gr : ws.graph)
        if (gr) (void)hipGraphExecDestroy(gr), gr = nullptr;
    upload_levels(ws, s);
    HIPCHK(hipStreamSynchronize(s));
}

// Calibration kernels of dtd_time_kernel: a float4 copy, and five float4 input streams combined
// into one output (the smoother's stream count without its stencil).
__global__ void k_copy4(long n4, const float4 *__restrict__ a, float4 *__restrict__ o) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k < n4) o[k] = a[k];
}
__global__ void k_stream5(long n4, const float4 *__restrict__ a, const float4 *__restrict__ b,
                          const float4 *__restrict__ c, const float4 *__restrict__ d,
                          const float4 *__restrict__ e, float4 *__restrict__ o) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n4) return;
    const float4 p = a[k], q = b[k], r = c[k], t = d[k], u = e[k];
    o[k] = make_float4(p.x + q.x * (r.x - t.x * u.x), p.y + q.y * (r.y - t.y * u.y),
                       p.z + q.z * (r.z - t.z * u.z), p.w + q.w * (r.w - t.w * u.w));
}

// The surface stage's diffusivity before its solve (surface/mod.rs:1982-1992): per cell,
// `submarine_kd` (marine.rs:44-59) of the water depth sea - h, with the land value
// kd_land * diffu[class] ashore.  The device's tanh and exp are not glibc's: values agree to
// ~1e-14 relative where they are used (an ulp of tanh near -1 grows in 1 + tanh), exactly
// elsewhere.
__global__ __launch_bounds__(kBlock) void k_fill_kd(long n, const double *__restrict__ h,
                                                    const unsigned char *__restrict__ cls,
                                                    const double *__restrict__ diffu,
                                                    double kd_land, double sea, double d0,
                                                    double wave_base, double tidal_range,
                                                    double *__restrict__ kd) {
    const long c = (long)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const double z = sea - h[c];
    const double land = kd_land * diffu[cls[c]];
    double f1;
    if (tidal_range > 0.0)
        f1 = (tanh(z / tidal_range) + 1.0) / 2.0;
    else
        f1 = z >= 0.0 ? 1.0 : 0.0;
    const double f2 = (z > wave_base && wave_base > 0.0) ? exp(-(z - wave_base) / wave_base) : 1.0;
    kd[c] = d0 * f1 * f2 + land * (1.0 - f1);
}

}  // namespace dtd

#include "batch.hpp"
#include "jacobi.hpp"
#include "failure.hpp"
#include "cube.hpp"
#include "nn.hpp"

using namespace dtd;

extern "C" {

void dtd_default_opts(dtd_opts *o) {
    if (!o) return;
    o->device = 0;
    o->omega = 0.8f;
    o->nu = 2;
    o->scale = 0.5f;
    o->max_iter = 400;
    o->cycle_precision = DTD_CYCLE_F32;
    o->guess = DTD_GUESS_LOD;
    o->transfer = DTD_TRANSFER_AGGREGATE;
    o->coarsest = DTD_COARSEST_SWEEPS;
    o->sums = DTD_SUMS_BLOCKS;
}

int dtd_create(const dtd_opts *opts, dtd_ctx **out) {
    if (!out) return DTD_EINVAL;
    *out = nullptr;
    dtd_ctx *ctx = new dtd_ctx();
    if (opts)
        ctx->opts = *opts;
    else
        dtd_default_opts(&ctx->opts);
    if (ctx->opts.nu < 1 || ctx->opts.max_iter < 1 ||
        (ctx->opts.cycle_precision != DTD_CYCLE_F32 && ctx->opts.cycle_precision != DTD_CYCLE_F64 &&
         ctx->opts.cycle_precision != DTD_CYCLE_BF16) ||
        (ctx->opts.guess != DTD_GUESS_LOD && ctx->opts.guess != DTD_GUESS_RHS) ||
        (ctx->opts.transfer != DTD_TRANSFER_AGGREGATE &&
         ctx->opts.transfer != DTD_TRANSFER_BILINEAR) ||
        (ctx->opts.coarsest != DTD_COARSEST_SWEEPS && ctx->opts.coarsest != DTD_COARSEST_EXACT) ||
        (ctx->opts.transfer == DTD_TRANSFER_BILINEAR && ctx->opts.cycle_precision == DTD_CYCLE_BF16) ||
        (ctx->opts.sums != DTD_SUMS_BLOCKS && ctx->opts.sums != DTD_SUMS_STRIPS) ||
        (ctx->opts.sums == DTD_SUMS_STRIPS &&
         (ctx->opts.cycle_precision != DTD_CYCLE_F32 ||
          ctx->opts.transfer != DTD_TRANSFER_AGGREGATE ||
          ctx->opts.coarsest != DTD_COARSEST_SWEEPS))) {
        delete ctx;
        return DTD_EINVAL;
    }
    const char *ng = getenv("DTD_NO_GRAPH");
    ctx->use_graph = !(ng && *ng && *ng != '0');
    const char *nk = getenv("DTD_NO_KD");
    ctx->use_kd = !(nk && *nk && *nk != '0');
    const char *nl = getenv("DTD_NO_LIVE");
    ctx->use_live = !(nl && *nl && *nl != '0');
    // The reference's switch for its audit (D:127-130): set at all, it is on.
    ctx->audit_on = getenv("DEEPTIME_DIFFUSION_AUDIT") != nullptr;
    int rc = guard(ctx, [&] {
        HIPCHK(hipSetDevice(ctx->opts.device));
        scratch_init(ctx->opts.device);
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        for (auto &e : ctx->ev) HIPCHK(hipEventCreate(&e));
        for (auto &e : ctx->flag_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPCHK(hipHostMalloc((void **)&ctx->host_st, sizeof(CgState), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&ctx->host_flag, 2 * sizeof(int),
                             hipHostMallocMapped | hipHostMallocCoherent));
        HIPCHK(hipHostGetDevicePointer((void **)&ctx->dev_flag, ctx->host_flag, 0));
        HIPCHK(hipHostMalloc((void **)&ctx->host_pair, 2 * sizeof(double), hipHostMallocDefault));
    });
    if (rc != DTD_OK) {
        dtd_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return DTD_OK;
}

void dtd_destroy(dtd_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->opts.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_ws(ctx->ws);
    free_batch(ctx);
    free_fail(ctx);
    if (ctx->audit_buf) (void)hipFree(ctx->audit_buf);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->flag_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->host_st) (void)hipHostFree(ctx->host_st);
    if (ctx->host_flag) (void)hipHostFree(ctx->host_flag);
    if (ctx->host_pair) (void)hipHostFree(ctx->host_pair);
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *dtd_last_error(const dtd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// ---- the opt-in audit (`diffusion::audit`, D:118-199) ------------------------------------------

// Per block, over its cells (grid-stride), for the pairs (a, b), (b, c), (a, c): the largest
// |difference| and the sum of squared differences, into parts[6 * block + 2 * pair + {0, 1}].
__global__ __launch_bounds__(kBlock) void k_audit_diff(long n, const double *__restrict__ a,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ c,
                                                       double *__restrict__ parts) {
    __shared__ double lds[kBlock / 64];
    double mx[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        const double d[3] = {fabs(a[i] - b[i]), fabs(b[i] - c[i]), fabs(a[i] - c[i])};
#pragma unroll
        for (int k = 0; k < 3; k++) mx[k] = fmax(mx[k], d[k]), sq[k] += d[k] * d[k];
    }
    for (int k = 0; k < 3; k++) {
        double m = mx[k];
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) lds[wid] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 1; q < kBlock / 64; q++) m = fmax(m, lds[q]);
            parts[6 * blockIdx.x + 2 * k] = m;
        }
        __syncthreads();
        const double t = block_sum<kBlock>(sq[k], lds);
        if (threadIdx.x == 0) parts[6 * blockIdx.x + 2 * k + 1] = t;
    }
}

constexpr double kAuditTightM = 1.0e-9;  // `audit::TIGHT_M` (D:124): the tight solve's RMS residual
constexpr int kAuditBlocks = 1024;

// After a multigrid solve of (b, kd) whose answer is in ws.x: the Jacobi solve and a tight
// multigrid solve of the same step (on the same workspace), the three differences, added to the
// context's sums (`audit::record`, D:148-176).  Leaves the audited answer in ws.x again.
static void audit_record(dtd_ctx *ctx, const double *b, const double *kd, double dx, double dt,
                         int wrap_x, size_t its) {
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)ws.n;
    if (ctx->audit_n < n) {
        if (ctx->audit_buf) HIPCHK(hipFree(ctx->audit_buf));
        ctx->audit_buf = nullptr, ctx->audit_n = 0;
        HIPCHK(hipMalloc((void **)&ctx->audit_buf, (2 * n + 6 * kAuditBlocks) * sizeof(double)));
        ctx->audit_n = n;
    }
    double *nw = ctx->audit_buf, *old = nw + ctx->audit_n, *parts = old + ctx->audit_n;
    HIPCHK(hipMemcpyAsync(nw, ws.x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    const size_t its_old = jacobi_solve(ctx, b, kd, dx, dt, wrap_x, nullptr);
    HIPCHK(hipMemcpyAsync(old, ws.x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    solve_core(ctx, b, kd, dx, dt, wrap_x, kAuditTightM, 0.0, true, kGuessLod, nullptr, nullptr,
               nullptr);
    const int nb = (int)std::min<size_t>(kAuditBlocks, (n + kBlock - 1) / kBlock);
    k_audit_diff<<<nb, kBlock, 0, s>>>((long)n, nw, old, ws.x, parts);
    LAUNCH_CHECK();
    std::vector<double> hp(6 * (size_t)nb);
    HIPCHK(hipMemcpyAsync(hp.data(), parts, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ws.x, nw, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    AuditSums &a = ctx->audit;
    a.solves += 1, a.its_new += its, a.its_old += its_old, a.cells += (double)n;
    for (int k = 0; k < 3; k++) {
        double mx = 0.0, sq = 0.0;
        for (int q = 0; q < nb; q++) mx = std::max(mx, hp[6 * q + 2 * k]), sq += hp[6 * q + 2 * k + 1];
        a.max[k] = std::max(a.max[k], mx);
        a.max_rms[k] = std::max(a.max_rms[k], std::sqrt(sq / (double)n));
        a.sq[k] += sq;
    }
}

// A number as Rust's `{:.3e}` writes it: three decimals, the exponent without sign padding.
static std::string rust_e3(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.3e", v);
    std::string t(b);
    const size_t e = t.find('e');
    if (e == std::string::npos) return t;
    const int ex = std::atoi(t.c_str() + e + 1);
    return t.substr(0, e) + "e" + std::to_string(ex);
}

int dtd_audit(dtd_ctx *ctx, int on, int reset) {
    if (!ctx) return DTD_EINVAL;
    ctx->audit_on = on != 0;
    if (reset) ctx->audit = AuditSums{};
    return DTD_OK;
}

int dtd_audit_line(const dtd_ctx *ctx, char *buf, size_t cap) {
    if (!ctx || !buf || cap == 0) return DTD_EINVAL;
    const AuditSums &s = ctx->audit;
    buf[0] = 0;
    if (s.solves == 0) return DTD_ESTATE;
    const double cells = std::max(s.cells, 1.0);
    auto rms = [&](int k) { return std::sqrt(s.sq[k] / cells); };
    auto per = [&](unsigned long long v) { return (double)v / (double)s.solves; };
    char its[64];
    std::snprintf(its, sizeof its, "its_multigrid=%.1f its_jacobi=%.1f", per(s.its_new),
                  per(s.its_old));
    std::string line = "diffusion_audit solves=" + std::to_string(s.solves) + " " + its;
    const char *names[3] = {"new_vs_old", "old_err", "new_err"};
    for (int k = 0; k < 3; k++) {
        line += std::string(" ") + names[k] + "_max_m=" + rust_e3(s.max[k]);
        line += std::string(" ") + names[k] + "_rms_m=" + rust_e3(rms(k));
        line += std::string(" ") + names[k] + "_worst_rms_m=" + rust_e3(s.max_rms[k]);
    }
    if (line.size() + 1 > cap) return DTD_EINVAL;
    std::memcpy(buf, line.c_str(), line.size() + 1);
    return DTD_OK;
}

int dtd_solve(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_inout,
              const double *kd, double dt, int solver, double tol_m, size_t *iterations) {
    return guard(ctx, [&] {
        begin(ctx);
        if (iterations) *iterations = 0;
        if (solver != DTD_SOLVER_MULTIGRID && solver != DTD_SOLVER_JACOBI)
            throw Error{DTD_EINVAL, "unknown solver"};
        // The Jacobi solve keeps its own tolerance (D:72-74), as the reference's does.
        if (solver == DTD_SOLVER_MULTIGRID && !(std::isfinite(tol_m) && tol_m > 0.0))
            throw Error{DTD_EINVAL, "tol_m must be finite and positive"};
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!h_inout || !kd) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        Workspace &ws = ctx->ws;
        upload(ctx, ws.kd, kd, n * sizeof(double));
        upload(ctx, ws.b, h_inout, n * sizeof(double));
        const size_t its =
            solver == DTD_SOLVER_JACOBI
                ? jacobi_solve(ctx, ws.b, ws.kd, dx, dt, wrap_x, nullptr)
                : solve_core(ctx, ws.b, ws.kd, dx, dt, wrap_x, tol_m, 0.0, true, kGuessLod,
                             nullptr, nullptr, nullptr);
        if (solver == DTD_SOLVER_MULTIGRID && ctx->audit_on)
            audit_record(ctx, ws.b, ws.kd, dx, dt, wrap_x, its);
        // The answer comes back through pinned staging and reaches h_inout only once every
        // device call has succeeded, so h_inout keeps h0 whatever fails.
        if (ctx->stage_n < n) {
            if (ctx->stage) (void)hipHostFree(ctx->stage);
            ctx->stage = nullptr, ctx->stage_n = 0;
            HIPCHK(hipHostMalloc((void **)&ctx->stage, n * sizeof(double), hipHostMallocDefault));
            ctx->stage_n = n;
        }
        download(ctx, ctx->stage, ws.x, n * sizeof(double));
        std::memcpy(h_inout, ctx->stage, n * sizeof(double));
        if (iterations) *iterations = its;
    });
}

int dtd_solve_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_dev,
                     const double *kd_dev, double dt, double tol_m, dtd_stats *stats) {
    return guard(ctx, [&] {
        begin(ctx);
        if (stats) *stats = dtd_stats{};
        if (!(std::isfinite(tol_m) && tol_m > 0.0))
            throw Error{DTD_EINVAL, "tol_m must be finite and positive"};
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!h_dev || !kd_dev) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        double rel = 0.0;
        const size_t its = solve_core(ctx, h_dev, kd_dev, dx, dt, wrap_x, tol_m, 0.0, true,
                                      kGuessLod, stats, &rel, nullptr);
        if (stats) stats->rel_residual = rel;
        if (ctx->audit_on) audit_record(ctx, h_dev, kd_dev, dx, dt, wrap_x, its);
        copy_answer(ctx, h_dev);
    });
}

int dtd_solve_device_noalloc(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                             double *h_dev, const double *kd_dev, double dt, double tol_m,
                             dtd_stats *stats) {
    return guard(ctx, [&] {
        begin(ctx);
        if (stats) *stats = dtd_stats{};
        if (!(std::isfinite(tol_m) && tol_m > 0.0))
            throw Error{DTD_EINVAL, "tol_m must be finite and positive"};
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!h_dev || !kd_dev) throw Error{DTD_EINVAL, "null array"};
        if (ctx->opts.sums == DTD_SUMS_STRIPS)
            throw Error{DTD_EINVAL, "the reference's summation order needs the allocating entry points"};
        const int wrap = wrap_x && w > 1;
        // The size check first, by a carve that only counts.
        Workspace probe;
        Carver count{nullptr, kScratchBytes};
        lay_out(probe, ctx->opts, (int)w, (int)h, wrap, true, std::ref(count));
        if (count.used > kScratchBytes)
            throw Error{DTD_ETOOBIG, "the grid's workspace (" + std::to_string(count.used) +
                                         " bytes) exceeds the fixed scratch array (" +
                                         std::to_string(kScratchBytes) + " bytes)"};
        unsigned char *base = (unsigned char *)scratch_of(ctx->opts.device);
        if (!base) throw Error{DTD_ESTATE, "the scratch array was not set up for this device"};
        Workspace sw;
        Carver carve{base, kScratchBytes};
        lay_out(sw, ctx->opts, (int)w, (int)h, wrap, true, std::ref(carve));
        auto it = ctx->plans.find(PlanKey{(int)w, (int)h, wrap, ctx->opts.cycle_precision});
        if (it != ctx->plans.end()) apply_plan(sw, it->second);  // launch shapes as dtd_solve_device's
        ScratchSwap swap(ctx, std::move(sw));
        upload_levels(ctx->ws, ctx->stream);
        double rel = 0.0;
        solve_core(ctx, h_dev, kd_dev, dx, dt, wrap_x, tol_m, 0.0, true, kGuessLod, stats, &rel,
                   nullptr);
        if (stats) stats->rel_residual = rel;
        copy_answer(ctx, h_dev);
    });
}

int dtd_solve_rel_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_dev,
                         const double *kd_dev, double dt, double rel_tol, dtd_stats *stats) {
    return guard(ctx, [&] {
        begin(ctx);
        dtd_stats local{};
        dtd_stats *st = stats ? stats : &local;
        *st = dtd_stats{};
        if (!(std::isfinite(rel_tol) && rel_tol > 0.0))
            throw Error{DTD_EINVAL, "rel_tol must be finite and positive"};
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!h_dev || !kd_dev) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        // The first solve stops at r.r <= rel^2 b.b (tol_m = rel ||b|| / sqrt(n)), b.b formed on
        // the device by the first residual's pass.
        double rel = 0.0, bnorm = 0.0;
        solve_core(ctx, h_dev, kd_dev, dx, dt, wrap_x, 0.0, rel_tol * rel_tol, true, kGuessLod,
                   st, &rel, &bnorm);
        double tol_m = rel_tol * bnorm / std::sqrt((double)n);
        while (rel > rel_tol && st->restarts < 4 && bnorm > 0.0) {
            // Solve again from the last answer (setup kept), aiming below the target by the
            // factor the last solve missed it by.
            tol_m *= 0.5 * rel_tol / rel;
            st->restarts++;
            solve_core(ctx, h_dev, kd_dev, dx, dt, wrap_x, tol_m, 0.0, false, kGuessKeep, st,
                       &rel, &bnorm);
        }
        st->rel_residual = rel;
        copy_answer(ctx, h_dev);
    });
}

int dtd_fill_kd(dtd_ctx *ctx, size_t n, const double *h_dev, const unsigned char *class_dev,
                const double *diffu_dev, double kd_land, double sea, double d0, double wave_base,
                double tidal_range, double *kd_dev) {
    return guard(ctx, [&] {
        begin(ctx);
        if (n == 0) return;
        if (!h_dev || !class_dev || !diffu_dev || !kd_dev) throw Error{DTD_EINVAL, "null array"};
        if (n > (size_t)INT32_MAX * kBlock) throw Error{DTD_ETOOBIG, "too many cells"};
        k_fill_kd<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(
            (long)n, h_dev, class_dev, diffu_dev, kd_land, sea, d0, wave_base, tidal_range, kd_dev);
        LAUNCH_CHECK();
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

static void check_failure(size_t w, size_t h, const dtd_failure *p, const void *sea,
                          const void *hgt, const void *avail) {
    check_grid(w, h);
    if (!p) throw Error{DTD_EINVAL, "null parameters"};
    if (w * h && (!sea || !hgt || !avail)) throw Error{DTD_EINVAL, "null array"};
    if (!(p->rate >= 0.0) || !std::isfinite(p->rate) || !std::isfinite(p->critical_slope))
        throw Error{DTD_EINVAL, "rate and critical slope must be finite, rate >= 0"};
}

int dtd_slope_failure_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                             const unsigned char *sea_dev, double sea_level, const dtd_failure *p,
                             double *h_dev, double *avail_dev, double *moved,
                             unsigned *iterations) {
    return guard(ctx, [&] {
        begin(ctx);
        if (moved) *moved = 0.0;
        if (iterations) *iterations = 0;
        check_failure(w, h, p, sea_dev, h_dev, avail_dev);
        if (w * h == 0 || p->iterations == 0) return;
        FailWs &f = fit_fail(ctx, w, h, false);
        const double m = slope_failure(ctx, f, w, h, dx, wrap_x, sea_dev, sea_level,
                                       p->critical_slope, p->rate, p->iterations, h_dev,
                                       avail_dev, iterations);
        if (moved) *moved = m;
    });
}

int dtd_cube_new(dtd_ctx *ctx, size_t n, const uint32_t *slots, const double *lap,
                 const double *area, double mean_area, dtd_cube **out) {
    return guard(ctx, [&] {
        begin(ctx);
        if (!out) throw Error{DTD_EINVAL, "null output"};
        *out = nullptr;
        if (n < 2) throw Error{DTD_EINVAL, "a cube face needs at least 2 x 2 cells"};
        if (n > 18000) throw Error{DTD_ETOOBIG, "the cube exceeds 2^31 cells"};
        if (!slots || !lap || !area) throw Error{DTD_EINVAL, "null array"};
        if (!(mean_area > 0.0) || !std::isfinite(mean_area))
            throw Error{DTD_EINVAL, "mean_area must be finite and > 0"};
        const size_t n6 = 6 * n * n;
        for (size_t c = 0; c < n6; c++)
            for (int q = 0; q < 8; q++) {
                const uint32_t m = slots[8 * c + q];
                if (m != DTD_NO_CELL && m >= n6)
                    throw Error{DTD_EINVAL, "a neighbour slot is outside the grid"};
                // Every cell has its four side neighbours; only a diagonal across a cube corner
                // goes missing (sphere/cube.rs:12-17).
                if (q % 2 == 0 && m == DTD_NO_CELL)
                    throw Error{DTD_EINVAL, "a side neighbour is missing"};
            }
        *out = cube_new(ctx, (int)n, slots, lap, area, mean_area);
    });
}

void dtd_cube_free(dtd_cube *g) { free_cube(g); }

int dtd_cube_solve_device(dtd_ctx *ctx, dtd_cube *g, double *h_dev, const double *kd_dev,
                          double dt, int solver, double tol_m, size_t *iterations) {
    return guard(ctx, [&] {
        begin(ctx);
        if (iterations) *iterations = 0;
        if (!g) throw Error{DTD_EINVAL, "null grid"};
        if (!h_dev || !kd_dev) throw Error{DTD_EINVAL, "null array"};
        if (solver != DTD_SOLVER_MULTIGRID && solver != DTD_SOLVER_JACOBI)
            throw Error{DTD_EINVAL, "unknown solver"};
        if (solver == DTD_SOLVER_MULTIGRID && !(std::isfinite(tol_m) && tol_m > 0.0))
            throw Error{DTD_EINVAL, "tol_m must be finite and positive"};
        const size_t its = cube_solve(ctx, g, h_dev, kd_dev, dt, solver, tol_m);
        if (iterations) *iterations = its;
    });
}

int dtd_cube_solve(dtd_ctx *ctx, dtd_cube *g, double *h_inout, const double *kd, double dt,
                   int solver, double tol_m, size_t *iterations) {
    return guard(ctx, [&] {
        begin(ctx);
        if (iterations) *iterations = 0;
        if (!g) throw Error{DTD_EINVAL, "null grid"};
        if (!h_inout || !kd) throw Error{DTD_EINVAL, "null array"};
        if (solver != DTD_SOLVER_MULTIGRID && solver != DTD_SOLVER_JACOBI)
            throw Error{DTD_EINVAL, "unknown solver"};
        if (solver == DTD_SOLVER_MULTIGRID && !(std::isfinite(tol_m) && tol_m > 0.0))
            throw Error{DTD_EINVAL, "tol_m must be finite and positive"};
        const size_t bytes = (size_t)g->n6 * sizeof(double);
        if (!g->hd) {
            HIPCHK(hipMalloc(&g->hd, bytes));
            g->owned.push_back(g->hd);
            HIPCHK(hipMalloc(&g->kdd, bytes));
            g->owned.push_back(g->kdd);
        }
        upload(ctx, g->hd, h_inout, bytes);
        upload(ctx, g->kdd, kd, bytes);
        const size_t its = cube_solve(ctx, g, g->hd, g->kdd, dt, solver, tol_m);
        download(ctx, h_inout, g->hd, bytes);
        if (iterations) *iterations = its;
    });
}

int dtd_slope_failure(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                      const unsigned char *sea, double sea_level, const dtd_failure *p,
                      double *h_inout, double *avail_inout, double *moved, unsigned *iterations) {
    return guard(ctx, [&] {
        begin(ctx);
        if (moved) *moved = 0.0;
        if (iterations) *iterations = 0;
        check_failure(w, h, p, sea, h_inout, avail_inout);
        const size_t n = w * h;
        if (n == 0 || p->iterations == 0) return;
        FailWs &f = fit_fail(ctx, w, h, true);
        upload(ctx, f.sd, sea, n);
        upload(ctx, f.hd, h_inout, n * sizeof(double));
        upload(ctx, f.ad, avail_inout, n * sizeof(double));
        const double m = slope_failure(ctx, f, w, h, dx, wrap_x, f.sd, sea_level,
                                       p->critical_slope, p->rate, p->iterations, f.hd, f.ad,
                                       iterations);
        download(ctx, h_inout, f.hd, n * sizeof(double));
        download(ctx, avail_inout, f.ad, n * sizeof(double));
        if (moved) *moved = m;
    });
}

int dtd_step_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_dev,
                    const double *kd_dev, double dt, dtd_stats *stats) {
    return guard(ctx, [&] {
        begin(ctx);
        if (stats) *stats = dtd_stats{};
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!h_dev || !kd_dev) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        jacobi_solve(ctx, h_dev, kd_dev, dx, dt, wrap_x, stats);
        copy_answer(ctx, h_dev);
    });
}

int dtd_solve_batch(dtd_ctx *ctx, size_t n, const dtd_grid *grids, double rel_tol,
                    size_t *iterations, double *rel_residuals, dtd_batch_stats *stats) {
    return guard(ctx, [&] {
        begin(ctx);
        if (stats) *stats = dtd_batch_stats{};
        if (n == 0) return;
        if (!grids) throw Error{DTD_EINVAL, "null grid array"};
        solve_batch(ctx, n, grids, rel_tol, iterations, rel_residuals, stats);
    });
}

int dtd_tune_batch(dtd_ctx *ctx, size_t n, const dtd_grid *grids, double rel_tol, int reps,
                   double *ms_rule, double *ms_tuned) {
    return guard(ctx, [&] {
        begin(ctx);
        if (ms_rule) *ms_rule = 0.0;
        if (ms_tuned) *ms_tuned = 0.0;
        if (n == 0) return;
        if (!grids) throw Error{DTD_EINVAL, "null grid array"};
        if (reps < 1) throw Error{DTD_EINVAL, "reps must be positive"};
        tune_batch(ctx, n, grids, rel_tol, std::min(reps, 20), ms_rule, ms_tuned);
    });
}

int dtd_operator(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *kd,
                 double dt, double *east, double *south, double *diag) {
    return guard(ctx, [&] {
        begin(ctx);
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!kd || !east || !south || !diag) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        Workspace &ws = ctx->ws;
        upload(ctx, ws.kd, kd, n * sizeof(double));
        const Grid g = grid_of(ws);
        k_operator<<<nblocks(n), kBlock, 0, ctx->stream>>>(g, ws.kd, dt, dx * dx, ws.es, nullptr,
                                                           nullptr);
        LAUNCH_CHECK();
        k_unpack_es<<<nblocks(n), kBlock, 0, ctx->stream>>>(g, ws.es, ws.p0, ws.p1, ws.ap);
        LAUNCH_CHECK();
        download(ctx, east, ws.p0, n * sizeof(double));
        download(ctx, south, ws.p1, n * sizeof(double));
        download(ctx, diag, ws.ap, n * sizeof(double));
    });
}

int dtd_apply(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *kd,
              double dt, const double *x, double *y) {
    return guard(ctx, [&] {
        begin(ctx);
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!kd || !x || !y) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        Workspace &ws = ctx->ws;
        upload(ctx, ws.kd, kd, n * sizeof(double));
        upload(ctx, ws.x, x, n * sizeof(double));
        const Grid g = grid_of(ws);
        k_operator<<<nblocks(n), kBlock, 0, ctx->stream>>>(g, ws.kd, dt, dx * dx, ws.es, nullptr,
                                                           nullptr);
        LAUNCH_CHECK();
        k_apply<<<nblocks(n), kBlock, 0, ctx->stream>>>(g, ws.es, ws.x, ws.ap);
        LAUNCH_CHECK();
        download(ctx, y, ws.ap, n * sizeof(double));
    });
}

int dtd_lod(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_inout,
            const double *kd, double dt) {
    return guard(ctx, [&] {
        begin(ctx);
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!kd || !h_inout) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        Workspace &ws = ctx->ws;
        upload(ctx, ws.kd, kd, n * sizeof(double));
        upload(ctx, ws.b, h_inout, n * sizeof(double));
        lod(ctx, ws.b, ws.x, ws.kd, dx, dt, wrap_x);
        download(ctx, h_inout, ws.x, n * sizeof(double));
    });
}

int dtd_time_kernel(dtd_ctx *ctx, int kernel, int reps, double *ms_avg, double *bytes) {
    return guard(ctx, [&] {
        begin(ctx);
        Workspace &ws = ctx->ws;
        if (ws.n == 0 || ws.prec == DTD_CYCLE_F64 || ws.lf.size() < 2)
            throw Error{DTD_ESTATE, "time a kernel after an f32-cycle solve of a multi-level grid"};
        if (reps < 1) throw Error{DTD_EINVAL, "reps must be positive"};
        hipStream_t s = ctx->stream;
        Lvl<float> &L = ws.lf[0];
        Lvl<float> &C = ws.lf[1];
        const int nb = nblocks(ws.n);
        const int nbc = nblocks(ws.n, kCgCells);
        const double n = (double)ws.n;
        double b = 0.0;
        // The workspace's kd buffer holds the last host solve's kd, and nothing defined after
        // solves on device arrays: the passes are timed on K = 1 (every face dt / dx^2 = 1; the
        // values do not change the work, while undefined ones could send the kd-formed faces
        // down their rare exact-division branch).
        if (!ws.kd) throw Error{DTD_ESTATE, "this workspace keeps no kd buffer"};
        k_fill64<<<nblocks(ws.n), kBlock, 0, s>>>(ws.kd, ws.n, 1.0);
        LAUNCH_CHECK();
        // The finest records on the same K (a solve with kd-formed records stores none).
        k_operator<<<nblocks(ws.n), kBlock, 0, s>>>(grid_of(ws), ws.kd, 1.0, 1.0, nullptr, L.coef,
                                                    nullptr);
        LAUNCH_CHECK();
        k_time_state<<<1, 64, 0, s>>>(ws.st, ws.kd);
        LAUNCH_CHECK();
        // The finest level's passes as the solve runs them (kd-formed records: no stored A p, no
        // records read by the fused smoothers).
        const bool kd = kd_records(ctx);
        auto one = [&] {
            switch (kernel) {
            case DTD_KERNEL_SWEEP:  // r, record, z in; z out (f32)
                k_sweep<float><<<nb, kBlock, 0, s>>>(L, L.z[1], L.z[0], nullptr, 0, nullptr,
                                                     nullptr, ws.st);
                b = 24.0 * n;
                break;
            case DTD_KERNEL_DIRECTION:  // z (f32), p_old, x, kd in; p, A p (unless kd), x out
                launch_dir<float>(ws, L.z[1], ws.p1, ws.p0, s, kd ? nullptr : ws.ap);
                b = (4.0 + 8.0 * 3 + 8.0 * (kd ? 2 : 3)) * n;
                break;
            case DTD_KERNEL_UPDATE:  // A p, r in; r, r32 out
                k_update<float><<<nbc, kBlock, 0, s>>>((int)ws.n, ws.ap, ws.r, L.r, ws.parts,
                                                       ws.st);
                b = (8.0 * 3 + 4.0) * n;
                break;
            case DTD_KERNEL_RESTRICT:  // fine r, z, record in; coarse r out
                k_restrict<float><<<nblocks(C.n), kBlock, 0, s>>>(L, L.z[1], C, ws.st);
                b = 20.0 * n + 4.0 * C.n;
                break;
            case DTD_KERNEL_PRE:  // r64, A p (or p and kd), record in; r64, z and the coarse r out
                if (L.w % 2 || !ws.r2) throw Error{DTD_ESTATE, "the finest width is odd"};
                if (kd) {
                    launch_pre<float, 2, kRUpdateK>(
                        L, C, L.z[0], ws.st, s, R64{ws.r, nullptr, ws.r2, 0.0, ws.p0, nullptr, 0.0, 0.0},
                        ws.parts, ws.nparts);
                    b = 36.0 * n + 4.0 * C.n;  // r64, p, kd in; r64, z out
                } else {
                    launch_pre<float, 2, kRUpdate>(L, C, L.z[0], ws.st, s,
                                                   R64{ws.r, ws.ap, ws.r2, 0.0}, ws.parts, ws.nparts);
                    b = (L.coefh ? 36.0 : 40.0) * n + 4.0 * C.n;  // bf16 records: 8 B, not 12
                }
                break;
            case DTD_KERNEL_POST:  // z, r64, record (or kd), coarse z in; z out
                if (L.w % 2) throw Error{DTD_ESTATE, "the finest width is odd"};
                if (kd) {
                    launch_post<float, 2, kR64K>(L, L.z[0], C.z[1], C.w, L.z[1], ws.r, ws.parts,
                                                 ws.nparts, ws.st, s);
                    b = 24.0 * n + 4.0 * C.n;
                } else {
                    launch_post<float, 2, kR64>(L, L.z[0], C.z[1], C.w, L.z[1], ws.r, ws.parts,
                                                ws.nparts, ws.st, s);
                    b = (L.coefh ? 24.0 : 28.0) * n + 4.0 * C.n;
                }
                break;
            case DTD_KERNEL_BANDCOPY: {  // r, record in; z out, the pre-smoother's walk
                const int th = band_height<2>((const void *)k_bandcopy<2>, L.w, L.h);
                k_bandcopy<2><<<stream_grid<2, 1>(L.w, L.h, th), kBlock, 0, s>>>(L, L.z[0], th);
                b = 20.0 * n;
                break;
            }
            case DTD_KERNEL_COPY:  // r32 in, z out
            case DTD_KERNEL_STREAM5: {  // five arrays of n floats in; one out
                const long n4 = ws.n / 4;
                const int nb4 = (int)((n4 + kBlock - 1) / kBlock);
                if (kernel == DTD_KERNEL_COPY) {
                    k_copy4<<<nb4, kBlock, 0, s>>>(n4, (const float4 *)L.r, (float4 *)L.z[0]);
                    b = 8.0 * 4 * n4;
                } else {
                    k_stream5<<<nb4, kBlock, 0, s>>>(n4, (const float4 *)L.r,
                                                     (const float4 *)L.z[1],
                                                     (const float4 *)ws.p0, (const float4 *)ws.p1,
                                                     (const float4 *)ws.ap, (float4 *)L.z[0]);
                    b = 24.0 * 4 * n4;
                }
                break;
            }
            default:
                throw Error{DTD_EINVAL, "unknown kernel"};
            }
            LAUNCH_CHECK();
        };
        one();  // warm
        HIPCHK(hipEventRecord(ctx->ev[4], s));
        for (int k = 0; k < reps; k++) one();
        HIPCHK(hipEventRecord(ctx->ev[5], s));
        HIPCHK(hipEventSynchronize(ctx->ev[5]));
        if (ms_avg) *ms_avg = elapsed(ctx->ev[4], ctx->ev[5]) / reps;
        if (bytes) *bytes = b;
    });
}

int dtd_cycle_new(dtd_ctx *ctx, size_t w, size_t h, int wrap_x, const double *east,
                  const double *south, const double *diag) {
    return guard(ctx, [&] {
        begin(ctx);
        ctx->cycle_levels = 0;
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) throw Error{DTD_EINVAL, "an empty grid has no cycle"};
        if (!east || !south || !diag) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        Workspace &ws = ctx->ws;
        upload(ctx, ws.p0, east, n * sizeof(double));
        upload(ctx, ws.p1, south, n * sizeof(double));
        upload(ctx, ws.ap, diag, n * sizeof(double));
        hipStream_t s = ctx->stream;
        if (ws.prec == DTD_CYCLE_F64) {
            k_finest64<<<nblocks(n), kBlock, 0, s>>>((int)n, ws.p0, ws.p1, ws.ap, ws.ld[0].coef);
            LAUNCH_CHECK();
            build_coarse<double>(ws.ld, (double)ctx->opts.scale, s);
        } else {
            k_finest32<<<nblocks(n), kBlock, 0, s>>>((int)n, ws.p0, ws.p1, ws.ap, ws.lf[0].coef);
            LAUNCH_CHECK();
            build_coarse<float>(ws.lf, ctx->opts.scale, s);
            pack_bf16(ws, s);
        }
        HIPCHK(hipStreamSynchronize(s));
        ctx->cycle_levels = ws.prec == DTD_CYCLE_F64 ? ws.ld.size() : ws.lf.size();
    });
}

size_t dtd_cycle_levels(const dtd_ctx *ctx) { return ctx ? ctx->cycle_levels : 0; }

int dtd_cycle_level(dtd_ctx *ctx, size_t l, size_t *w, size_t *h, int *wrap, float *east,
                    float *south, float *diag, float *wd, float *mass) {
    return guard(ctx, [&] {
        begin(ctx);
        if (l >= ctx->cycle_levels) throw Error{DTD_ESTATE, "no such level"};
        Workspace &ws = ctx->ws;
        if (ws.prec == DTD_CYCLE_F64) throw Error{DTD_ESTATE, "the levels are f64"};
        Lvl<float> &L = ws.lf[l];
        if (w) *w = (size_t)L.w;
        if (h) *h = (size_t)L.h;
        if (wrap) *wrap = L.wrap;
        const size_t n = (size_t)L.n;
        std::vector<float> rec(kRec * n);
        download(ctx, rec.data(), L.coef, kRec * n * sizeof(float));
        float *outs[3] = {east, south, diag};
        for (int q = 0; q < kRec; q++)
            if (outs[q])
                for (size_t c = 0; c < n; c++) outs[q][c] = rec[kRec * c + q];
        // The smoother's weight omega / diag is formed where it is used, in f32 as here.
        if (wd)
            for (size_t c = 0; c < n; c++) {
                const float d = rec[kRec * c + 2];
                wd[c] = d != 0.0f ? L.omega / d : 0.0f;
            }
        if (mass) {
            if (L.mass)
                download(ctx, mass, L.mass, n * sizeof(float));
            else
                for (size_t c = 0; c < n; c++) mass[c] = 1.0f;
        }
    });
}

int dtd_cycle_precondition(dtd_ctx *ctx, const float *r32, const double *r64, float *z32,
                           double *rz) {
    return guard(ctx, [&] {
        begin(ctx);
        if (ctx->cycle_levels == 0) throw Error{DTD_ESTATE, "dtd_cycle_new first"};
        if (!r32 || !r64 || !z32) throw Error{DTD_EINVAL, "null array"};
        Workspace &ws = ctx->ws;
        const size_t n = (size_t)ws.n;
        hipStream_t s = ctx->stream;
        upload(ctx, ws.r, r64, n * sizeof(double));
        int nbz;
        if (ws.prec == DTD_CYCLE_F64) {
            // An f64 cycle reads the f64 residual; r32 is its value rounded, kept for the ABI.
            nbz = vcycle<double>(ws.ld, 0, (int)ctx->opts.nu, ws.r, ws.parts, ws.nparts, nullptr,
                                 s, kRLevel, R64{}, ws.d_ld, ws.bottom);
        } else {
            upload(ctx, ws.lf[0].r, r32, n * sizeof(float));
            nbz = vcycle<float>(ws.lf, 0, (int)ctx->opts.nu, ws.r, ws.parts, ws.nparts, nullptr,
                                s, kRLevel, R64{}, ws.d_lf, ws.bottom);
        }
        k_sum1<<<1, kReduceBlock, 0, s>>>(ws.parts, nbz, ws.pair);
        LAUNCH_CHECK();
        if (ws.prec == DTD_CYCLE_F64) {
            std::vector<double> tmp(n);
            download(ctx, tmp.data(), ws.ld[0].z[1], n * sizeof(double));
            for (size_t c = 0; c < n; c++) z32[c] = (float)tmp[c];
        } else {
            download(ctx, z32, ws.lf[0].z[1], n * sizeof(float));
        }
        download(ctx, ctx->host_pair, ws.pair, sizeof(double));
        if (rz) *rz = ctx->host_pair[0];
    });
}

// ---- the learned first guess (no reference counterpart) ------------------------------------

int dtd_nn_conv3x3(dtd_ctx *ctx, int n, int h, int w, int cin, int cout, const uint16_t *x,
                   const uint16_t *wt, const float *bias, int relu, uint16_t *y, float *y32,
                   int y32c) {
    return guard(ctx, [&] {
        begin(ctx);
        nn::check_shape(n, h, w), nn::check_channels(cin), nn::check_channels(cout);
        if (!x || !wt || (!y && !y32)) throw Error{DTD_EINVAL, "null array"};
        if (y32 && (y32c < 1 || y32c > cout)) throw Error{DTD_EINVAL, "y32c out of range"};
        nn::check_aligned(x, 8), nn::check_aligned(wt, 8);
        nn::conv(ctx->stream, nn::Shape{n, h, w}, cin, cout, x, wt, bias, relu, y, y32, y32c);
    });
}

int dtd_nn_conv3x3_cat(dtd_ctx *ctx, int n, int h, int w, int cu, int cs, int cout,
                       const uint16_t *u, const uint16_t *skip, const uint16_t *wt,
                       const float *bias, int relu, uint16_t *y, float *y32, int y32c) {
    return guard(ctx, [&] {
        begin(ctx);
        nn::check_shape(n, h, w), nn::check_channels(cu), nn::check_channels(cs);
        nn::check_channels(cout);
        if (!u || !skip || !wt || !y) throw Error{DTD_EINVAL, "null array"};
        if (y32 && (y32c < 1 || y32c > cout)) throw Error{DTD_EINVAL, "y32c out of range"};
        if (!nn::conv_cat_ok(nn::Shape{n, h, w}, cu, cs, u, skip, wt))
            throw Error{DTD_EINVAL, "the in-place concatenation needs w a multiple of 128, h even, "
                                    "cu and cs multiples of 32 and 16-byte aligned tensors"};
        nn::conv_cat(ctx->stream, nn::Shape{n, h, w}, cu, cs, cout, u, skip, wt, bias, relu, y,
                     y32, y32c);
    });
}

int dtd_nn_conv3x3_bwd_input(dtd_ctx *ctx, int n, int h, int w, int cin, int cout,
                             const uint16_t *dz, const uint16_t *wt, uint16_t *wt_scratch,
                             uint16_t *dx) {
    return guard(ctx, [&] {
        begin(ctx);
        nn::check_shape(n, h, w), nn::check_channels(cin), nn::check_channels(cout);
        if (!dz || !wt || !wt_scratch || !dx) throw Error{DTD_EINVAL, "null array"};
        nn::check_aligned(dz, 8), nn::check_aligned(wt_scratch, 8);
        const long nw = (long)cin * 9 * cout;
        nn::k_conv_flip<<<nn::blocks(nw), 256, 0, ctx->stream>>>(cin, cout, wt, wt_scratch);
        LAUNCH_CHECK();
        nn::conv(ctx->stream, nn::Shape{n, h, w}, cout, cin, dz, wt_scratch, nullptr, 0, dx,
                nullptr, 0);
    });
}

size_t dtd_nn_bwd_weight_scratch(int n, int h, int w, int cin, int cout, size_t *f32_elems) {
    const long np = (long)n * h * w;
    const long chunks = (np + nn::kChunk - 1) / nn::kChunk;
    if (f32_elems) *f32_elems = (size_t)chunks * cout * 9 * cin;
    return (size_t)(3 * cin + cout) * np;  // x with its columns moved by -1, 0, 1, and dz
}

int dtd_nn_conv3x3_bwd_weight(dtd_ctx *ctx, int n, int h, int w, int cin, int cout,
                              const uint16_t *x, const uint16_t *dz, uint16_t *planes,
                              size_t planes_elems, float *part, size_t part_elems, float *dw,
                              float *db) {
    return guard(ctx, [&] {
        begin(ctx);
        nn::check_shape(n, h, w), nn::check_channels(cin), nn::check_channels(cout);
        if (!x || !dz || !planes || !part || !dw) throw Error{DTD_EINVAL, "null array"};
        nn::check_aligned(planes, 16);
        size_t need_part = 0;
        const size_t need_planes = dtd_nn_bwd_weight_scratch(n, h, w, cin, cout, &need_part);
        if (planes_elems < need_planes || part_elems < need_part)
            throw Error{DTD_EINVAL, "scratch too small"};
        hipStream_t s = ctx->stream;
        const long np = (long)n * h * w;
        const nn::Shape sh{n, h, w};
        uint16_t *xs = planes, *dzt = planes + (size_t)3 * cin * np;
        const dim3 tg((unsigned)((np + 63) / 64), (unsigned)(cin / 16));
        for (int dx = -1; dx <= 1; dx++) {
            nn::k_to_planes<<<tg, 256, 0, s>>>(sh, cin, dx, x, xs + (size_t)(dx + 1) * cin * np);
            LAUNCH_CHECK();
        }
        nn::k_to_planes<<<dim3((unsigned)((np + 63) / 64), (unsigned)(cout / 16)), 256, 0, s>>>(
            sh, cout, 0, dz, dzt);
        LAUNCH_CHECK();
        const int chunks = (int)((np + nn::kChunk - 1) / nn::kChunk);
        const int combos = (cout / 16) * (cin / 16) * 9;
        const dim3 grid((unsigned)chunks, (unsigned)((combos + 3) / 4));
        if (np % 32 == 0)
            nn::k_conv_dw<32><<<grid, 256, 0, s>>>(sh, cin, cout, xs, dzt, nn::kChunk, part);
        else
            nn::k_conv_dw<16><<<grid, 256, 0, s>>>(sh, cin, cout, xs, dzt, nn::kChunk, part);
        LAUNCH_CHECK();
        const long nw = (long)cout * 9 * cin;
        nn::k_dw_reduce<<<nn::blocks(nw), 256, 0, s>>>(nw, chunks, part, dw);
        LAUNCH_CHECK();
        if (db) {
            nn::k_db<<<(unsigned)cout, 256, 0, s>>>(np, dzt, db);
            LAUNCH_CHECK();
        }
    });
}

int dtd_nn_relu_bwd(dtd_ctx *ctx, size_t n, const uint16_t *y, const uint16_t *dy, uint16_t *dz) {
    return guard(ctx, [&] {
        begin(ctx);
        if (!y || !dy || !dz) throw Error{DTD_EINVAL, "null array"};
        nn::check_aligned(y, 16), nn::check_aligned(dy, 16), nn::check_aligned(dz, 16);
        if (n == 0) return;
        nn::k_relu_bwd<<<nn::blocks(((long)n + 7) / 8), 256, 0, ctx->stream>>>((long)n, y, dy,
                                                                                dz);
        LAUNCH_CHECK();
    });
}

int dtd_nn_pool(dtd_ctx *ctx, int n, int h, int w, int c, const uint16_t *x, uint16_t *y) {
    return guard(ctx, [&] {
        begin(ctx);
        nn::check_shape(n, h, w), nn::check_channels(c);
        if (h % 2 || w % 2) throw Error{DTD_EINVAL, "pooling needs even sizes"};
        if (!x || !y) throw Error{DTD_EINVAL, "null array"};
        nn::check_aligned(x, 16), nn::check_aligned(y, 16);
        nn::k_pool<<<nn::blocks((long)n * (h / 2) * (w / 2) * (c / 8)), 256, 0, ctx->stream>>>(
            nn::Shape{n, h, w}, c, x, y);
        LAUNCH_CHECK();
    });
}

int dtd_nn_up_cat(dtd_ctx *ctx, int n, int h, int w, int cu, int cs, const uint16_t *u,
                  const uint16_t *skip, uint16_t *cat) {
    return guard(ctx, [&] {
        begin(ctx);
        nn::check_shape(n, h, w), nn::check_channels(cu), nn::check_channels(cs);
        if (h % 2 || w % 2) throw Error{DTD_EINVAL, "upsampling needs even sizes"};
        if (!u || !skip || !cat) throw Error{DTD_EINVAL, "null array"};
        nn::check_aligned(u, 16), nn::check_aligned(skip, 16), nn::check_aligned(cat, 16);
        nn::k_up_cat<<<nn::blocks((long)n * h * w * ((cu + cs) / 8)), 256, 0, ctx->stream>>>(
            nn::Shape{n, h, w}, cu, cs, u, skip, cat);
        LAUNCH_CHECK();
    });
}

int dtd_nn_up_cat_bwd(dtd_ctx *ctx, int n, int h, int w, int cu, int cs, const uint16_t *dcat,
                      const uint16_t *dpool, uint16_t *du, uint16_t *dskip) {
    return guard(ctx, [&] {
        begin(ctx);
        nn::check_shape(n, h, w), nn::check_channels(cu), nn::check_channels(cs);
        if (h % 2 || w % 2) throw Error{DTD_EINVAL, "upsampling needs even sizes"};
        if (!dcat) throw Error{DTD_EINVAL, "null array"};
        nn::check_aligned(dcat, 16), nn::check_aligned(dpool, 16);
        nn::check_aligned(du, 16), nn::check_aligned(dskip, 16);
        const nn::Shape sh{n, h, w};
        if (du) {
            nn::k_up_cat_bwd_u<<<nn::blocks((long)n * (h / 2) * (w / 2) * (cu / 8)), 256, 0,
                                 ctx->stream>>>(sh, cu, cs, dcat, du);
            LAUNCH_CHECK();
        }
        if (dskip) {
            nn::k_skip_bwd<<<nn::blocks((long)n * h * w * (cs / 8)), 256, 0, ctx->stream>>>(
                sh, cu, cs, dcat, dpool, dskip);
            LAUNCH_CHECK();
        }
    });
}

int dtd_nn_input(dtd_ctx *ctx, size_t np, const double *h, const double *kd, const double *base,
                 double dt, double dx, double scale, uint16_t *x) {
    return guard(ctx, [&] {
        begin(ctx);
        if (!h || !kd || !x) throw Error{DTD_EINVAL, "null array"};
        if (!(scale > 0.0) || !(dt > 0.0) || !(dx > 0.0))
            throw Error{DTD_EINVAL, "scale, dt and dx must be positive"};
        nn::check_aligned(x, 16);
        if (np == 0) return;
        nn::k_nn_input<<<nn::blocks((long)np), 256, 0, ctx->stream>>>((long)np, h, kd, base, dt,
                                                                      dx * dx, scale, x);
        LAUNCH_CHECK();
    });
}

int dtd_nn_loss(dtd_ctx *ctx, size_t np, const float *out, int outc, const double *t,
                const double *base, double scale, float *parts, uint16_t *dy, double *loss) {
    return guard(ctx, [&] {
        begin(ctx);
        if (!out || !t || !base || !parts || !dy || !loss) throw Error{DTD_EINVAL, "null array"};
        if (np == 0 || outc < 1) throw Error{DTD_EINVAL, "empty loss"};
        const unsigned nb = nn::blocks((long)np);
        nn::k_nn_loss<<<nb, 256, 0, ctx->stream>>>((long)np, out, outc, t, base, scale, parts,
                                                   dy);
        LAUNCH_CHECK();
        std::vector<float> hp(nb);
        download(ctx, hp.data(), parts, nb * sizeof(float));
        double s = 0.0;
        for (float v : hp) s += v;
        *loss = s / (double)np;
    });
}

int dtd_nn_guess(dtd_ctx *ctx, size_t np, const double *base, const float *out, int outc,
                 double scale, double *g) {
    return guard(ctx, [&] {
        begin(ctx);
        if (!base || !out || !g) throw Error{DTD_EINVAL, "null array"};
        if (outc < 1) throw Error{DTD_EINVAL, "outc must be at least 1"};
        if (np == 0) return;
        nn::k_nn_guess<<<nn::blocks((long)np), 256, 0, ctx->stream>>>((long)np, base, out, outc,
                                                                      scale, g);
        LAUNCH_CHECK();
    });
}

int dtd_nn_block_mean(dtd_ctx *ctx, size_t w, size_t h, int f, int log_mean, const double *in,
                      double *out) {
    return guard(ctx, [&] {
        begin(ctx);
        check_grid(w, h);
        if (f < 1 || w % f || h % f) throw Error{DTD_EINVAL, "w and h must be multiples of f"};
        if (!in || !out) throw Error{DTD_EINVAL, "null array"};
        const long nc = (long)(w / f) * (long)(h / f);
        if (nc == 0) return;
        nn::k_block_mean<<<nn::blocks(nc), 256, 0, ctx->stream>>>((int)w, (int)h, f, log_mean, in,
                                                                   out);
        LAUNCH_CHECK();
    });
}

int dtd_nn_guess_coarse(dtd_ctx *ctx, size_t w, size_t h, int f, const double *base,
                        const float *out, int outc, double scale, double *g) {
    return guard(ctx, [&] {
        begin(ctx);
        check_grid(w, h);
        if (f < 1 || w % f || h % f) throw Error{DTD_EINVAL, "w and h must be multiples of f"};
        if (!base || !out || !g) throw Error{DTD_EINVAL, "null array"};
        if (outc < 1) throw Error{DTD_EINVAL, "outc must be at least 1"};
        if (w * h == 0) return;
        nn::k_nn_guess_coarse<<<nn::blocks((long)(w * h)), 256, 0, ctx->stream>>>(
            (int)w, (int)h, f, base, out, outc, scale, g);
        LAUNCH_CHECK();
    });
}

int dtd_lod_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *h_dev,
                   const double *kd_dev, double dt, double *out_dev) {
    return guard(ctx, [&] {
        begin(ctx);
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!h_dev || !kd_dev || !out_dev) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        Workspace &ws = ctx->ws;
        lod(ctx, h_dev, ws.x, kd_dev, dx, dt, wrap_x);
        HIPCHK(hipMemcpyAsync(out_dev, ws.x, n * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
    });
}

int dtd_nn_adam(dtd_ctx *ctx, size_t n, float *w, float *m, float *v, const float *g, float lr,
                float beta1, float beta2, float eps, int step, uint16_t *wb) {
    return guard(ctx, [&] {
        begin(ctx);
        if (!w || !m || !v || !g) throw Error{DTD_EINVAL, "null array"};
        if (step < 1) throw Error{DTD_EINVAL, "step counts from 1"};
        if (n == 0) return;
        const float bc1 = 1.0f - std::pow(beta1, (float)step);
        const float bc2 = 1.0f - std::pow(beta2, (float)step);
        nn::k_adam<<<nn::blocks((long)n), 256, 0, ctx->stream>>>((long)n, w, m, v, g, lr, beta1,
                                                                beta2, eps, bc1, bc2, wb);
        LAUNCH_CHECK();
    });
}

int dtd_nn_to_bf16(dtd_ctx *ctx, size_t n, const float *a, uint16_t *b) {
    return guard(ctx, [&] {
        begin(ctx);
        if (!a || !b) throw Error{DTD_EINVAL, "null array"};
        if (n == 0) return;
        nn::k_to_bf16<<<nn::blocks((long)n), 256, 0, ctx->stream>>>((long)n, a, b);
        LAUNCH_CHECK();
    });
}

int dtd_sync(dtd_ctx *ctx) {
    return guard(ctx, [&] { HIPCHK(hipStreamSynchronize(ctx->stream)); });
}

int dtd_solve_rel_guess_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                               double *h_dev, const double *kd_dev, double dt, double rel_tol,
                               const double *x0_dev, dtd_stats *stats) {
    return guard(ctx, [&] {
        begin(ctx);
        dtd_stats local{};
        dtd_stats *st = stats ? stats : &local;
        *st = dtd_stats{};
        if (!(std::isfinite(rel_tol) && rel_tol > 0.0))
            throw Error{DTD_EINVAL, "rel_tol must be finite and positive"};
        check_grid(w, h);
        const size_t n = w * h;
        if (n == 0) return;
        if (!h_dev || !kd_dev || !x0_dev) throw Error{DTD_EINVAL, "null array"};
        fit(ctx, (int)w, (int)h, wrap_x && w > 1);
        Workspace &ws = ctx->ws;
        HIPCHK(hipMemcpyAsync(ws.x, x0_dev, n * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
        double rel = 0.0, bnorm = 0.0;
        solve_core(ctx, h_dev, kd_dev, dx, dt, wrap_x, 0.0, rel_tol * rel_tol, true, kGuessKeep,
                   st, &rel, &bnorm);
        double tol_m = rel_tol * bnorm / std::sqrt((double)n);
        while (rel > rel_tol && st->restarts < 4 && bnorm > 0.0) {
            tol_m *= 0.5 * rel_tol / rel;
            st->restarts++;
            solve_core(ctx, h_dev, kd_dev, dx, dt, wrap_x, tol_m, 0.0, false, kGuessKeep, st,
                       &rel, &bnorm);
        }
        st->rel_residual = rel;
        copy_answer(ctx, h_dev);
    });
}

int dtd_tune(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *h_dev,
             const double *kd_dev, double dt, int reps, double *ms_rule, double *ms_tuned) {
    return guard(ctx, [&] {
        begin(ctx);
        if (ms_rule) *ms_rule = 0.0;
        if (ms_tuned) *ms_tuned = 0.0;
        check_grid(w, h);
        const size_t n = w * h;
        const int wrap = wrap_x && w > 1;
        const PlanKey key{(int)w, (int)h, wrap, ctx->opts.cycle_precision};
        if (reps < 1) {
            // Forget the shape's plan: the built-in rules from the next fit on.
            ctx->plans.erase(key);
            if (ctx->ws.w == (int)w && ctx->ws.h == (int)h && ctx->ws.wrap == wrap) {
                apply_plan(ctx->ws, Plan{});
                for (auto &g : ctx->ws.graph)
                    if (g) (void)hipGraphExecDestroy(g), g = nullptr;
                upload_levels(ctx->ws, ctx->stream);
                HIPCHK(hipStreamSynchronize(ctx->stream));
            }
            return;
        }
        if (n == 0) return;
        if (!h_dev || !kd_dev) throw Error{DTD_EINVAL, "null array"};
        if (!(std::isfinite(dx) && dx > 0.0 && std::isfinite(dt) && dt > 0.0))
            throw Error{DTD_EINVAL, "dx and dt must be finite and positive"};
        fit(ctx, (int)w, (int)h, wrap);
        Plan plan;
        tune(ctx, h_dev, kd_dev, dx, dt, wrap_x, std::min(reps, 50), plan);
        ctx->plans[key] = plan;
        if (ms_rule) *ms_rule = plan.ms_rule;
        if (ms_tuned) *ms_tuned = plan.ms_tuned;
    });
}

int dtd_tune_plan(const dtd_ctx *ctx, size_t w, size_t h, int wrap_x, char *buf, size_t cap) {
    if (!ctx || !buf || cap == 0) return DTD_EINVAL;
    const int wrap = wrap_x && w > 1;
    auto it = ctx->plans.find(PlanKey{(int)w, (int)h, wrap, ctx->opts.cycle_precision});
    std::string t;
    if (it == ctx->plans.end()) {
        t = "{}";
    } else {
        const Plan &p = it->second;
        t = "{\"levels\": [";
        for (size_t l = 0; l < p.lvl.size(); l++)
            t += (l ? ", [" : "[") + std::to_string(p.lvl[l][0]) + ", " + std::to_string(p.lvl[l][1]) + "]";
        t += "], \"direction\": " + std::to_string(p.th_dir) + ", \"residual\": " +
             std::to_string(p.th_resid) + ", \"finish1\": " + std::to_string(p.th_fin1) +
             ", \"finish2\": " + std::to_string(p.th_fin2) + ", \"bottom\": " +
             std::to_string(p.bottom) + ", \"ms_rule\": " + std::to_string(p.ms_rule) +
             ", \"ms_tuned\": " + std::to_string(p.ms_tuned) + "}";
    }
    if (t.size() + 1 > cap) return DTD_EINVAL;
    std::memcpy(buf, t.c_str(), t.size() + 1);
    return DTD_OK;
}

int dtd_inject_fault(dtd_ctx *ctx, long after, long *calls) {
    if (!ctx) return DTD_EINVAL;
    ctx->fault.after = after < 0 ? -1 : after;
    if (calls) *calls = ctx->fault.calls;
    return DTD_OK;
}

}  // extern "C"
