// Host-side state shared by the library's translation units: errors, the context and its
// single-grid workspace, and the launch helpers of the streamed kernels.
#pragma once

#include "../../include/dtdiff.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace dtd {

struct Error {
    int code;
    std::string msg;
    bool hip = false;  // raised by a HIP call (the context's device state is then dropped)
};

// Failure injection (a test hook, dtd_inject_fault): counts the checked HIP calls made for a
// context and, once armed, skips the call `after` further calls on and reports it as a device
// error, so the error paths run without any real device fault.
struct FaultHook {
    long after = -1;  // < 0: disarmed
    long calls = 0;
};
inline thread_local FaultHook *t_fault = nullptr;  // the hook of the entry point running here
inline bool fault_now() {
    FaultHook *f = t_fault;
    if (!f) return false;
    f->calls++;
    if (f->after < 0) return false;
    if (f->after-- > 0) return false;
    f->after = -1;
    return true;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        if (::dtd::fault_now())                                                          \
            throw ::dtd::Error{DTD_EDEVICE, std::string(#expr) + ": injected failure",   \
                               true};                                                    \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::dtd::Error{e_ == hipErrorOutOfMemory ? DTD_ETOOBIG : DTD_EDEVICE,    \
                               std::string(#expr) + ": " + hipGetErrorString(e_), true}; \
    } while (0)

#define LAUNCH_CHECK() HIPCHK(hipGetLastError())

inline int nblocks(long n, int bs = kBlock) { return (int)((n + bs - 1) / bs); }

template <typename T>
T *dalloc(size_t n) {
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    return (T *)p;
}

// The level shapes `Level::coarsen` produces (MG:117-123; `shapes`, MGGPU:336-347): halve,
// rounding up, until a level holds at most 64 cells or stops shrinking; a coarse level wraps
// when the finer one does and it is more than one cell wide.
struct Shape {
    int w, h, wrap;
};
inline std::vector<Shape> shapes(int w, int h, int wrap) {
    std::vector<Shape> out{{w, h, wrap}};
    for (;;) {
        Shape s = out.back();
        int wc = (s.w + 1) / 2, hc = (s.h + 1) / 2;
        if ((long)s.w * s.h <= 64 || (wc == s.w && hc == s.h)) return out;
        out.push_back({wc, hc, s.wrap && wc > 1});
    }
}

// Device memory of one grid shape (the reference's Workspace, D:213-260, and the device cycle's
// levels, MGGPU:350-558, in one place), kept while the shape stays.
struct Workspace {
    int w = 0, h = 0, wrap = 0, prec = -1;
    long n = 0;
    double *b = nullptr, *x = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr;
    double *r2 = nullptr;  // the residual's second buffer (the f32 cycle's fused update)
    double *ap = nullptr, *kd = nullptr, *parts = nullptr, *pair = nullptr;
    long nparts = 0;  // blocks whose partial sums parts holds (two doubles a block)
    double2 *es = nullptr;  // the f64 operator: (east, south) of each cell
    double *sums = nullptr;  // DTD_SUMS_STRIPS: strip sums (two sets) and the put-back's row pairs
    CgState *st = nullptr;
    std::vector<Lvl<float>> lf;
    std::vector<Lvl<double>> ld;
    Lvl<float> *d_lf = nullptr;  // device copies of the level tables (the bottom-of-cycle kernel)
    Lvl<double> *d_ld = nullptr;
    std::vector<void *> owned;  // every allocation of the levels
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // Band heights of the finest level's f64 passes (direction, first residual, the two finish
    // passes) chosen by the tuner, and the level size from which the one-block bottom of the
    // cycle takes over; 0 and -1: the built-in rules.
    int th_dir = 0, th_resid = 0, th_fin1 = 0, th_fin2 = 0, bottom = -1;
    // The Jacobi solve's extra buffers (diag, two row sums a row, its scalars) and graphs.
    double *jdiag = nullptr, *jrows = nullptr;
    void *jst = nullptr;
    // The rows' live strips of the Jacobi loop (one byte a strip of 8 rows, and one a row for the
    // zero-row test) and its flags (any zero row, a zero row's elevation not normal, the visited
    // strips, a visited value not finite).
    unsigned char *jzero = nullptr, *jlive = nullptr;
    int *jflags = nullptr, *jlist = nullptr;  // jlist: the visited strips in order
    hipGraphExec_t jgraph[2] = {nullptr, nullptr};
};

// A tuned launch plan for one grid shape (dtd_tune): per level the fused smoothers' band
// heights, the finest level's f64 passes' band heights and the bottom's size.
struct Plan {
    std::vector<std::array<int, 2>> lvl;  // (th_pre, th_post) per level
    int th_dir = 0, th_resid = 0, th_fin1 = 0, th_fin2 = 0, bottom = -1;
    double ms_rule = 0.0, ms_tuned = 0.0;  // whole solves timed with the built-in rules and the plan
};
using PlanKey = std::array<int, 4>;  // w, h, wrap, cycle precision

// The batched solve's band heights by class of level (dtd_tune_batch): levels of at least 64
// rows, of 16 to 63 rows, of fewer; and the f64 passes of grids of at least 64 rows.
struct BatchPolicy {
    int tall = 16, mid = 8, small = 4, f64 = 16;
    // The finest level's fused pre- and post-smoothing (records formed from kd) of grids of at
    // least 64 rows: their own heights, the pre-smoothing holding more registers.
    int fine_pre = 32, fine_post = 24;
    bool operator==(const BatchPolicy &o) const {
        return tall == o.tall && mid == o.mid && small == o.small && f64 == o.f64 &&
               fine_pre == o.fine_pre && fine_post == o.fine_post;
    }
    int level(int h) const { return h >= 64 ? tall : (h >= 16 ? mid : small); }
    int pass(int h) const { return h >= 64 ? f64 : (h >= 16 ? mid : small); }
    int finest(bool pre, int h) const { return h >= 64 ? (pre ? fine_pre : fine_post) : level(h); }
};

struct BatchWs;
struct FailWs;

}  // namespace dtd

namespace dtd {
// The opt-in audit of the multigrid solve (`diffusion::audit`, D:118-199): over the audited
// solves, for the pairs [multigrid - Jacobi, Jacobi - tight, multigrid - tight], the largest
// |difference| of any cell, the largest RMS of one solve and the sum of squares.
struct AuditSums {
    unsigned long long solves = 0, its_new = 0, its_old = 0;
    double max[3] = {0.0, 0.0, 0.0}, max_rms[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
    double cells = 0.0;
};
}  // namespace dtd

struct dtd_ctx {
    dtd_opts opts;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    hipEvent_t flag_ev[2] = {};
    dtd::CgState *host_st = nullptr;  // pinned
    int *host_flag = nullptr;         // pinned and coherent, two slots the device writes
    int *dev_flag = nullptr;          // host_flag's device address
    double *host_pair = nullptr;      // pinned, two doubles
    dtd::Workspace ws;
    std::string err;
    size_t cycle_levels = 0;  // levels built by dtd_cycle_new
    bool use_graph = true;
    // The finest level's fused smoothers form their records (and the pre-smoothing A p) from kd
    // and p instead of reading them (DTD_NO_KD=1 reads them, for comparison).
    bool use_kd = true;
    bool use_live = true;  // the Jacobi loop skips the rows `Live` skips (DTD_NO_LIVE=1: it visits all)
    dtd::FaultHook fault;         // dtd_inject_fault
    std::map<dtd::PlanKey, dtd::Plan> plans;  // tuned plans by grid shape (dtd_tune)
    // Tuned band-height policies of the batched solve by shape list (dtd_tune_batch), and the one
    // a batch plan is built with.
    std::map<std::vector<long>, dtd::BatchPolicy> batch_plans;
    dtd::BatchPolicy batch_policy;
    bool batch_tuning = false;  // dtd_tune_batch sets batch_policy itself
    double *stage = nullptr;      // pinned host staging of dtd_solve's answer
    size_t stage_n = 0;
    dtd::BatchWs *batch = nullptr;  // the batched solve's buffers (batch.hpp)
    dtd::FailWs *fail = nullptr;    // the slope failure's buffers (failure.hpp)
    // The audit (DEEPTIME_DIFFUSION_AUDIT at creation, or dtd_audit): its sums, and its device
    // buffer (the audited answer, the Jacobi answer and the per-block partials).
    bool audit_on = false;
    dtd::AuditSums audit;
    double *audit_buf = nullptr;
    size_t audit_n = 0;
};

namespace dtd {

void free_ws(Workspace &ws);
void free_batch(dtd_ctx *ctx);  // drops the batched solve's buffers (batch.hpp)
void free_fail(dtd_ctx *ctx);   // drops the slope failure's buffers (failure.hpp)

// Band height for a streaming kernel on a w x h level (a multiple of 4 rows; DTD_BAND_ROWS
// overrides it).
template <int V, int H = LaneGeom<V>::halo>
int band_height(const void *kernel, int w, int h) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    static const int forced = [] {
        const char *v = getenv("DTD_BAND_ROWS");
        return v ? atoi(v) : 0;
    }();
    if (forced > 0) return std::max(4, (forced + 3) / 4 * 4);
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0));
    const long resident = (long)std::max(per_cu, 1) * cus * (kBlock / 64);
    // Short bands in many rounds keep the waves in flight on a compact stretch of rows (the best
    // measured on the finest levels); levels too small for two rounds take 4 rows.
    for (int th = 16; th >= 8; th >>= 1)
        if ((long)stream_waves<V, H>(w, h, th) >= 2 * resident) return th;
    return 4;
}

template <int V, int H>
int stream_grid(int w, int h, int th) {
    return (stream_waves<V, H>(w, h, th) + kBlock / 64 - 1) / (kBlock / 64);
}

// The most blocks a launch over a w x h grid can have when each block leaves a partial sum: the
// cell-parallel passes take 256 cells a block; the streamed passes at their narrowest bands
// (60 owned columns, bands of at least 4 rows, 4 waves a block) take more on narrow grids.
inline long parts_cap(int w, int h) {
    const long waves = (long)((w + 59) / 60) * ((h + 3) / 4);
    return std::max<long>(nblocks((long)w * h), (waves + kBlock / 64 - 1) / (kBlock / 64));
}

// A launch of nb blocks that each write a partial sum (or a pair) must fit the buffer.
inline void need_parts(long nb, long cap) {
    if (nb > cap)
        throw Error{DTD_ESTATE, "a launch would write more partial sums than its buffer holds"};
}

inline float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    return ms;
}

inline void check_grid(size_t w, size_t h) {
    // Cell indices are 32-bit; keep a margin for the last block's threads.
    const size_t cap = 0x7fff0000ull;
    if (w > cap || h > cap || (w && h > cap / w))
        throw Error{DTD_ETOOBIG, "the grid exceeds 2^31 - 2^16 cells"};
}

inline void begin(dtd_ctx *ctx) {
    if (!ctx) throw Error{DTD_EINVAL, "null context"};
    ctx->err.clear();
    HIPCHK(hipSetDevice(ctx->opts.device));
}

// Device copy of a host array on the context's stream.
inline void upload(dtd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
}
inline void download(dtd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

// After a device error: end a stream capture the error interrupted, let the work already
// queued finish, and drop every buffer, level and graph, so the next call builds them afresh.
inline void drop_device_state(dtd_ctx *ctx) {
    if (!ctx->stream) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipGetLastError();
    free_ws(ctx->ws);
    free_batch(ctx);
    free_fail(ctx);
    if (ctx->audit_buf) (void)hipFree(ctx->audit_buf);
    ctx->audit_buf = nullptr, ctx->audit_n = 0;
    ctx->cycle_levels = 0;
}

// Runs one entry point's body: errors become a code and the context's message; an error a HIP
// call raised (a device error, or running out of device memory) or a host exception drops the
// context's device state, which the next call builds afresh.
template <class F>
int guard(dtd_ctx *ctx, F f) {
    FaultHook *outer = t_fault;
    t_fault = ctx ? &ctx->fault : nullptr;
    int code = DTD_OK;
    bool drop = false;
    try {
        f();
    } catch (const Error &e) {
        if (ctx) ctx->err = e.msg;
        code = e.code;
        drop = e.hip;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        code = DTD_EDEVICE;
        drop = true;
    }
    t_fault = outer;
    if (drop && ctx) drop_device_state(ctx);
    return code;
}

}  // namespace dtd
