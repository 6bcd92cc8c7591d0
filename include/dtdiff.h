/*
 * dtdiff: the implicit hillslope/marine diffusion solve of deep-time on AMD Instinct MI355X
 * (gfx950), behind a C ABI.
 *
 * The system is (I + dt * div K grad) h1 = h0 on a w x h raster of square cells of side dx,
 * periodic in x when wrap_x (and w > 1), with no flux through the first and last rows.  Face
 * couplings are 0.5 * (K_i + K_j) * dt / dx^2.  It is solved by f64 conjugate gradients
 * preconditioned by one multigrid V-cycle (2x2 aggregation, coarse faces scaled by 1/2, damped
 * Jacobi with omega = 0.8, two sweeps before and after, 16 sweeps on the coarsest level of at
 * most 64 cells), from the locally one-dimensional splitting's first guess, to an RMS residual
 * of tol_m metres, after which the residual's sum is returned to the cells in proportion to
 * their coupling, so sum(h1) == sum(h0) to round-off.
 *
 * Reference interfaces replaced (paths relative to the deep-time repository root):
 *   dtd_solve / dtd_solve_device      crates/deeptime-core/src/surface/diffusion.rs:77-116
 *                                     (`solve` with DiffusionSolver::Multigrid -> `multigrid_in`
 *                                      -> `lod` + `multigrid::cg` -> `cg_on`, multigrid.rs:562-763)
 *   dtd_create / dtd_destroy          the session's device slot, compute.rs:110-122
 *                                     (`Ambient::kept`), whose levels stay resident between solves
 *   dtd_cycle_new / dtd_cycle_precondition
 *                                     `DeviceCycle::new` / `DeviceCycle::precondition`,
 *                                     multigrid/gpu.rs:587 and :661 (the device seam of `Pre`,
 *                                     multigrid.rs:385-458)
 *   dtd_cycle_level                   the levels `Level::finest`/`Level::coarsen` build,
 *                                     multigrid.rs:96-199
 *   dtd_operator / dtd_apply          `Operator::new` / `Operator::apply`, diffusion.rs:532-663
 *   dtd_lod                           `lod`, diffusion.rs:370-489 (every row visited)
 *
 * Errors: every entry point returns DTD_OK (0) or a negative DTD_E* code; the message is kept
 * per context (dtd_last_error).  On error the caller's output arrays are left as they were.
 * No exception crosses the ABI.  A context is not thread-safe; distinct contexts are
 * independent.  Results do not depend on timing: every device reduction sums in a fixed order.
 */
#ifndef DTDIFF_H
#define DTDIFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTD_OK 0
#define DTD_EINVAL (-1)   /* a bad argument (NULL pointer, non-finite or non-positive tol) */
#define DTD_ETOOBIG (-2)  /* the grid exceeds 2^31 - 1 cells or the device memory */
#define DTD_EDEVICE (-3)  /* a HIP runtime error */
#define DTD_ESTATE (-4)   /* a call out of order (dtd_cycle_precondition before dtd_cycle_new) */

/* `DiffusionSolver`, crates/deeptime-core/src/surface/params.rs:154. */
#define DTD_SOLVER_MULTIGRID 0
#define DTD_SOLVER_JACOBI 1

/* Precision of the V-cycle that preconditions the f64 conjugate gradients. */
#define DTD_CYCLE_F32 0  /* the reference's: every level in f32 (multigrid.rs:24-27) */
#define DTD_CYCLE_F64 1  /* every level in f64 (the plain FP64 multigrid) */
#define DTD_CYCLE_BF16 2 /* the f32 cycle, its two finest levels' smoothers reading their
                            operator records stored in bf16 (a reduced-storage variant) */

/* Transfers between the cycle's levels. */
#define DTD_TRANSFER_AGGREGATE 0 /* the reference's: 2x2 aggregation (restriction sums the block,
                                    prolongation copies the coarse value; multigrid.rs:117-356) */
#define DTD_TRANSFER_BILINEAR 1  /* cell-centred bilinear prolongation (weights 3/4, 1/4 along each
                                    axis from the two nearest coarse centres) and its transpose as
                                    the restriction (full weighting); no reference counterpart */

/* Solve of the cycle's coarsest level (at most 64 cells). */
#define DTD_COARSEST_SWEEPS 0 /* the reference's: 16 damped-Jacobi sweeps in the cycle's precision */
#define DTD_COARSEST_EXACT 1  /* exact: Cholesky in f64 (factorised once a solve); no reference
                                 counterpart */

/* First guess of a solve. */
#define DTD_GUESS_LOD 0  /* the splitting, as `multigrid_in` (diffusion.rs:110-114) */
#define DTD_GUESS_RHS 1  /* h0 itself */

/* Order of the multigrid solve's sums (r.r, p.A p, r.z, the put-back's two). */
#define DTD_SUMS_BLOCKS 0 /* per block of the fused passes, the blocks in order: run to run the
                             same, fast, but not the reference's rounding (its iterates agree to
                             ~1e-15 and its iteration count can differ by one or two) */
#define DTD_SUMS_STRIPS 1 /* the reference's: per strip of 8 rows in row-major order, then the
                             strips in order (multigrid.rs:596-723, 314-318), the put-back per row
                             then the rows (multigrid.rs:740-755), with the splitting's own
                             recurrences as the first guess: the reference's iterates and
                             iteration count bit for bit.  With the reference's cycle only
                             (DTD_CYCLE_F32, aggregation, 16 coarsest sweeps); several times
                             slower. */

typedef struct dtd_ctx dtd_ctx;

/* Settings of a context; dtd_default_opts gives the reference's (multigrid.rs:45-56,
 * diffusion.rs:497). */
typedef struct dtd_opts {
    int device;           /* HIP device ordinal */
    float omega;          /* Jacobi damping, 0.8 */
    unsigned nu;          /* sweeps before and after the coarse correction, 2 */
    float scale;          /* coarse coupling share, 0.5 */
    unsigned max_iter;    /* conjugate-gradient iteration cap, 400 */
    int cycle_precision;  /* DTD_CYCLE_F32 */
    int guess;            /* DTD_GUESS_LOD */
    int transfer;         /* DTD_TRANSFER_AGGREGATE (DTD_TRANSFER_BILINEAR: not with DTD_CYCLE_BF16,
                             nor in the batched solve) */
    int coarsest;         /* DTD_COARSEST_SWEEPS (DTD_COARSEST_EXACT: not in the batched solve) */
    int sums;             /* DTD_SUMS_BLOCKS (DTD_SUMS_STRIPS: single-grid multigrid solves only) */
} dtd_opts;

/* What a solve did (all times in milliseconds, from HIP events on the context's stream). */
typedef struct dtd_stats {
    size_t iterations;    /* conjugate-gradient iterations, summed over restarts */
    size_t restarts;      /* extra solves dtd_solve_rel_device took to meet its target */
    double rel_residual;  /* ||h0 - A h1||_2 / ||h0||_2 in f64 after the mass put-back */
    double ms_setup;      /* operator, levels */
    double ms_guess;      /* first guess */
    double ms_cg;         /* conjugate gradients */
    double ms_total;      /* the whole solve on the device */
    size_t rows_visited;  /* the Jacobi step: rows its loop visits (h unless `Live` skips some,
                             diffusion.rs:285-367); 0 for the multigrid solves */
} dtd_stats;

void dtd_default_opts(dtd_opts *opts);

/* A context on opts->device (NULL opts: the defaults on device 0). */
int dtd_create(const dtd_opts *opts, dtd_ctx **out);
void dtd_destroy(dtd_ctx *ctx);
/* The last error's message ("" when none); valid until the next call on ctx. */
const char *dtd_last_error(const dtd_ctx *ctx);

/* `diffusion::solve` (diffusion.rs:77-97) on host arrays: h_inout holds h0 on entry and h1 on
 * return; kd (m^2/yr) has w*h cells; dt in years.  DTD_SOLVER_MULTIGRID stops at the RMS
 * residual tol_m (metres); DTD_SOLVER_JACOBI is `step_in` (diffusion.rs:262-283), which keeps the
 * reference's own 1e-6 m (tol_m is not used).  *iterations (may be NULL) receives the
 * conjugate-gradient iterations.  w*h == 0 returns 0 iterations. */
int dtd_solve(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_inout,
              const double *kd, double dt, int solver, double tol_m, size_t *iterations);

/* `step_in` (diffusion.rs:262-283, the Jacobi-preconditioned solve) on device-resident arrays:
 * h_dev holds h0 on entry and h1 on return.  Its conjugate-gradient sums are the reference's, in
 * its order (per row in column order, then rows in order), so its iterations and iterates are the
 * reference's; the put-back's sums are taken in another order (its step differs by rounding).
 * stats (may be NULL) gets the iterations and device times. */
int dtd_step_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_dev,
                    const double *kd_dev, double dt, dtd_stats *stats);

/* The same solve on device-resident arrays (h_dev in/out and kd_dev: device pointers of w*h
 * doubles on the context's device).  stats may be NULL. */
int dtd_solve_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_dev,
                     const double *kd_dev, double dt, double tol_m, dtd_stats *stats);

/* dtd_solve_device for callers that cannot allocate device memory (code that must not call
 * hipMalloc: tight serving loops, harnesses that forbid it).  No counterpart in the reference.
 * Same arguments, kernels and results (bit for bit); the caller still passes h_dev and kd_dev.
 * The workspace (f64 vectors, partial sums, the cycle's levels) is laid out in one fixed 64 MiB
 * device array per GPU, which the library holds from its load until the process ends, instead of
 * the context's buffers; nothing is allocated or freed, and the iterations are launched directly
 * (no graphs).  A grid whose workspace exceeds 64 MiB (about 800 000 cells with the f32 cycle,
 * e.g. 1024 x 768) is refused with DTD_ETOOBIG and a message giving both sizes.  Every call on one
 * GPU, from any context or thread, shares that GPU's array and overwrites what the last call left
 * there: calls that overlap in time on one GPU race and must not be made. */
int dtd_solve_device_noalloc(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                             double *h_dev, const double *kd_dev, double dt, double tol_m,
                             dtd_stats *stats);

/* The solve to a true relative residual: ||h0 - A h1||_2 / ||h0||_2 <= rel_tol in f64, measured
 * after the mass put-back.  Runs the solve at tol_m = rel_tol * ||h0|| / sqrt(n) and, while the
 * measured residual misses the target, solves again from the last answer at a tighter tol_m
 * (at most 4 restarts). */
int dtd_solve_rel_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_dev,
                         const double *kd_dev, double dt, double rel_tol, dtd_stats *stats);

/* `Operator::new` (host arrays): east, south, diag of w*h doubles each. */
int dtd_operator(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *kd,
                 double dt, double *east, double *south, double *diag);

/* `Operator::apply` with the operator built from kd: y = (I + dt L) x (host arrays). */
int dtd_apply(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *kd,
              double dt, const double *x, double *y);

/* `lod` on host arrays: h_inout becomes the splitting's first guess. */
int dtd_lod(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, double *h_inout,
            const double *kd, double dt);

/* `DeviceCycle::new`: the cycle's levels for the f64 operator (host arrays of w*h doubles). */
int dtd_cycle_new(dtd_ctx *ctx, size_t w, size_t h, int wrap_x, const double *east,
                  const double *south, const double *diag);
/* Number of levels of the cycle dtd_cycle_new built (0 before). */
size_t dtd_cycle_levels(const dtd_ctx *ctx);
/* Level l's shape and f32 operator (NULL outputs are skipped; mass is 1 on level 0). */
int dtd_cycle_level(dtd_ctx *ctx, size_t l, size_t *w, size_t *h, int *wrap, float *east,
                    float *south, float *diag, float *wd, float *mass);
/* `DeviceCycle::precondition`: z32 = M^-1 r32 by one V-cycle from zero; *rz = sum r64 * z32. */
int dtd_cycle_precondition(dtd_ctx *ctx, const float *r32, const double *r64, float *z32,
                           double *rz);

/* The surface stage's diffusivity fill before its solve (crates/deeptime-core/src/surface/
 * mod.rs:1982-1992) on device arrays of n cells: kd[c] = `submarine_kd` (surface/marine.rs:44-59)
 * of the water depth sea - h[c], blending to kd_land * diffu[class[c]] ashore.  diffu is the rock
 * classes' multiplier table (device, indexed by class).  The device's tanh and exp are not glibc's:
 * values agree with the CPU's to ~1e-14 relative where they are used, exactly elsewhere. */
int dtd_fill_kd(dtd_ctx *ctx, size_t n, const double *h_dev, const unsigned char *class_dev,
                const double *diffu_dev, double kd_land, double sea, double d0, double wave_base,
                double tidal_range, double *kd_dev);

/* The submarine slope failure's parameters (`FailureParams`, crates/deeptime-core/src/surface/
 * marine.rs:254-262; the stage uses rate 0.05 and 40 iterations, marine.rs:231-232). */
typedef struct dtd_failure {
    double critical_slope; /* the steepest gradient sediment holds under the sea (tan) */
    double rate;           /* share of a pair's excess over the critical drop moved per iteration */
    unsigned iterations;
} dtd_failure;

/* `marine::slope_failure` (crates/deeptime-core/src/surface/marine.rs:292-580), the stage's step
 * after the solve (surface/mod.rs:2012-2023), on device arrays of w*h cells: sea (0/1 a cell),
 * h and avail (in/out).  Sediment on the sea floor steeper than the critical slope slides to its
 * wet D8 neighbours, a cell giving at most its avail; an iteration that moves nothing ends the
 * loop.  The heights and avail are the reference's bit for bit; *moved (may be NULL) gets the
 * metres moved (sum |dz| / 2 over the movers, equal to rounding: the blocks' sums are added in
 * order, the reference's cell by cell), *iterations (may be NULL) the iterations run.  The
 * reference's own device path (marine/gpu.rs) rounds in f32; this one keeps f64. */
int dtd_slope_failure_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                             const unsigned char *sea_dev, double sea_level, const dtd_failure *p,
                             double *h_dev, double *avail_dev, double *moved,
                             unsigned *iterations);
/* The same on host arrays (copied to the device and back). */
int dtd_slope_failure(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                      const unsigned char *sea, double sea_level, const dtd_failure *p,
                      double *h_inout, double *avail_inout, double *moved, unsigned *iterations);

/* A cube grid on the device: what `CubeGrid` (crates/deeptime-core/src/sphere/cube.rs) hands the
 * cube's diffusion solve.  n x n cells on each of 6 faces, cell (face, i, j) at face n^2 + j n + i;
 * slots: each cell's 8 neighbours in D8 order in its face frame (E, NE, N, NW, W, SW, S, SE),
 * DTD_NO_CELL across a cube corner (`CubeTopology::slots`); lap: the Laplacian's 8 couplings a
 * cell (`CubeGrid::lap`, symmetric); area: the cells' areas and mean_area their mean
 * (`CubeGrid::areas`, `mean_area`).  The arrays are copied; the V-cycle's levels (their neighbour
 * lists and the finer entries each coarse value sums, cube/multigrid.rs:65-186) are built here,
 * once.  Free with dtd_cube_free (before the context). */
#define DTD_NO_CELL 0xFFFFFFFFu
typedef struct dtd_cube dtd_cube;
int dtd_cube_new(dtd_ctx *ctx, size_t n, const uint32_t *slots, const double *lap,
                 const double *area, double mean_area, dtd_cube **out);
void dtd_cube_free(dtd_cube *g);

/* `surface::cube::diffusion::solve` (crates/deeptime-core/src/surface/cube/diffusion.rs:227-314)
 * on host arrays of 6 n^2 cells: h (in/out), kd (m^2/yr), dt (yr).  DTD_SOLVER_MULTIGRID: every
 * cell, the face-wise first guess, CG preconditioned by the cube's f32 V-cycle to an RMS residual
 * of tol_m; DTD_SOLVER_JACOBI: `step_in` (CD:249-259), the cells with a coupling, the diagonal
 * preconditioner, its own 1e-6 m.  Then the residual's sum put back (Σ a h kept).  Every sum is
 * taken in the reference's order, so the iterations and h are the reference's.  dt NaN or <= 0
 * does nothing.  *iterations (may be NULL) gets the CG iterations. */
int dtd_cube_solve(dtd_ctx *ctx, dtd_cube *g, double *h_inout, const double *kd, double dt,
                   int solver, double tol_m, size_t *iterations);
/* The same on device arrays. */
int dtd_cube_solve_device(dtd_ctx *ctx, dtd_cube *g, double *h_dev, const double *kd_dev,
                          double dt, int solver, double tol_m, size_t *iterations);

/* One grid of a batched solve: a dtd_solve_device call's arguments. */
typedef struct dtd_grid {
    size_t w, h;        /* cells */
    double dx;          /* metres */
    int wrap_x;         /* periodic in x */
    double dt;          /* years */
    double tol_m;       /* RMS residual target in metres (used when rel_tol <= 0) */
    double *h_dev;      /* device: h0 on entry, h1 on return */
    const double *kd_dev; /* device: diffusivity, m^2/yr */
} dtd_grid;

typedef struct dtd_batch_stats {
    size_t grids;           /* non-empty grids solved */
    size_t cells;           /* their cells */
    size_t iterations_max;  /* most CG iterations of one grid (restarts included) */
    size_t iterations_sum;
    size_t restarts;        /* grid restarts of the relative target */
    double rel_residual_max; /* largest true relative residual after the put-back */
    double ms_setup, ms_guess, ms_cg, ms_total; /* device time */
} dtd_batch_stats;

/* Many independent grids solved together, each as dtd_solve_device (rel_tol <= 0: grid k stops
 * at its tol_m) or dtd_solve_rel_device (rel_tol > 0: every grid to a true relative residual of
 * rel_tol, with the same restarts) would solve it alone, with one launch per pass and level over
 * all the grids.  No reference counterpart: the reference solves one grid per call of
 * `surface::diffusion::solve` (crates/deeptime-core/src/surface/diffusion.rs:77-116); this is
 * that call over many regional grids at once.  The context must use the f32 cycle with nu = 2.
 * Empty grids are skipped (0 iterations).  iterations[k] and rel_residuals[k] (either may be
 * NULL) get grid k's iteration count and true relative residual.  A grid's answer does not depend
 * on the other grids of the batch.  The device buffers of the last batch shape stay resident. */
int dtd_solve_batch(dtd_ctx *ctx, size_t n, const dtd_grid *grids, double rel_tol,
                    size_t *iterations, double *rel_residuals, dtd_batch_stats *stats);

/* Kernels dtd_time_kernel can time on the workspace of the last f32-cycle solve. */
#define DTD_KERNEL_SWEEP 0      /* one damped-Jacobi sweep of the finest level (f32) */
#define DTD_KERNEL_DIRECTION 1  /* the conjugate gradients' direction pass (f64) */
#define DTD_KERNEL_UPDATE 2     /* the residual update pass (f64) */
#define DTD_KERNEL_RESTRICT 3   /* the finest level's residual restriction (f32) */
#define DTD_KERNEL_PRE 4        /* fused pre-smoothing + restriction of the finest level (f32) */
#define DTD_KERNEL_POST 5       /* fused post-smoothing (two sweeps + dot) of the finest level */
#define DTD_KERNEL_COPY 6       /* calibration: a float4 copy of one finest-level array */
#define DTD_KERNEL_STREAM5 7    /* calibration: five finest-level arrays combined into one */
#define DTD_KERNEL_BANDCOPY 8   /* calibration: the pre-smoother's band walk without its stencil */

/* The opt-in audit of the multigrid solve, `diffusion::audit`
 * (crates/deeptime-core/src/surface/diffusion.rs:118-199; switched on there by the environment
 * variable DEEPTIME_DIFFUSION_AUDIT, D:127-130, which dtd_create reads too).  While it is on, each
 * multigrid solve of dtd_solve and dtd_solve_device is followed by the Jacobi solve and by a
 * multigrid solve to an RMS residual of 1e-9 m (`audit::TIGHT_M`) of the same step, on the same
 * context, and the context adds, for [multigrid - Jacobi, Jacobi - tight, multigrid - tight], the
 * largest |difference| of any cell, the largest RMS of one solve and the sum of squares
 * (`audit::record`, D:148-176).  The caller's answer is the multigrid solve's, as without the
 * audit.  dtd_audit switches it (reset != 0 also clears the sums). */
int dtd_audit(dtd_ctx *ctx, int on, int reset);

/* `audit::audit_line` (D:179-199): the sums so far as one text line in buf (cap bytes),
 * "diffusion_audit solves=... its_multigrid=... its_jacobi=... new_vs_old_max_m=..." with the
 * reference's fields and number formats; DTD_ESTATE (and "") when no solve has been audited,
 * where the reference returns None. */
int dtd_audit_line(const dtd_ctx *ctx, char *buf, size_t cap);

/* Measurement hook, no reference counterpart: launches `kernel` reps times back to back on the
 * context's stream over the resident data of the last solve and returns the average duration
 * between two HIP events (ms) and the kernel's compulsory bytes per launch (every array it
 * touches read or written once, at its storage precision).  The workspace's vectors are
 * overwritten: solve again before reading results. */
int dtd_time_kernel(dtd_ctx *ctx, int kernel, int reps, double *ms_avg, double *bytes);

/* The launch tuner (the project's own; the reference tunes its cycle's settings by a hand-run
 * sweep, `tune_the_cycle`, multigrid/tests.rs:303-360, whose settings stay the reference's here).
 * For the w x h grid of (h_dev, kd_dev, dt) it searches how the solve's work is cut into launches:
 * the band height (rows a wave walks) of every streamed launch, the fused pre- and
 * post-smoothing of each level and the finest level's f64 passes, each candidate launched `reps`
 * times on the resident data and timed with HIP events; then the level size from which the
 * one-block bottom of the cycle takes over, by timing whole solves.  The plan is kept only if
 * whole solves with it are faster than with the built-in rules; either way it is kept in the
 * context for that shape (and cycle precision) and applied whenever the workspace takes the shape
 * again.  h_dev and kd_dev are only read.  *ms_rule and *ms_tuned (optional) get the median
 * device time of a whole solve with the built-in rules and with the plan kept.  reps < 1 forgets
 * the shape's plan.  A tuned solve's partial sums cover other cells, so its sums round differently
 * (the cycle's correction is unchanged); it meets the same stopping test. */
int dtd_tune(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *h_dev,
             const double *kd_dev, double dt, int reps, double *ms_rule, double *ms_tuned);

/* dtd_tune for the batched solve: the band heights of its launches by class of level (levels of
 * at least 64 rows, of 16-63 rows, of fewer, and the f64 passes of grids of at least 64 rows),
 * chosen by timing whole dtd_solve_batch calls on these grids (median of reps after one untimed
 * each), one class at a time, and kept in the context for this list of shapes; later batches of
 * the same shapes use it.  The grids' h arrays get their h0 back before every timed solve and at
 * the end.  *ms_rule and *ms_tuned (optional) get the batch's device time with the built-in
 * heights and with the kept ones. */
int dtd_tune_batch(dtd_ctx *ctx, size_t n, const dtd_grid *grids, double rel_tol, int reps,
                   double *ms_rule, double *ms_tuned);

/* The plan dtd_tune keeps for a shape, as JSON text in buf (cap bytes): per level the pre- and
 * post-smoothing band heights (0: the built-in rule), the f64 passes' band heights, the bottom's
 * size (-1: built-in) and the two timings; "{}" when the shape has no plan. */
int dtd_tune_plan(const dtd_ctx *ctx, size_t w, size_t h, int wrap_x, char *buf, size_t cap);

/* Test hook for the failure seam of the reference's device cycle (`Pre`, multigrid.rs:385-472,
 * whose recovery drops the device state and builds it again, multigrid.rs:443-455).  Arms a
 * simulated device error: of the checked HIP calls the context makes from now on, the one after
 * `after` others is not made, and the entry point that reaches it returns DTD_EDEVICE with a
 * message naming that call ("... injected failure"), as if HIP had failed there.  Nothing runs on
 * the device that would not run otherwise.  As after any device error, the caller's output arrays
 * keep their values, the context's buffers, levels and graphs are dropped, and the next call
 * builds them again.  after < 0 disarms the hook.  *calls (optional) gets the number of checked
 * HIP calls the context has made so far, so a test can step the failure through every call. */
int dtd_inject_fault(dtd_ctx *ctx, long after, long *calls);

/* ---- The learned first guess ---------------------------------------------------------------
 * No reference counterpart: the reference starts conjugate gradients from its splitting
 * (diffusion.rs:110-114).  These entry points run a small convolutional network that predicts h1
 * from (h0, ln kd, dt) and a solve that starts from a given guess; the same stopping test ends
 * the solve, so a guess can only change how many iterations it takes.
 *
 * Network tensors are device arrays of bf16 (uint16_t, the upper half of an f32) in NHWC order:
 * n images of h x w pixels with c channels, element (img, y, x, ch) at ((img h + y) w + x) c + ch;
 * w and every c must be multiples of 16.  A 3x3 convolution wraps in x and reads zeros past the
 * first and last rows, as the planet grid; its weights are wt[cout][9][cin] (tap 3 (dy + 1) +
 * dx + 1), bf16, its bias f32.  All calls are asynchronous on the context's stream except
 * dtd_nn_loss (which returns the loss) and dtd_sync. */

/* y = relu?(bias + conv3x3(x, wt)) (bias may be NULL); y32 (optional) also gets the first y32c
 * output channels in f32, y32[p y32c + ch].  y may be NULL when y32 is given (only the f32
 * channels are wanted). */
int dtd_nn_conv3x3(dtd_ctx *ctx, int n, int h, int w, int cin, int cout, const uint16_t *x,
                   const uint16_t *wt, const float *bias, int relu, uint16_t *y, float *y32,
                   int y32c);

/* dtd_nn_conv3x3 of the concatenation [nearest upsampling of u (h/2 x w/2, cu channels), skip
 * (h x w, cs channels)] without forming it (the weights as for cin = cu + cs, in that order); the
 * same outputs as dtd_nn_up_cat then dtd_nn_conv3x3.  Needs w a multiple of 128, h even, cu and cs
 * multiples of 32 and 16-byte aligned tensors. */
int dtd_nn_conv3x3_cat(dtd_ctx *ctx, int n, int h, int w, int cu, int cs, int cout,
                       const uint16_t *u, const uint16_t *skip, const uint16_t *wt,
                       const float *bias, int relu, uint16_t *y, float *y32, int y32c);

/* The input's gradient dx (cin channels) of a convolution from its output's gradient dz (cout
 * channels); wt_scratch holds cin * 9 * cout bf16. */
int dtd_nn_conv3x3_bwd_input(dtd_ctx *ctx, int n, int h, int w, int cin, int cout,
                             const uint16_t *dz, const uint16_t *wt, uint16_t *wt_scratch,
                             uint16_t *dx);

/* The scratch dtd_nn_conv3x3_bwd_weight needs: returns the bf16 elements of `planes`, and
 * *f32_elems gets the f32 elements of `part`. */
size_t dtd_nn_bwd_weight_scratch(int n, int h, int w, int cin, int cout, size_t *f32_elems);

/* The weights' gradient dw[cout][9][cin] (f32) and the bias's db[cout] (optional) from the
 * layer's input x and output gradient dz; sums in a fixed order. */
int dtd_nn_conv3x3_bwd_weight(dtd_ctx *ctx, int n, int h, int w, int cin, int cout,
                              const uint16_t *x, const uint16_t *dz, uint16_t *planes,
                              size_t planes_elems, float *part, size_t part_elems, float *dw,
                              float *db);

/* dz = dy where y > 0, else 0 (n elements). */
int dtd_nn_relu_bwd(dtd_ctx *ctx, size_t n, const uint16_t *y, const uint16_t *dy, uint16_t *dz);

/* y = 2x2 mean of x (h x w, even) into h/2 x w/2. */
int dtd_nn_pool(dtd_ctx *ctx, int n, int h, int w, int c, const uint16_t *x, uint16_t *y);

/* cat (h x w, cu + cs channels) = [nearest upsampling of u (h/2 x w/2, cu), skip (cs)]. */
int dtd_nn_up_cat(dtd_ctx *ctx, int n, int h, int w, int cu, int cs, const uint16_t *u,
                  const uint16_t *skip, uint16_t *cat);

/* Its backward pass: du = the upsampling's adjoint of dcat's first cu channels; dskip = dcat's
 * last cs channels plus 0.25 x the pooling's gradient dpool (h/2 x w/2, cs; NULL: none) at each
 * pixel's parent.  du or dskip may be NULL (that part is skipped). */
int dtd_nn_up_cat_bwd(dtd_ctx *ctx, int n, int h, int w, int cu, int cs, const uint16_t *dcat,
                      const uint16_t *dpool, uint16_t *du, uint16_t *dskip);

/* The network's input x (16 channels, 16-byte aligned) from f64 h, kd and the base guess
 * (optional): h / scale, ln kd, ln(dt kd / dx^2) / 4, base / scale (0 when base is NULL), zeros. */
int dtd_nn_input(dtd_ctx *ctx, size_t np, const double *h, const double *kd, const double *base,
                 double dt, double dx, double scale, uint16_t *x);

/* Mean squared error of the network's out[p outc] against (t - base) / scale (formed in f64) into
 * *loss, and its gradient dy (16 channels, channel 0 only); parts holds np / 256 + 1 floats. */
int dtd_nn_loss(dtd_ctx *ctx, size_t np, const float *out, int outc, const double *t,
                const double *base, double scale, float *parts, uint16_t *dy, double *loss);

/* The guess g = base + scale out[p outc] in f64. */
int dtd_nn_guess(dtd_ctx *ctx, size_t np, const double *base, const float *out, int outc,
                 double scale, double *g);

/* The splitting's first guess (as dtd_lod) of the device array h_dev into out_dev (device),
 * asynchronous on the context's stream. */
int dtd_lod_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x, const double *h_dev,
                   const double *kd_dev, double dt, double *out_dev);

/* One Adam step (Kingma & Ba 2015, step counted from 1) on f32 masters w with moments m, v and
 * gradient g; wb (optional) gets bf16(w). */
int dtd_nn_adam(dtd_ctx *ctx, size_t n, float *w, float *m, float *v, const float *g, float lr,
                float beta1, float beta2, float eps, int step, uint16_t *wb);

/* b = bf16(a). */
int dtd_nn_to_bf16(dtd_ctx *ctx, size_t n, const float *a, uint16_t *b);

/* Waits for the context's stream. */
int dtd_sync(dtd_ctx *ctx);

/* The learned coarse correction: the network runs at 1/f of the grid's resolution (f x f blocks)
 * and predicts the block means of the splitting's error; these two calls move fields between the
 * resolutions.  dtd_nn_block_mean: out (w/f x h/f) = each block's mean of in, its cells added in
 * row-major order; log_mean: exp of the mean of log(max(in, 1e-30)) (kd's geometric mean).
 * dtd_nn_guess_coarse: g = base + scale * (P out[:, 0]), P the bilinear interpolation between the
 * coarse cells' centres (periodic in x, constant past the outer rows' centres), out holding outc
 * f32 channels a coarse pixel.  w and h must be multiples of f; asynchronous on the stream. */
int dtd_nn_block_mean(dtd_ctx *ctx, size_t w, size_t h, int f, int log_mean, const double *in,
                      double *out);
int dtd_nn_guess_coarse(dtd_ctx *ctx, size_t w, size_t h, int f, const double *base,
                        const float *out, int outc, double scale, double *g);

/* dtd_solve_rel_device started from the guess x0_dev (device, w h f64) instead of the
 * splitting. */
int dtd_solve_rel_guess_device(dtd_ctx *ctx, size_t w, size_t h, double dx, int wrap_x,
                               double *h_dev, const double *kd_dev, double dt, double rel_tol,
                               const double *x0_dev, dtd_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
