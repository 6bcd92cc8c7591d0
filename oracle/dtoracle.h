/*
 * CPU oracle for the implicit diffusion solve: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of deep-time's CPU solve of (I + dt*L(K)) h1 = h0 on the x-periodic,
 * row-closed raster: the five-point operator, the locally one-dimensional first guess (Thomas
 * and Sherman-Morrison cyclic solves), the f32 multigrid V-cycle (aggregation levels, damped
 * Jacobi) and the f64 conjugate gradients it preconditions, with the residual's sum put back.
 * Each function cites the reference file:line it restates (paths relative to the reference
 * repository root).  Arithmetic follows the reference cell by cell and operation by operation
 * (built with -ffp-contract=off, no fast-math), and every reduction keeps the reference order:
 * a sequential sum per strip of 8 rows (or per row where the reference sums per row), then the
 * partial sums added in ascending order from -0.0.  OpenMP only schedules strips; no result
 * depends on the thread count.
 *
 * Only the test suite, the smoke check and the benchmark's CPU-baseline leg may load this
 * library.  The product (the HIP library) never calls it.
 *
 * Parity pins (tests/test_oracle.py): the closed-form eigen-solution for constant K, dense
 * Cholesky on small grids, the reference's own definitional oracles (apply_reference, plain CG,
 * bound), its known-answer ring test, the SplitMix64 published vectors, and invariants (A*1 = 1,
 * symmetry, mass, an SPD cycle).  Functions without such a pin are marked "parity unpinned".
 */
#ifndef DTORACLE_H
#define DTORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fixtures ------------------------------------------------------------------------------ */
uint64_t or_splitmix_next(uint64_t *state);
double or_splitmix_f64(uint64_t *state);
void or_terrain(size_t w, size_t h, uint64_t seed, double *z);
double or_submarine_kd(double z, double d0, double wave_base, double tidal_range, double land);
void or_shelf_kd(size_t n, const double *z, uint64_t seed, double *kd);
void or_fractal(size_t w, size_t h, uint64_t seed, double *z);

/* ---- operator ------------------------------------------------------------------------------ */
void or_operator(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                 double *east, double *south, double *diag);
void or_apply(size_t w, size_t h, int wrap_x, const double *east, const double *south,
              const double *diag, const double *x, double *y);
void or_apply_reference(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                        const double *x, double *y);

/* ---- first guess --------------------------------------------------------------------------- */
void or_thomas(size_t n, const double *lo, const double *di, const double *up, double *b);
void or_cyclic(size_t n, const double *lo, const double *di, const double *up, double alpha,
               double beta, double *b);
void or_lod(size_t w, size_t h, double dx, int wrap_x, double *hgt, const double *kd, double dt);

/* ---- multigrid cycle ----------------------------------------------------------------------- */
typedef struct or_hier or_hier;
or_hier *or_hier_new(size_t w, size_t h, int wrap, const double *east, const double *south,
                     const double *diag, float omega, size_t nu, float scale);
void or_hier_free(or_hier *hy);
size_t or_hier_levels(const or_hier *hy);
/* Level l's shape and its operator arrays (NULL pointers are skipped; mass is 1 on level 0). */
void or_hier_level(const or_hier *hy, size_t l, size_t *w, size_t *h, int *wrap, float *east,
                   float *south, float *diag, float *wd, float *mass);
/* z = M^-1 r for r32 on the finest level; returns sum r64 * z (reference order). */
double or_hier_precondition(or_hier *hy, const float *r32, const double *r64, float *z);
/* The cycle's options (no counterpart in the reference): transfer 1 replaces 2x2 aggregation by
 * the cell-centred bilinear prolongation P and its transpose as the restriction; coarsest 1
 * solves the coarsest level exactly (Cholesky in f64) instead of COARSE_SWEEPS sweeps. */
or_hier *or_hier_new_opts(size_t w, size_t h, int wrap, const double *east, const double *south,
                          const double *diag, float omega, size_t nu, float scale, int transfer,
                          int coarsest);
/* The bilinear transfers alone: out = P zc (fine wf x hf, coarse wc x hc, cwrap: the coarse
 * level wraps) and rc = P^T res (wrapf: the fine level wraps). */
void or_bil_prolong(size_t wf, size_t hf, size_t wc, size_t hc, int cwrap, const float *zc,
                    float *out);
void or_bil_restrict(size_t wf, size_t hf, int wrapf, size_t wc, size_t hc, int cwrap,
                     const float *res, float *rc);
/* The exact coarsest solve alone: z = A^-1 r for a level's f32 operator, in f64. */
void or_coarse_exact(size_t w, size_t h, int wrap, const float *east, const float *south,
                     const float *diag, const float *r, float *z);

/* ---- solves -------------------------------------------------------------------------------- */
size_t or_cg_on(size_t w, size_t h, double dx, int wrap_x, double *x, const double *b,
                const double *kd, double dt, double tol_m, float omega, size_t nu, float scale);
size_t or_solve(size_t w, size_t h, double dx, int wrap_x, double *hgt, const double *kd,
                double dt, double tol_m);
size_t or_jacobi_cg(size_t w, size_t h, double dx, int wrap_x, double *x, const double *b,
                    const double *kd, double dt, double tol_m);
int or_live_of(size_t w, size_t h, double dx, const double *hgt, const double *kd, double dt,
               unsigned char *strip_on);
size_t or_step_in(size_t w, size_t h, double dx, int wrap_x, double *hgt, const double *kd,
                  double dt);
double or_spread_diffusivity(double z, double strength, double d0);
size_t or_reference_cg(size_t w, size_t h, double dx, int wrap_x, const double *b,
                       const double *kd, double dt, double *x);
double or_bound(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                double tol);
double or_rel_residual(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                       const double *b, const double *x);
int or_threads(void);

/* ---- submarine slope failure (MAR:254-655) ---------------------------------------------------- */
long or_d8_neighbour(size_t w, size_t h, int wrap_x, size_t c, int k);
size_t or_cube_solve(size_t n, const uint32_t *slots, const double *lap, const double *area,
                     double mean_area, double *hgt, const double *kd, double dt, int solver,
                     double tol_m);
double or_slope_failure(size_t w, size_t h, double dx, int wrap_x, const unsigned char *sea,
                        double sea_level, double critical_slope, double rate, unsigned iterations,
                        double *hgt, double *avail);

#ifdef __cplusplus
}
#endif
#endif
