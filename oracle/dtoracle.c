/*
 * CPU oracle for the implicit diffusion solve: TEST INFRASTRUCTURE ONLY (see dtoracle.h).
 *
 * Reference files (relative to the reference repository root):
 *   D   crates/deeptime-core/src/surface/diffusion.rs
 *   MG  crates/deeptime-core/src/surface/diffusion/multigrid.rs
 *   MGT crates/deeptime-core/src/surface/diffusion/multigrid/tests.rs
 *   PAR crates/deeptime-core/src/surface/diffusion/parity.rs
 *   SM  crates/deeptime-core/src/sphere/mod.rs
 *   MAR crates/deeptime-core/src/surface/marine.rs
 *   GRID crates/deeptime-core/src/surface/grid.rs
 *   CD  crates/deeptime-core/src/surface/cube/diffusion.rs
 *   CM  crates/deeptime-core/src/surface/cube/multigrid.rs
 *   CUBE crates/deeptime-core/src/surface/cube.rs
 *   PROJ crates/deeptime-format/src/projection.rs
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fno-fast-math (oracle/Makefile).  With SSE2
 * arithmetic (x86-64) float and double operations round exactly once each, so the cell-by-cell
 * arithmetic is the reference's bit for bit; the reductions keep the reference's order.
 */
#include "dtoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Rows per strip of the reference's fused passes (D:502); every strip sum is taken over its
 * rows in row-major order and the strip sums are then added in ascending strip order. */
#define STRIP_ROWS 8
/* CG iteration cap (D:497). */
#define CG_MAX_ITER 400
/* The Jacobi solve's RMS tolerance, metres (D:495). */
#define CG_TOL_M 1.0e-6
/* Cycle constants (MG:45-56). */
#define OMEGA 0.8f
#define NU 2
#define SCALE 0.5f
#define COARSEST 64
#define COARSE_SWEEPS 16

int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* The value an empty f64 sum starts from in the reference (D:507-510): -0.0. */
static inline double sum_zero(void) { return -0.0; }

static size_t div_ceil(size_t a, size_t b) { return (a + b - 1) / b; }

/* ---- fixtures ------------------------------------------------------------------------------ */

/* SplitMix64::next_u64, SM:52-58 (Steele, Lea & Flood 2014). */
uint64_t or_splitmix_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* SplitMix64::next_f64, SM:61-63: 53 random bits scaled into [0, 1). */
double or_splitmix_f64(uint64_t *state) {
    return (double)(or_splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* Rust's f64::fract: x - trunc(x). */
static inline double fract(double x) { return x - trunc(x); }

/* `terrain`, MGT:77-98: six octaves of bilinear value noise, metres. */
void or_terrain(size_t w, size_t h, uint64_t seed, double *z) {
    uint64_t st = seed;
    memset(z, 0, w * h * sizeof(double));
    for (int oct = 0; oct < 6; oct++) {
        size_t g = (size_t)2 << oct;
        double *grid = (double *)malloc(g * g * sizeof(double));
        for (size_t k = 0; k < g * g; k++) grid[k] = or_splitmix_f64(&st) - 0.5;
        double amp = 3000.0 / (double)(1 << oct);
        size_t hmax = h > 1 ? h : 1;
        for (size_t j = 0; j < h; j++) {
            double fy = (double)j / (double)hmax * (double)(g - 1);
            size_t j0 = (size_t)floor(fy);
            double ty = fract(fy);
            for (size_t i = 0; i < w; i++) {
                double fx = (double)i / (double)w * (double)g;
                size_t i0 = (size_t)floor(fx);
                double tx = fract(fx);
#define AT(a, b) grid[((b) < g - 1 ? (b) : g - 1) * g + (a) % g]
                double top = AT(i0, j0) * (1.0 - tx) + AT(i0 + 1, j0) * tx;
                double bot = AT(i0, j0 + 1) * (1.0 - tx) + AT(i0 + 1, j0 + 1) * tx;
#undef AT
                z[j * w + i] += amp * (top * (1.0 - ty) + bot * ty);
            }
        }
        free(grid);
    }
}

/* `fractal`, PAR:340-359: eight octaves of value noise on a 2:1 lattice. */
void or_fractal(size_t w, size_t h, uint64_t seed, double *z) {
    uint64_t st = seed;
    memset(z, 0, w * h * sizeof(double));
    for (int oct = 0; oct < 8; oct++) {
        size_t gw = (size_t)4 << oct, gh = (size_t)2 << oct;
        double *g = (double *)malloc(gw * gh * sizeof(double));
        for (size_t k = 0; k < gw * gh; k++) g[k] = or_splitmix_f64(&st) - 0.5;
        double amp = 4000.0 / pow((double)(1 << oct), 0.9);
        for (size_t j = 0; j < h; j++) {
            double fy = (double)j / (double)h * (double)gh;
            size_t j0 = (size_t)floor(fy);
            double ty = fract(fy);
            for (size_t i = 0; i < w; i++) {
                double fx = (double)i / (double)w * (double)gw;
                size_t i0 = (size_t)floor(fx);
                double tx = fract(fx);
#define AT(a, b) g[((b) < gh - 1 ? (b) : gh - 1) * gw + (a) % gw]
                double top = AT(i0, j0) * (1.0 - tx) + AT(i0 + 1, j0) * tx;
                double bot = AT(i0, j0 + 1) * (1.0 - tx) + AT(i0 + 1, j0 + 1) * tx;
#undef AT
                z[j * w + i] += amp * (top * (1.0 - ty) + bot * ty);
            }
        }
        free(g);
    }
}

/* `submarine_kd`, MAR:44-59: depth-dependent marine diffusivity blended to `land` ashore. */
double or_submarine_kd(double z, double d0, double wave_base, double tidal_range, double land) {
    double f1;
    if (tidal_range > 0.0)
        f1 = (tanh(z / tidal_range) + 1.0) / 2.0;
    else
        f1 = z >= 0.0 ? 1.0 : 0.0;
    double f2 = (z > wave_base && wave_base > 0.0) ? exp(-(z - wave_base) / wave_base) : 1.0;
    return d0 * f1 * f2 + land * (1.0 - f1);
}

/* `shelf_kd`, MGT:102-110. */
void or_shelf_kd(size_t n, const double *z, uint64_t seed, double *kd) {
    uint64_t st = seed ^ 0x5eedull;
    for (size_t c = 0; c < n; c++) {
        double land = 2.0 * (0.25 + 4.0 * or_splitmix_f64(&st));
        kd[c] = or_submarine_kd(-z[c], 100.0, 60.0, 2.0, land);
    }
}

/* ---- operator ------------------------------------------------------------------------------ */

/* `Operator::new`, D:532-603 (every row visited): east/south face couplings
 * 0.5*(K_i + K_j)*dt/dx^2 and diag = ((((1 + e) + w) + s) + n). */
void or_operator(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                 double *east, double *south, double *diag) {
    double s2 = dx * dx;
    int wrap = wrap_x && w > 1;
#pragma omp parallel for schedule(static)
    for (long jl = 0; jl < (long)h; jl++) {
        size_t j = (size_t)jl;
        const double *krow = kd + j * w;
        double *erow = east + j * w, *srow = south + j * w;
        for (size_t i = 0; i < w; i++) {
            if (i + 1 < w)
                erow[i] = 0.5 * (krow[i] + krow[i + 1]) * dt / s2;
            else if (wrap)
                erow[i] = 0.5 * (krow[i] + krow[0]) * dt / s2;
            else
                erow[i] = 0.0;
        }
        if (j + 1 < h) {
            const double *below = kd + (j + 1) * w;
            for (size_t i = 0; i < w; i++) srow[i] = 0.5 * (krow[i] + below[i]) * dt / s2;
        } else {
            for (size_t i = 0; i < w; i++) srow[i] = 0.0;
        }
    }
#pragma omp parallel for schedule(static)
    for (long jl = 0; jl < (long)h; jl++) {
        size_t j = (size_t)jl, c0 = j * w;
        for (size_t i = 0; i < w; i++) {
            double west = i > 0 ? east[c0 + i - 1] : (wrap ? east[c0 + w - 1] : 0.0);
            double north = j > 0 ? south[c0 - w + i] : 0.0;
            diag[c0 + i] = 1.0 + east[c0 + i] + west + south[c0 + i] + north;
        }
    }
}

/* One cell of y = A x in the order of `Operator::apply_row`, D:609-649: diagonal, then the east
 * neighbour (the wrapped one for the last cell), then the west neighbour (the wrapped one for the
 * first cell), then south, then north, each product subtracted from the running value. */
static inline double op_cell(size_t w, size_t h, int wrap, const double *east,
                             const double *south, const double *diag, const double *x, size_t j,
                             size_t i) {
    size_t c = j * w + i, c0 = j * w;
    double o = diag[c] * x[c];
    if (i + 1 < w)
        o -= east[c] * x[c + 1];
    else if (wrap)
        o -= east[c] * x[c0];
    if (i > 0)
        o -= east[c - 1] * x[c - 1];
    else if (wrap)
        o -= east[c0 + w - 1] * x[c0 + w - 1];
    if (j + 1 < h) o -= south[c] * x[c + w];
    if (j > 0) o -= south[c - w] * x[c - w];
    return o;
}

/* `Operator::apply`, D:652-663. */
void or_apply(size_t w, size_t h, int wrap_x, const double *east, const double *south,
              const double *diag, const double *x, double *y) {
    int wrap = wrap_x && w > 1;
#pragma omp parallel for schedule(static)
    for (long jl = 0; jl < (long)h; jl++)
        for (size_t i = 0; i < w; i++)
            y[(size_t)jl * w + i] = op_cell(w, h, wrap, east, south, diag, x, (size_t)jl, i);
}

/* `apply_reference`, MGT:17-43: (I + dt L) x from the definition, cell by cell. */
void or_apply_reference(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                        const double *x, double *y) {
    double s2 = dx * dx;
    int wrap = wrap_x && w > 1;
    for (size_t j = 0; j < h; j++) {
        for (size_t i = 0; i < w; i++) {
            size_t c = j * w + i;
            long nb[4];
            nb[0] = i + 1 < w ? (long)(c + 1) : (wrap ? (long)(j * w) : -1);
            nb[1] = i > 0 ? (long)(c - 1) : (wrap ? (long)(j * w + w - 1) : -1);
            nb[2] = j + 1 < h ? (long)(c + w) : -1;
            nb[3] = j > 0 ? (long)(c - w) : -1;
            double v = x[c];
            for (int k = 0; k < 4; k++) {
                if (nb[k] < 0) continue;
                size_t m = (size_t)nb[k];
                v += 0.5 * (kd[c] + kd[m]) * dt / s2 * (x[c] - x[m]);
            }
            y[c] = v;
        }
    }
}

/* ---- first guess --------------------------------------------------------------------------- */

/* `thomas`, D:1029-1046. */
void or_thomas(size_t n, const double *lo, const double *di, const double *up, double *b) {
    if (n == 0) return;
    double *c = (double *)calloc(n, sizeof(double));
    double beta = di[0];
    b[0] /= beta;
    for (size_t i = 1; i < n; i++) {
        c[i] = up[i - 1] / beta;
        beta = di[i] - lo[i] * c[i];
        b[i] = (b[i] - lo[i] * b[i - 1]) / beta;
    }
    for (size_t i = n - 1; i-- > 0;) {
        double next = b[i + 1];
        b[i] -= c[i + 1] * next;
    }
    free(c);
}

/* `cyclic`, D:1050-1071: Sherman-Morrison over two Thomas solves; `alpha` couples row n-1 to
 * column 0, `beta` row 0 to column n-1. */
void or_cyclic(size_t n, const double *lo, const double *di, const double *up, double alpha,
               double beta, double *b) {
    double gamma = -di[0];
    double *bb = (double *)malloc(n * sizeof(double));
    double *l = (double *)malloc(n * sizeof(double));
    double *u = (double *)malloc(n * sizeof(double));
    double *x = (double *)malloc(n * sizeof(double));
    double *z = (double *)calloc(n, sizeof(double));
    memcpy(bb, di, n * sizeof(double));
    bb[0] = di[0] - gamma;
    bb[n - 1] = di[n - 1] - alpha * beta / gamma;
    memcpy(l, lo, n * sizeof(double));
    memcpy(u, up, n * sizeof(double));
    l[0] = 0.0;
    u[n - 1] = 0.0;
    memcpy(x, b, n * sizeof(double));
    or_thomas(n, l, bb, u, x);
    z[0] = gamma;
    z[n - 1] = alpha;
    or_thomas(n, l, bb, u, z);
    double fact = (x[0] + beta * x[n - 1] / gamma) / (1.0 + z[0] + beta * z[n - 1] / gamma);
    for (size_t i = 0; i < n; i++) b[i] = x[i] - fact * z[i];
    free(bb);
    free(l);
    free(u);
    free(x);
    free(z);
}

/* `lod`, D:370-489: an implicit sweep along x (a cyclic tridiagonal system per row when the grid
 * wraps and is wider than two cells), then along the columns.  row_on (NULL: every row) marks the
 * visited rows (`Live`): a skipped row is left as it is by the x sweep, and each column is solved
 * one run of visited rows at a time (D:430-488). */
static void lod_live(size_t w, size_t h, double dx, int wrap_x, double *hgt, const double *kd,
                     double dt, const unsigned char *row_on) {
    double s2 = dx * dx;
    if (w == 0 || h == 0) return;
#pragma omp parallel
    {
        double *lo = (double *)malloc(w * sizeof(double));
        double *di = (double *)malloc(w * sizeof(double));
        double *up = (double *)malloc(w * sizeof(double));
#pragma omp for schedule(static)
        for (long jl = 0; jl < (long)h; jl++) {
            size_t j = (size_t)jl;
            if (row_on && !row_on[j]) continue;
            const double *krow = kd + j * w;
            double *row = hgt + j * w;
#define FACE(i)                                                                                \
    ((i) + 1 < w ? 0.5 * (krow[(i)] + krow[(i) + 1]) * dt / s2                                  \
                 : ((wrap_x && w > 1) ? 0.5 * (krow[w - 1] + krow[0]) * dt / s2 : 0.0))
            for (size_t i = 0; i < w; i++) {
                double west = i == 0 ? FACE(w - 1) : FACE(i - 1);
                double east = FACE(i);
                lo[i] = -west;
                up[i] = -east;
                di[i] = 1.0 + west + east;
            }
            if (wrap_x && w > 2) {
                double corner_lo = lo[0], corner_up = up[w - 1];
                or_cyclic(w, lo, di, up, corner_up, corner_lo, row);
            } else {
                if (w == 2 && wrap_x) {
                    double f = FACE(0) + FACE(1);
                    lo[1] = -f;
                    up[0] = -f;
                    di[0] = 1.0 + f;
                    di[1] = 1.0 + f;
                }
                lo[0] = 0.0;
                up[w - 1] = 0.0;
                if (w == 1) di[0] = 1.0;
                or_thomas(w, lo, di, up, row);
            }
#undef FACE
        }
        free(lo);
        free(di);
        free(up);
    }
#pragma omp parallel
    {
        double *col = (double *)malloc(h * sizeof(double));
        double *lo = (double *)malloc(h * sizeof(double));
        double *di = (double *)malloc(h * sizeof(double));
        double *up = (double *)malloc(h * sizeof(double));
#pragma omp for schedule(static)
        for (long il = 0; il < (long)w; il++) {
            size_t i = (size_t)il;
#define FACEZ(j) ((j) + 1 < h ? 0.5 * (kd[(j) * w + i] + kd[((j) + 1) * w + i]) * dt / s2 : 0.0)
            /* The runs of visited rows, [a, e), in order. */
            for (size_t a = 0; a < h;) {
                if (row_on && !row_on[a]) {
                    a++;
                    continue;
                }
                size_t e = a + 1;
                while (e < h && (!row_on || row_on[e])) e++;
                size_t len = e - a;
                for (size_t jj = 0; jj < len; jj++) {
                    size_t j = a + jj;
                    col[jj] = hgt[j * w + i];
                    double north = j == 0 ? 0.0 : FACEZ(j - 1);
                    double south = FACEZ(j);
                    lo[jj] = -north;
                    up[jj] = -south;
                    di[jj] = 1.0 + north + south;
                }
                or_thomas(len, lo, di, up, col);
                for (size_t jj = 0; jj < len; jj++) hgt[(a + jj) * w + i] = col[jj];
                a = e;
            }
#undef FACEZ
        }
        free(col);
        free(lo);
        free(di);
        free(up);
    }
}

void or_lod(size_t w, size_t h, double dx, int wrap_x, double *hgt, const double *kd, double dt) {
    lod_live(w, h, dx, wrap_x, hgt, kd, dt, NULL);
}

/* ---- multigrid cycle (MG:73-523) ----------------------------------------------------------- */

typedef struct {
    size_t w, h;
    int wrap;
    float *east, *south, *diag, *wd, *mass; /* mass == NULL on the finest level (1 a cell) */
    float *r, *z, *t, *zp;
} lvl;

struct or_hier {
    size_t n;
    lvl *lv;
    float omega;
    size_t nu;
    float scale;
    int bil;       /* bilinear prolongation and its transpose as restriction (an option) */
    double *chol;  /* the coarsest level's Cholesky factor, for its exact f64 solve (an option) */
};

static float *fzalloc(size_t n) { return (float *)calloc(n ? n : 1, sizeof(float)); }

/* `Level::finest`, MG:96-114: f32 copies of the operator, wd = omega / diag. */
static void level_finest(lvl *L, size_t w, size_t h, int wrap, const double *east,
                         const double *south, const double *diag, float omega) {
    size_t n = w * h;
    L->w = w;
    L->h = h;
    L->wrap = wrap;
    L->east = fzalloc(n);
    L->south = fzalloc(n);
    L->diag = fzalloc(n);
    L->wd = fzalloc(n);
    L->mass = NULL;
    L->r = fzalloc(n);
    L->z = fzalloc(n);
    L->t = fzalloc(n);
    L->zp = fzalloc(n);
    for (size_t c = 0; c < n; c++) {
        L->east[c] = (float)east[c];
        L->south[c] = (float)south[c];
        L->diag[c] = (float)diag[c];
        L->wd[c] = omega / L->diag[c];
    }
}

/* `Level::coarsen`, MG:117-199: 2x2 aggregation (a lone last row/column when odd); mass summed,
 * coarse faces = scale * the fine faces across them; returns 0 when this level is the coarsest. */
static int level_coarsen(const lvl *F, lvl *C, float omega, float scale) {
    size_t w = F->w, h = F->h;
    size_t wc = div_ceil(w, 2), hc = div_ceil(h, 2);
    if (w * h <= COARSEST || (wc == w && hc == h)) return 0;
    size_t nc = wc * hc;
    int wrap = F->wrap && wc > 1;
    C->w = wc;
    C->h = hc;
    C->wrap = wrap;
    C->east = fzalloc(nc);
    C->south = fzalloc(nc);
    C->diag = fzalloc(nc);
    C->wd = fzalloc(nc);
    C->mass = fzalloc(nc);
    C->r = fzalloc(nc);
    C->z = fzalloc(nc);
    C->t = fzalloc(nc);
    C->zp = fzalloc(nc);
    for (size_t jc = 0; jc < hc; jc++) {
        size_t r0 = 2 * jc, r1 = 2 * jc + 2 < h ? 2 * jc + 2 : h;
        size_t last_row = r1 - 1;
        for (size_t ic = 0; ic < wc; ic++) {
            size_t c0 = 2 * ic, c1 = 2 * ic + 2 < w ? 2 * ic + 2 : w;
            size_t last_col = c1 - 1;
            float m = 0.0f;
            for (size_t j = r0; j < r1; j++)
                for (size_t i = c0; i < c1; i++) m += F->mass ? F->mass[j * w + i] : 1.0f;
            C->mass[jc * wc + ic] = m;
            if (ic + 1 < wc || wrap) {
                float e = 0.0f;
                for (size_t j = r0; j < r1; j++) e += F->east[j * w + last_col];
                C->east[jc * wc + ic] = scale * e;
            }
            if (jc + 1 < hc) {
                float s = 0.0f;
                for (size_t i = c0; i < c1; i++) s += F->south[last_row * w + i];
                C->south[jc * wc + ic] = scale * s;
            }
        }
    }
    for (size_t j = 0; j < hc; j++) {
        size_t c0 = j * wc;
        for (size_t i = 0; i < wc; i++) {
            float west = i > 0 ? C->east[c0 + i - 1] : (wrap ? C->east[c0 + wc - 1] : 0.0f);
            float north = j > 0 ? C->south[c0 - wc + i] : 0.0f;
            C->diag[c0 + i] = C->mass[c0 + i] + C->east[c0 + i] + west + C->south[c0 + i] + north;
        }
    }
    for (size_t c = 0; c < nc; c++) C->wd[c] = omega / C->diag[c];
    return 1;
}

/* One cell of A x on a level, in the order of `Level::apply_row`, MG:203-240: diagonal, east
 * (non-wrapped), west (non-wrapped), then the wrapped east term of the last cell and the wrapped
 * west term of the first, then south, then north. */
static inline float lv_cell(const lvl *L, const float *x, size_t j, size_t i) {
    size_t w = L->w, c = j * w + i, c0 = j * w;
    float o = L->diag[c] * x[c];
    if (i + 1 < w) o -= L->east[c] * x[c + 1];
    if (i > 0) o -= L->east[c - 1] * x[c - 1];
    if (L->wrap) {
        if (i == w - 1) o -= L->east[c0 + w - 1] * x[c0];
        if (i == 0) o -= L->east[c0 + w - 1] * x[c0 + w - 1];
    }
    if (j + 1 < L->h) o -= L->south[c] * x[c + w];
    if (j > 0) o -= L->south[c - w] * x[c - w];
    return o;
}

/* `Level::first_sweep`, MG:243-256: z = wd * r. */
static void first_sweep(lvl *L) {
    size_t n = L->w * L->h;
#pragma omp parallel for schedule(static) if (n > 65536)
    for (long c = 0; c < (long)n; c++) L->z[c] = L->wd[c] * L->r[c];
}

/* `Level::sweep`, MG:260-329: z <- z' + wd (r - A z') with z' = z + P zc when `zc` is given
 * (piecewise-constant prolongation); with `dot`, returns sum dot * z_new, summed per strip of
 * STRIP_ROWS rows in row-major order and the strips added in order. */
static double sweep(lvl *L, const float *zc, size_t wc, const double *dot) {
    size_t w = L->w, h = L->h, n = w * h;
    const float *zp = L->z;
    if (zc) {
        for (size_t j = 0; j < h; j++)
            for (size_t i = 0; i < w; i++)
                L->zp[j * w + i] = L->z[j * w + i] + zc[(j / 2) * wc + i / 2];
        zp = L->zp;
    }
    size_t ns = div_ceil(h, STRIP_ROWS);
    double *acc = (double *)malloc((ns ? ns : 1) * sizeof(double));
#pragma omp parallel for schedule(static) if (n > 65536)
    for (long sl = 0; sl < (long)ns; sl++) {
        size_t s = (size_t)sl, j0 = s * STRIP_ROWS;
        size_t j1 = j0 + STRIP_ROWS < h ? j0 + STRIP_ROWS : h;
        double a = sum_zero();
        for (size_t j = j0; j < j1; j++) {
            for (size_t i = 0; i < w; i++) {
                size_t c = j * w + i;
                float ax = lv_cell(L, zp, j, i);
                L->t[c] = zp[c] + L->wd[c] * (L->r[c] - ax);
            }
            if (dot)
                for (size_t i = 0; i < w; i++) a += dot[j * w + i] * (double)L->t[j * w + i];
        }
        acc[s] = a;
    }
    double total = sum_zero();
    for (size_t s = 0; s < ns; s++) total += acc[s];
    free(acc);
    float *tmp = L->z;
    L->z = L->t;
    L->t = tmp;
    return total;
}

/* `Level::restrict_into`, MG:332-356: rc = sum over the 2x2 block of (r - A z), fine row j0 then
 * j0 + 1, columns ascending, from +0. */
static void restrict_into(const lvl *F, lvl *C) {
    size_t w = F->w, h = F->h, wc = C->w, hc = C->h;
#pragma omp parallel for schedule(static) if (w * h > 65536)
    for (long jcl = 0; jcl < (long)hc; jcl++) {
        size_t jc = (size_t)jcl;
        float *rc = C->r + jc * wc;
        for (size_t ic = 0; ic < wc; ic++) rc[ic] = 0.0f;
        size_t j1 = 2 * jc + 2 < h ? 2 * jc + 2 : h;
        for (size_t j = 2 * jc; j < j1; j++)
            for (size_t i = 0; i < w; i++) rc[i / 2] += F->r[j * w + i] - lv_cell(F, F->z, j, i);
    }
}

/* ---- options of the cycle with no counterpart in the reference ------------------------------
 * Bilinear transfers: the cell-centred bilinear prolongation P and its transpose as the
 * restriction (full weighting), in place of 2x2 aggregation; the coarse levels' operators stay
 * the aggregation ones.  Exact coarsest solve: the coarsest level (at most 64 cells) solved by a
 * Cholesky factorisation in f64 in place of COARSE_SWEEPS damped-Jacobi sweeps.  The device
 * kernels do the same operations in the same order. */

/* Along one axis, fine index i lies in coarse cell *c = i / 2 and nearest to coarse cell
 * *nb = c + 1 (i odd) or c - 1 (i even); *nb wraps on a periodic axis of nc cells and otherwise
 * falls back to c. */
static inline void bil_axis(long i, long nc, int wrap, long *c, long *nb) {
    long cc = i / 2, m = (i & 1) ? cc + 1 : cc - 1;
    if (wrap)
        m = (m + nc) % nc;
    else if (m < 0 || m >= nc)
        m = cc;
    *c = cc;
    *nb = m;
}

/* P's weight of coarse index C for fine index i along one axis: 3/4 for the cell i lies in, 1/4
 * for its nearest neighbour, summed when they coincide. */
static inline float bil_weight(long i, long C, long nc, int wrap) {
    long c, nb;
    bil_axis(i, nc, wrap, &c, &nb);
    return (c == C ? 0.75f : 0.0f) + (nb == C ? 0.25f : 0.0f);
}

/* (P zc) at fine cell (i, j): 3/4 (3/4 zc[cy][cx] + 1/4 zc[cy][nx]) + 1/4 (3/4 zc[ny][cx] +
 * 1/4 zc[ny][nx]); cwrap: the coarse level wraps in x. */
static inline float bil_at(const float *zc, size_t wc, size_t hc, int cwrap, size_t i, size_t j) {
    long cx, nx, cy, ny;
    bil_axis((long)i, (long)wc, cwrap, &cx, &nx);
    bil_axis((long)j, (long)hc, 0, &cy, &ny);
    float a = 0.75f * zc[cy * (long)wc + cx] + 0.25f * zc[cy * (long)wc + nx];
    float b = 0.75f * zc[ny * (long)wc + cx] + 0.25f * zc[ny * (long)wc + nx];
    return 0.75f * a + 0.25f * b;
}

/* (P^T res) at coarse cell (I, J): over the fine rows 2J-1 .. 2J+2 on the level and the fine
 * columns 2I-1 .. 2I+2 (wrapped when the fine level wraps, each column once), the terms
 * (weight_y * weight_x) * res with a non-zero weight, added from +0, rows outer. */
static inline float bil_restrict_at(const float *res, size_t w, size_t h, int wrapf, size_t wc,
                                    size_t hc, int cwrap, size_t I, size_t J) {
    float acc = 0.0f;
    for (long t = -1; t <= 2; t++) {
        long j = 2 * (long)J + t;
        if (j < 0 || j >= (long)h) continue;
        float wy = bil_weight(j, (long)J, (long)hc, 0);
        if (wy == 0.0f) continue;
        long seen[4];
        int ns = 0;
        for (long u = -1; u <= 2; u++) {
            long i = 2 * (long)I + u;
            if (wrapf)
                i = ((i % (long)w) + (long)w) % (long)w;
            else if (i < 0 || i >= (long)w)
                continue;
            int dup = 0;
            for (int q = 0; q < ns; q++) dup |= seen[q] == i;
            if (dup) continue;
            seen[ns++] = i;
            float wx = bil_weight(i, (long)I, (long)wc, cwrap);
            if (wx == 0.0f) continue;
            acc += (wy * wx) * res[j * (long)w + i];
        }
    }
    return acc;
}

void or_bil_prolong(size_t wf, size_t hf, size_t wc, size_t hc, int cwrap, const float *zc,
                    float *out) {
    for (size_t j = 0; j < hf; j++)
        for (size_t i = 0; i < wf; i++) out[j * wf + i] = bil_at(zc, wc, hc, cwrap, i, j);
}

void or_bil_restrict(size_t wf, size_t hf, int wrapf, size_t wc, size_t hc, int cwrap,
                     const float *res, float *rc) {
    for (size_t J = 0; J < hc; J++)
        for (size_t I = 0; I < wc; I++)
            rc[J * wc + I] = bil_restrict_at(res, wf, hf, wrapf, wc, hc, cwrap, I, J);
}

/* The coarsest level's matrix in f64 (row c: the terms of lv_cell in its order, accumulated from
 * 0), factorised in place: A = L L^T with L in the lower triangle (n x n, row-major). */
static void coarse_factor(size_t w, size_t h, int wrap, const float *east, const float *south,
                          const float *diag, double *A) {
    size_t n = w * h;
    memset(A, 0, n * n * sizeof(double));
    for (size_t j = 0; j < h; j++)
        for (size_t i = 0; i < w; i++) {
            size_t c = j * w + i, c0 = j * w;
            double *row = A + c * n;
            row[c] += (double)diag[c];
            if (i + 1 < w) row[c + 1] -= (double)east[c];
            if (i > 0) row[c - 1] -= (double)east[c - 1];
            if (wrap) {
                if (i == w - 1) row[c0] -= (double)east[c0 + w - 1];
                if (i == 0) row[c0 + w - 1] -= (double)east[c0 + w - 1];
            }
            if (j + 1 < h) row[c + w] -= (double)south[c];
            if (j > 0) row[c - w] -= (double)south[c - w];
        }
    for (size_t k = 0; k < n; k++) {
        A[k * n + k] = sqrt(A[k * n + k]);
        for (size_t i = k + 1; i < n; i++) A[i * n + k] = A[i * n + k] / A[k * n + k];
        for (size_t i = k + 1; i < n; i++)
            for (size_t j = k + 1; j <= i; j++) A[i * n + j] -= A[i * n + k] * A[j * n + k];
    }
}

/* z = (L L^T)^-1 r in f64 (forward, then backward substitution by columns), rounded to f32. */
static void coarse_solve(const double *L, size_t n, const float *r, float *z) {
    double y[64];
    for (size_t c = 0; c < n; c++) y[c] = (double)r[c];
    for (size_t k = 0; k < n; k++) {
        y[k] = y[k] / L[k * n + k];
        for (size_t i = k + 1; i < n; i++) y[i] -= L[i * n + k] * y[k];
    }
    for (size_t kk = n; kk-- > 0;) {
        y[kk] = y[kk] / L[kk * n + kk];
        for (size_t i = 0; i < kk; i++) y[i] -= L[kk * n + i] * y[kk];
    }
    for (size_t c = 0; c < n; c++) z[c] = (float)y[c];
}

void or_coarse_exact(size_t w, size_t h, int wrap, const float *east, const float *south,
                     const float *diag, const float *r, float *z) {
    size_t n = w * h;
    double *A = (double *)malloc((n > 0 ? n * n : 1) * sizeof(double));
    coarse_factor(w, h, wrap, east, south, diag, A);
    coarse_solve(A, n, r, z);
    free(A);
}

/* sum dot * z per strip of STRIP_ROWS rows from -0.0, the strips added in order. */
static double strip_dot(size_t w, size_t h, const double *dot, const float *z) {
    double total = sum_zero();
    for (size_t j0 = 0; j0 < h; j0 += STRIP_ROWS) {
        double a = sum_zero();
        size_t j1 = j0 + STRIP_ROWS < h ? j0 + STRIP_ROWS : h;
        for (size_t c = j0 * w; c < j1 * w; c++) a += dot[c] * (double)z[c];
        total += a;
    }
    return total;
}

/* `vcycle`, MG:500-523: from z = 0; on the coarsest level a first sweep and COARSE_SWEEPS - 1
 * sweeps; otherwise nu sweeps, restriction, the coarse cycle, the corrected sweep and nu - 1
 * sweeps.  With `dot`, returns sum dot * z of the last sweep. */
static double vcycle(or_hier *hy, size_t l, const double *dot) {
    lvl *L = &hy->lv[l];
    if (l + 1 == hy->n) {
        if (hy->chol) {
            coarse_solve(hy->chol, L->w * L->h, L->r, L->z);
            return dot ? strip_dot(L->w, L->h, dot, L->z) : 0.0;
        }
        first_sweep(L);
        double acc = 0.0;
        size_t ks = COARSE_SWEEPS > 2 ? COARSE_SWEEPS : 2;
        for (size_t k = 1; k < ks; k++) acc = sweep(L, NULL, 0, k + 1 == ks ? dot : NULL);
        return acc;
    }
    first_sweep(L);
    for (size_t k = 1; k < hy->nu; k++) sweep(L, NULL, 0, NULL);
    lvl *C = &hy->lv[l + 1];
    if (hy->bil) {
        size_t w = L->w, h = L->h;
        for (size_t j = 0; j < h; j++)
            for (size_t i = 0; i < w; i++) L->t[j * w + i] = L->r[j * w + i] - lv_cell(L, L->z, j, i);
        or_bil_restrict(w, h, L->wrap, C->w, C->h, C->wrap, L->t, C->r);
    } else {
        restrict_into(L, C);
    }
    vcycle(hy, l + 1, NULL);
    size_t wc = C->w;
    if (hy->bil) {
        /* z' = z + P zc, then the sweep from z' (as the aggregation path adds zc per cell). */
        for (size_t j = 0; j < L->h; j++)
            for (size_t i = 0; i < L->w; i++)
                L->z[j * L->w + i] = L->z[j * L->w + i] + bil_at(C->z, C->w, C->h, C->wrap, i, j);
    }
    double acc = sweep(L, hy->bil ? NULL : C->z, wc, hy->nu == 1 ? dot : NULL);
    for (size_t k = 1; k < hy->nu; k++) acc = sweep(L, NULL, 0, k + 1 == hy->nu ? dot : NULL);
    return acc;
}

/* `Hierarchy::new`, MG:481-487, with the cycle's options (transfer 1: bilinear; coarsest 1: the
 * exact f64 solve). */
or_hier *or_hier_new_opts(size_t w, size_t h, int wrap_x, const double *east, const double *south,
                          const double *diag, float omega, size_t nu, float scale, int transfer,
                          int coarsest);

or_hier *or_hier_new(size_t w, size_t h, int wrap_x, const double *east, const double *south,
                     const double *diag, float omega, size_t nu, float scale) {
    return or_hier_new_opts(w, h, wrap_x, east, south, diag, omega, nu, scale, 0, 0);
}

or_hier *or_hier_new_opts(size_t w, size_t h, int wrap_x, const double *east, const double *south,
                          const double *diag, float omega, size_t nu, float scale, int transfer,
                          int coarsest) {
    or_hier *hy = (or_hier *)calloc(1, sizeof(or_hier));
    hy->bil = transfer == 1;
    size_t cap = 4;
    hy->lv = (lvl *)calloc(cap, sizeof(lvl));
    hy->omega = omega;
    hy->nu = nu;
    hy->scale = scale;
    level_finest(&hy->lv[0], w, h, wrap_x && w > 1, east, south, diag, omega);
    hy->n = 1;
    for (;;) {
        if (hy->n == cap) {
            cap *= 2;
            hy->lv = (lvl *)realloc(hy->lv, cap * sizeof(lvl));
        }
        lvl next;
        memset(&next, 0, sizeof next);
        if (!level_coarsen(&hy->lv[hy->n - 1], &next, omega, scale)) break;
        hy->lv[hy->n++] = next;
    }
    if (coarsest == 1) {
        const lvl *L = &hy->lv[hy->n - 1];
        size_t n = L->w * L->h;
        hy->chol = (double *)malloc((n > 0 ? n * n : 1) * sizeof(double));
        coarse_factor(L->w, L->h, L->wrap, L->east, L->south, L->diag, hy->chol);
    }
    return hy;
}

void or_hier_free(or_hier *hy) {
    if (!hy) return;
    for (size_t l = 0; l < hy->n; l++) {
        lvl *L = &hy->lv[l];
        free(L->east);
        free(L->south);
        free(L->diag);
        free(L->wd);
        free(L->mass);
        free(L->r);
        free(L->z);
        free(L->t);
        free(L->zp);
    }
    free(hy->lv);
    free(hy->chol);
    free(hy);
}

size_t or_hier_levels(const or_hier *hy) { return hy->n; }

void or_hier_level(const or_hier *hy, size_t l, size_t *w, size_t *h, int *wrap, float *east,
                   float *south, float *diag, float *wd, float *mass) {
    const lvl *L = &hy->lv[l];
    size_t n = L->w * L->h;
    if (w) *w = L->w;
    if (h) *h = L->h;
    if (wrap) *wrap = L->wrap;
    if (east) memcpy(east, L->east, n * sizeof(float));
    if (south) memcpy(south, L->south, n * sizeof(float));
    if (diag) memcpy(diag, L->diag, n * sizeof(float));
    if (wd) memcpy(wd, L->wd, n * sizeof(float));
    if (mass) {
        if (L->mass)
            memcpy(mass, L->mass, n * sizeof(float));
        else
            for (size_t c = 0; c < n; c++) mass[c] = 1.0f;
    }
}

/* `Hierarchy::precondition`, MG:489-492. */
double or_hier_precondition(or_hier *hy, const float *r32, const double *r64, float *z) {
    size_t n = hy->lv[0].w * hy->lv[0].h;
    memcpy(hy->lv[0].r, r32, n * sizeof(float));
    double rz = vcycle(hy, 0, r64);
    if (z) memcpy(z, hy->lv[0].z, n * sizeof(float));
    return rz;
}

/* ---- solves -------------------------------------------------------------------------------- */

/* The residual's sum put back in proportion to each cell's coupling, MG:737-761: per row the
 * sums of (b - A x) and (diag - 1) from -0.0 in column order, rows folded in order. */
static void put_back(size_t w, size_t h, int wrap, const double *east, const double *south,
                     const double *diag, const double *b, double *x, double *ap) {
    or_apply(w, h, wrap, east, south, diag, x, ap);
    double lost = sum_zero(), weight = sum_zero();
    for (size_t j = 0; j < h; j++) {
        double l = sum_zero(), g = sum_zero();
        for (size_t i = 0; i < w; i++) {
            size_t c = j * w + i;
            l += b[c] - ap[c];
            g += diag[c] - 1.0;
        }
        lost = lost + l;
        weight = weight + g;
    }
    if (weight > 0.0 && lost != 0.0) {
        double k = lost / weight;
        for (size_t c = 0; c < w * h; c++) x[c] += k * (diag[c] - 1.0);
    }
}

/* `cg_on`, MG:562-763: f64 conjugate gradients preconditioned by one f32 V-cycle to an RMS
 * residual of tol_m, then the mass put-back.  Returns the iterations.  Reproduces the reference's
 * loop exactly, including that the pending x += alpha p is not cleared by the direction pass
 * (MG:650-657), so a break on p.A.p <= 0 applies it a second time after the loop (MG:731-735). */
size_t or_cg_on(size_t w, size_t h, double dx, int wrap_x, double *x, const double *b,
                const double *kd, double dt, double tol_m, float omega, size_t nu, float scale) {
    size_t n = w * h;
    if (n == 0) return 0;
    int wrap = wrap_x && w > 1;
    double *east = (double *)malloc(n * sizeof(double));
    double *south = (double *)malloc(n * sizeof(double));
    double *diag = (double *)malloc(n * sizeof(double));
    double *r = (double *)malloc(n * sizeof(double));
    double *ap = (double *)malloc(n * sizeof(double));
    double *pold = (double *)calloc(n, sizeof(double));
    double *pnew = (double *)calloc(n, sizeof(double));
    or_operator(w, h, dx, wrap_x, kd, dt, east, south, diag);
    or_hier *hy = or_hier_new(w, h, wrap_x, east, south, diag, omega, nu, scale);
    float *r32 = hy->lv[0].r;
    size_t ns = div_ceil(h, STRIP_ROWS);
    double *acc = (double *)malloc(ns * sizeof(double));

    /* r = b - A x, its f32 copy, r.r (MG:596-619). */
#pragma omp parallel for schedule(static)
    for (long sl = 0; sl < (long)ns; sl++) {
        size_t j0 = (size_t)sl * STRIP_ROWS, j1 = j0 + STRIP_ROWS < h ? j0 + STRIP_ROWS : h;
        double a = sum_zero();
        for (size_t j = j0; j < j1; j++)
            for (size_t i = 0; i < w; i++) {
                size_t c = j * w + i;
                double rv = b[c] - op_cell(w, h, wrap, east, south, diag, x, j, i);
                r[c] = rv;
                r32[c] = (float)rv;
                a += rv * rv;
            }
        acc[sl] = a;
    }
    double rr0 = sum_zero();
    for (size_t s = 0; s < ns; s++) rr0 += acc[s];
    double tol2 = tol_m * tol_m * (double)n;
    size_t iterations = 0;
    if (rr0 > tol2) {
        double rz = vcycle(hy, 0, r);
        int have_beta = 0, have_pending = 0;
        double beta = 0.0, alpha = 0.0;
        while (iterations < CG_MAX_ITER) {
            iterations++;
            const float *z = hy->lv[0].z;
            /* p = z + beta p_old, A p, p.A p, and the pending x += alpha p_old (MG:633-695). */
#pragma omp parallel for schedule(static)
            for (long c = 0; c < (long)n; c++) {
                if (have_pending) x[c] += alpha * pold[c];
                pnew[c] = have_beta ? (double)z[c] + beta * pold[c] : (double)z[c];
            }
#pragma omp parallel for schedule(static)
            for (long sl = 0; sl < (long)ns; sl++) {
                size_t j0 = (size_t)sl * STRIP_ROWS;
                size_t j1 = j0 + STRIP_ROWS < h ? j0 + STRIP_ROWS : h;
                double a = sum_zero();
                for (size_t j = j0; j < j1; j++)
                    for (size_t i = 0; i < w; i++) {
                        size_t c = j * w + i;
                        double av = op_cell(w, h, wrap, east, south, diag, pnew, j, i);
                        ap[c] = av;
                        a += pnew[c] * av;
                    }
                acc[sl] = a;
            }
            double pap = sum_zero();
            for (size_t s = 0; s < ns; s++) pap += acc[s];
            double *tp = pold;
            pold = pnew;
            pnew = tp;
            if (pap <= 0.0 || !isfinite(pap)) break;
            alpha = rz / pap;
            have_pending = 1;
            /* r -= alpha A p, its f32 copy, r.r (MG:703-723). */
#pragma omp parallel for schedule(static)
            for (long sl = 0; sl < (long)ns; sl++) {
                size_t c0 = (size_t)sl * STRIP_ROWS * w;
                size_t c1 = c0 + STRIP_ROWS * w < n ? c0 + STRIP_ROWS * w : n;
                double a = sum_zero();
                for (size_t c = c0; c < c1; c++) {
                    r[c] -= alpha * ap[c];
                    r32[c] = (float)r[c];
                    a += r[c] * r[c];
                }
                acc[sl] = a;
            }
            double rr = sum_zero();
            for (size_t s = 0; s < ns; s++) rr += acc[s];
            if (rr <= tol2) break;
            double rz_new = vcycle(hy, 0, r);
            if (!(isfinite(rz_new) && rz > 0.0)) break;
            beta = rz_new / rz;
            have_beta = 1;
            rz = rz_new;
        }
        if (have_pending)
            for (size_t c = 0; c < n; c++) x[c] += alpha * pold[c];
    }
    put_back(w, h, wrap, east, south, diag, b, x, ap);
    or_hier_free(hy);
    free(acc);
    free(east);
    free(south);
    free(diag);
    free(r);
    free(ap);
    free(pold);
    free(pnew);
    return iterations;
}

/* `solve` with the multigrid solver, D:77-116 (`multigrid_in`): the right-hand side is h, the
 * first guess the splitting, then the preconditioned conjugate gradients. */
size_t or_solve(size_t w, size_t h, double dx, int wrap_x, double *hgt, const double *kd,
                double dt, double tol_m) {
    size_t n = w * h;
    if (n == 0) return 0;
    double *b = (double *)malloc(n * sizeof(double));
    memcpy(b, hgt, n * sizeof(double));
    or_lod(w, h, dx, wrap_x, hgt, kd, dt);
    size_t its = or_cg_on(w, h, dx, wrap_x, hgt, b, kd, dt, tol_m, OMEGA, NU, SCALE);
    free(b);
    return its;
}

/* The Jacobi-preconditioned conjugate gradients `cg`, D:927-1026, to an RMS residual of tol_m:
 * per-row sums (ascending columns from -0.0), rows added in order; z = r / diag formed where
 * used; x += alpha p deferred to the next direction pass (taken once there).  strip_on (NULL:
 * every strip) marks the visited strips of STRIP_ROWS rows (`Live`): on a skipped strip the
 * operator keeps +0 faces and diagonal 1 (D:585-600), r, p, A p and the sums are +0 (D:757-772,
 * 799-807) and x is left alone; the put-back adds one +0 per skipped row (D:1000-1012). */
static size_t jacobi_cg_live(size_t w, size_t h, double dx, int wrap_x, double *x,
                             const double *b, const double *kd, double dt, double tol_m,
                             const unsigned char *strip_on) {
    size_t n = w * h;
    if (n == 0) return 0;
    int wrap = wrap_x && w > 1;
    size_t nstrips = div_ceil(h, STRIP_ROWS);
#define ROW_ON(j) (!strip_on || strip_on[(j) / STRIP_ROWS])
    double *east = (double *)malloc(n * sizeof(double));
    double *south = (double *)malloc(n * sizeof(double));
    double *diag = (double *)malloc(n * sizeof(double));
    double *r = (double *)malloc(n * sizeof(double));
    double *ap = (double *)malloc(n * sizeof(double));
    double *pold = (double *)calloc(n, sizeof(double));
    double *pnew = (double *)calloc(n, sizeof(double));
    double *s0 = (double *)malloc(h * sizeof(double));
    double *s1 = (double *)malloc(h * sizeof(double));
    double *pr = (double *)malloc(h * sizeof(double));
    or_operator(w, h, dx, wrap_x, kd, dt, east, south, diag);
    for (size_t j = 0; j < h; j++)
        if (!ROW_ON(j))
            for (size_t i = 0; i < w; i++) {
                east[j * w + i] = 0.0;
                south[j * w + i] = 0.0;
                diag[j * w + i] = 1.0;
            }
    /* pass_init, D:747-780. */
    for (size_t j = 0; j < h; j++) {
        if (!ROW_ON(j)) {
            for (size_t i = 0; i < w; i++) r[j * w + i] = 0.0;
            s0[j] = 0.0;
            s1[j] = 0.0;
            continue;
        }
        double a = sum_zero(), bsum = sum_zero();
        for (size_t i = 0; i < w; i++) {
            size_t c = j * w + i;
            double rv = b[c] - op_cell(w, h, wrap, east, south, diag, x, j, i);
            r[c] = rv;
            a += rv * rv;
            bsum += rv * (rv / diag[c]);
        }
        s0[j] = a;
        s1[j] = bsum;
    }
    double tol2 = tol_m * tol_m * (double)n;
    size_t iterations = 0;
    double rr0 = sum_zero();
    for (size_t j = 0; j < h; j++) rr0 += s0[j];
    if (rr0 > tol2) {
        double rz = sum_zero();
        for (size_t j = 0; j < h; j++) rz += s1[j];
        int have_beta = 0, pending = 0;
        double beta = 0.0, alpha = 0.0;
        while (iterations < CG_MAX_ITER) {
            iterations++;
            /* pass_direction, D:782-880: the pending x update and p on the visited strips,
             * +0 on the others; then A p and p.A p per row. */
            for (size_t s = 0; s < nstrips; s++) {
                size_t j0 = s * STRIP_ROWS, j1 = j0 + STRIP_ROWS < h ? j0 + STRIP_ROWS : h;
                for (size_t c = j0 * w; c < j1 * w; c++) {
                    if (strip_on && !strip_on[s]) {
                        pnew[c] = 0.0;
                        continue;
                    }
                    if (pending) x[c] += alpha * pold[c];
                    pnew[c] = have_beta ? r[c] / diag[c] + beta * pold[c] : r[c] / diag[c];
                }
            }
            pending = 0;
            double pap = sum_zero();
            for (size_t j = 0; j < h; j++) {
                if (!ROW_ON(j)) {
                    for (size_t i = 0; i < w; i++) ap[j * w + i] = 0.0;
                    pr[j] = 0.0;
                } else {
                    double a = sum_zero();
                    for (size_t i = 0; i < w; i++) {
                        size_t c = j * w + i;
                        ap[c] = op_cell(w, h, wrap, east, south, diag, pnew, j, i);
                        a += pnew[c] * ap[c];
                    }
                    pr[j] = a;
                }
                pap += pr[j];
            }
            double *tp = pold;
            pold = pnew;
            pnew = tp;
            if (pap <= 0.0) break;
            alpha = rz / pap;
            /* pass_update, D:883-905: skipped strips keep r and their sums. */
            for (size_t j = 0; j < h; j++) {
                if (!ROW_ON(j)) continue;
                double a = sum_zero(), bsum = sum_zero();
                for (size_t i = 0; i < w; i++) {
                    size_t c = j * w + i;
                    r[c] -= alpha * ap[c];
                    a += r[c] * r[c];
                    bsum += r[c] * (r[c] / diag[c]);
                }
                s0[j] = a;
                s1[j] = bsum;
            }
            pending = 1;
            double rr = sum_zero();
            for (size_t j = 0; j < h; j++) rr += s0[j];
            if (rr <= tol2) break;
            double rz_new = sum_zero();
            for (size_t j = 0; j < h; j++) rz_new += s1[j];
            beta = rz_new / rz;
            have_beta = 1;
            rz = rz_new;
        }
        /* pass_x, D:907-919. */
        if (pending)
            for (size_t c = 0; c < n; c++)
                if (ROW_ON(c / w)) x[c] += alpha * pold[c];
    }
    /* Put-back, D:1000-1023: one running sum over the cells in order (a skipped row adds +0). */
    double lost = sum_zero(), weight = sum_zero();
    for (size_t j = 0; j < h; j++) {
        if (!ROW_ON(j)) {
            lost += 0.0;
            weight += 0.0;
            continue;
        }
        for (size_t i = 0; i < w; i++) {
            size_t c = j * w + i;
            ap[c] = op_cell(w, h, wrap, east, south, diag, x, j, i);
        }
        for (size_t i = 0; i < w; i++) lost += b[j * w + i] - ap[j * w + i];
        for (size_t i = 0; i < w; i++) weight += diag[j * w + i] - 1.0;
    }
    if (weight > 0.0 && lost != 0.0) {
        double k = lost / weight;
        for (size_t c = 0; c < n; c++)
            if (ROW_ON(c / w)) x[c] += k * (diag[c] - 1.0);
    }
#undef ROW_ON
    free(east);
    free(south);
    free(diag);
    free(r);
    free(ap);
    free(pold);
    free(pnew);
    free(s0);
    free(s1);
    free(pr);
    return iterations;
}

size_t or_jacobi_cg(size_t w, size_t h, double dx, int wrap_x, double *x, const double *b,
                    const double *kd, double dt, double tol_m) {
    return jacobi_cg_live(w, h, dx, wrap_x, x, b, kd, dt, tol_m, NULL);
}

/* `Live::of`, D:308-355: whether a step may skip rows, and which strips it visits.  A row is a
 * zero row when its kd is +0 throughout (bit pattern 0); a row is skipped when it and its
 * neighbours are zero rows; a strip is visited when it holds a row that is not skipped.  The
 * skipping is off (returns 0) unless some strip is skipped, and unless every zero row's elevations
 * v have 0.5 v normal. */
int or_live_of(size_t w, size_t h, double dx, const double *hgt, const double *kd, double dt,
               unsigned char *strip_on) {
    double s2 = dx * dx;
    if (w == 0 || h == 0 || !(isfinite(dt) && dt > 0.0 && isfinite(s2) && s2 > 0.0)) return 0;
    unsigned char *zero = (unsigned char *)malloc(h);
    int any = 0;
    for (size_t j = 0; j < h; j++) {
        int z = 1;
        for (size_t i = 0; i < w && z; i++) {
            uint64_t bits;
            memcpy(&bits, &kd[j * w + i], 8);
            z = bits == 0;
        }
        zero[j] = (unsigned char)z;
        any |= z;
    }
    if (!any) {
        free(zero);
        return 0;
    }
    size_t nstrips = div_ceil(h, STRIP_ROWS);
    int all = 1;
    for (size_t s = 0; s < nstrips; s++) {
        size_t j1 = (s + 1) * STRIP_ROWS < h ? (s + 1) * STRIP_ROWS : h;
        int on = 0;
        for (size_t j = s * STRIP_ROWS; j < j1; j++) {
            int skip = zero[j] && (j == 0 || zero[j - 1]) && (j + 1 == h || zero[j + 1]);
            on |= !skip;
        }
        strip_on[s] = (unsigned char)on;
        all &= on;
    }
    int normal = 1;
    for (size_t j = 0; j < h && normal; j++)
        if (zero[j])
            for (size_t i = 0; i < w && normal; i++)
                normal = fpclassify(0.5 * hgt[j * w + i]) == FP_NORMAL;
    free(zero);
    return !all && normal;
}

/* `step_in`, D:262-283: the right-hand side is h, the splitting's first guess and the Jacobi
 * conjugate gradients to CG_TOL_M over the strips `Live` visits; when rows were skipped and a
 * visited value ends up not finite, the step is solved again on every row. */
size_t or_step_in(size_t w, size_t h, double dx, int wrap_x, double *hgt, const double *kd,
                  double dt) {
    size_t n = w * h;
    if (n == 0) return 0;
    size_t nstrips = div_ceil(h, STRIP_ROWS);
    unsigned char *strip_on = (unsigned char *)malloc(nstrips);
    unsigned char *row_on = (unsigned char *)malloc(h);
    int live = or_live_of(w, h, dx, hgt, kd, dt, strip_on);
    for (size_t j = 0; j < h; j++) row_on[j] = !live || strip_on[j / STRIP_ROWS];
    double *b = (double *)malloc(n * sizeof(double));
    memcpy(b, hgt, n * sizeof(double));
    lod_live(w, h, dx, wrap_x, hgt, kd, dt, live ? row_on : NULL);
    size_t its = jacobi_cg_live(w, h, dx, wrap_x, hgt, b, kd, dt, CG_TOL_M, live ? strip_on : NULL);
    if (live) {
        int finite = 1;
        for (size_t c = 0; c < n && finite; c++)
            if (row_on[c / w]) finite = isfinite(hgt[c]);
        if (!finite) {
            memcpy(hgt, b, n * sizeof(double));
            lod_live(w, h, dx, wrap_x, hgt, kd, dt, NULL);
            its = jacobi_cg_live(w, h, dx, wrap_x, hgt, b, kd, dt, CG_TOL_M, NULL);
        }
    }
    free(b);
    free(strip_on);
    free(row_on);
    return its;
}

/* `spread::diffusivity`, spread.rs:76-84: zero up to the strength height, d0 x^3 above it with x
 * the excess over it (relative, at most MAX_EXCESS = 2, spread.rs:60). */
double or_spread_diffusivity(double z, double strength, double d0) {
    if (z > strength && strength > 0.0) {
        double x = z / strength - 1.0;
        if (x > 2.0) x = 2.0;
        return d0 * x * x * x;
    }
    return 0.0;
}

/* `reference`, MGT:46-74: unpreconditioned CG on the definitional operator to a relative
 * residual of 1e-13. */
size_t or_reference_cg(size_t w, size_t h, double dx, int wrap_x, const double *b,
                       const double *kd, double dt, double *x) {
    size_t n = w * h, its = 0;
    double *ax = (double *)malloc((n ? n : 1) * sizeof(double));
    double *r = (double *)malloc((n ? n : 1) * sizeof(double));
    double *p = (double *)malloc((n ? n : 1) * sizeof(double));
    memcpy(x, b, n * sizeof(double));
    or_apply_reference(w, h, dx, wrap_x, kd, dt, x, ax);
    double rr = sum_zero(), bb = sum_zero();
    for (size_t k = 0; k < n; k++) {
        r[k] = b[k] - ax[k];
        p[k] = r[k];
        rr += r[k] * r[k];
        bb += b[k] * b[k];
    }
    if (bb < 1.0) bb = 1.0;
    for (size_t it = 0; it < 20 * n + 100; it++) {
        if (rr <= 1e-26 * bb) break;
        its++;
        or_apply_reference(w, h, dx, wrap_x, kd, dt, p, ax);
        double pap = sum_zero();
        for (size_t k = 0; k < n; k++) pap += p[k] * ax[k];
        double alpha = rr / pap;
        for (size_t k = 0; k < n; k++) {
            x[k] += alpha * p[k];
            r[k] -= alpha * ax[k];
        }
        double rr_new = sum_zero();
        for (size_t k = 0; k < n; k++) rr_new += r[k] * r[k];
        double beta = rr_new / rr;
        rr = rr_new;
        for (size_t k = 0; k < n; k++) p[k] = r[k] + beta * p[k];
    }
    free(ax);
    free(r);
    free(p);
    return its;
}

/* `bound`, MGT:157-164: the RMS error a solve stopped at RMS residual `tol` may have. */
double or_bound(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                double tol) {
    size_t n = w * h;
    double *e = (double *)malloc(n * sizeof(double));
    double *s = (double *)malloc(n * sizeof(double));
    double *g = (double *)malloc(n * sizeof(double));
    or_operator(w, h, dx, wrap_x, kd, dt, e, s, g);
    double m = sum_zero(), q = sum_zero();
    for (size_t c = 0; c < n; c++) m += g[c] - 1.0;
    for (size_t c = 0; c < n; c++) q += (g[c] - 1.0) * (g[c] - 1.0);
    double mean = m / (double)n, rms = sqrt(q / (double)n);
    free(e);
    free(s);
    free(g);
    return tol * (1.0 + (mean > 0.0 ? rms / mean : 0.0)) * 1.01;
}

/* ||b - A x||_2 / ||b||_2 in f64 (the benchmark's convergence measure; not a reference function). */
double or_rel_residual(size_t w, size_t h, double dx, int wrap_x, const double *kd, double dt,
                       const double *b, const double *x) {
    size_t n = w * h;
    double *e = (double *)malloc(n * sizeof(double));
    double *s = (double *)malloc(n * sizeof(double));
    double *g = (double *)malloc(n * sizeof(double));
    double *y = (double *)malloc(n * sizeof(double));
    or_operator(w, h, dx, wrap_x, kd, dt, e, s, g);
    or_apply(w, h, wrap_x, e, s, g, x, y);
    double rr = 0.0, bb = 0.0;
    for (size_t c = 0; c < n; c++) {
        double d = b[c] - y[c];
        rr += d * d;
        bb += b[c] * b[c];
    }
    free(e);
    free(s);
    free(g);
    free(y);
    return bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
}

/* ---- submarine slope failure (MAR:254-580) ---------------------------------------------------- */

/* D8 offsets in the grid's fixed order E, NE, N, NW, W, SW, S, SE ("north" is row j - 1) and
 * whether each is diagonal (GRID:21-31). */
static const int D8_DI[8] = {1, 1, 0, -1, -1, -1, 0, 1};
static const int D8_DJ[8] = {0, -1, -1, -1, 0, 1, 1, 1};
static const int D8_DIAG[8] = {0, 1, 0, 1, 0, 1, 0, 1};
/* The double nearest sqrt 2 (0x3FF6A09E667F3BCD, Rust's std::f64::consts::SQRT_2). */
static const double SQRT_2 = 1.4142135623730951;

/* `Dims::neighbour` (GRID:46-60) by the projection's rule (PROJ:350-368): rows clamp, columns wrap
 * when wrap_x (rem_euclid), and the cell itself is no neighbour.  -1 when there is none. */
long or_d8_neighbour(size_t w, size_t h, int wrap_x, size_t c, int k) {
    long i = (long)(c % w), j = (long)(c / w);
    long nj = j + D8_DJ[k], ni = i + D8_DI[k];
    if (nj < 0 || nj >= (long)h) return -1;
    if (ni < 0 || ni >= (long)w) {
        if (!wrap_x) return -1;
        ni = ((ni % (long)w) + (long)w) % (long)w;
    }
    long m = nj * (long)w + ni;
    return m == (long)c ? -1 : m;
}

/* The slope failure as a full sweep of every cell every iteration (the test's
 * `slope_failure_reference`, MAR:595-655, which the active-set path of MAR:292-580 equals bit for
 * bit).  Per iteration: each open-sea cell's share min(avail / give, 1) of what it would give,
 * give = sum over wet neighbours of rate * max(h_c - h_m - S_c d_k, 0); then each cell's change
 * sum over wet neighbours of rate * (in * share_m - out * share_c), neighbours in D8 order; the
 * movers (change != 0) in ascending cell order take h += dz, avail = max(avail + dz, 0), and
 * moved += |dz| / 2.  The loop stops after an iteration that moves nothing.  S_c d_k is
 * critical_slope * dx, or critical_slope * (dx * sqrt 2) for a diagonal (`step_len`,
 * GRID:65-71).  Returns the metres moved. */
double or_slope_failure(size_t w, size_t h, double dx, int wrap_x, const unsigned char *sea,
                        double sea_level, double critical_slope, double rate, unsigned iterations,
                        double *hgt, double *avail) {
    size_t n = w * h;
    if (n == 0) return 0.0;
    double crit[8];
    for (int k = 0; k < 8; k++) crit[k] = critical_slope * (D8_DIAG[k] ? dx * SQRT_2 : dx);
    double *scale = (double *)malloc(n * sizeof(double));
    double *dz = (double *)malloc(n * sizeof(double));
    unsigned char *wet = (unsigned char *)malloc(n);
    double moved = 0.0;
    for (unsigned it = 0; it < iterations; it++) {
#pragma omp parallel for schedule(static)
        for (long cl = 0; cl < (long)n; cl++) wet[cl] = sea[cl] && hgt[cl] < sea_level;
#pragma omp parallel for schedule(static)
        for (long cl = 0; cl < (long)n; cl++) {
            size_t c = (size_t)cl;
            double sc = 0.0;
            if (wet[c] && avail[c] > 0.0) {
                double give = 0.0;
                for (int k = 0; k < 8; k++) {
                    long m = or_d8_neighbour(w, h, wrap_x, c, k);
                    if (m >= 0 && wet[m]) give += rate * fmax(hgt[c] - hgt[m] - crit[k], 0.0);
                }
                if (give > 0.0) sc = fmin(avail[c] / give, 1.0);
            }
            scale[c] = sc;
        }
#pragma omp parallel for schedule(static)
        for (long cl = 0; cl < (long)n; cl++) {
            size_t c = (size_t)cl;
            double d = 0.0;
            if (wet[c])
                for (int k = 0; k < 8; k++) {
                    long m = or_d8_neighbour(w, h, wrap_x, c, k);
                    if (m < 0 || !wet[m]) continue;
                    double out = fmax(hgt[c] - hgt[m] - crit[k], 0.0);
                    double inn = fmax(hgt[m] - hgt[c] - crit[k], 0.0);
                    d += rate * (inn * scale[m] - out * scale[c]);
                }
            dz[c] = d;
        }
        int any = 0;
        for (size_t c = 0; c < n; c++)
            if (dz[c] != 0.0) {
                any = 1;
                hgt[c] += dz[c];
                avail[c] = fmax(avail[c] + dz[c], 0.0);
                moved += fabs(dz[c]) * 0.5;
            }
        if (!any) break;
    }
    free(scale);
    free(dz);
    free(wet);
    return moved;
}

/* ---- cube-grid diffusion (CD, CM) ------------------------------------------------------------ */

#define CUBE_CHUNK 4096          /* CUBE:48 */
#define CUBE_NO_CELL 0xFFFFFFFFu /* sphere/cube.rs:65 */

typedef struct {
    size_t l;           /* cells taking part */
    int all;            /* every cell takes part (cells[i] == i) */
    uint32_t *cells;    /* ascending */
    uint32_t *cols;     /* l x 8: each slot's neighbour as a position, CUBE_NO_CELL for none */
    double *k, *mu, *s, *diag; /* l x 8 couplings; mass; sum of couplings; mu + s */
} cube_op;

static void cube_op_free(cube_op *op) {
    free(op->cells), free(op->cols), free(op->k), free(op->mu), free(op->s), free(op->diag);
}

/* `couplings`, CD:94-119: a side pair's mean kd, a diagonal's least of its quad's four side
 * means, times scale = dt / mean area and the grid's coupling. */
static void cube_couplings(const uint32_t *slots, const double *lap, const double *kd,
                           double scale, size_t c, double *k) {
    const uint32_t *sl = slots + 8 * c;
    for (int q = 0; q < 8; q++) {
        k[q] = 0.0;
        if (sl[q] == CUBE_NO_CELL) continue;
        size_t m = sl[q];
        double d;
        if (q % 2 == 0) {
            d = 0.5 * (kd[c] + kd[m]);
        } else {
            size_t a = sl[q - 1], b = sl[(q + 1) % 8];
            d = fmin(fmin(fmin(0.5 * (kd[c] + kd[a]), 0.5 * (kd[a] + kd[m])), 0.5 * (kd[m] + kd[b])),
                     0.5 * (kd[b] + kd[c]));
        }
        k[q] = scale * d * lap[8 * c + q];
    }
}

/* `Workspace::build`, CD:148-224: every cell (all, or when every cell has a coupling) or the
 * cells with a coupling, ascending, their neighbours as positions among them. */
static void cube_build(size_t n6, const uint32_t *slots, const double *lap, const double *area,
                       double mean_area, const double *kd, double dt, int all, cube_op *op) {
    double scale = dt / mean_area;
    double *full = (double *)malloc(8 * n6 * sizeof(double));
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)n6; c++) cube_couplings(slots, lap, kd, scale, (size_t)c, full + 8 * c);
    int every = all;
    if (!every) {
        every = 1;
        for (size_t c = 0; c < n6 && every; c++) {
            int any = 0;
            for (int q = 0; q < 8; q++) any |= full[8 * c + q] != 0.0;
            every = any;
        }
    }
    op->all = every;
    uint32_t *pos = (uint32_t *)malloc(n6 * sizeof(uint32_t));
    size_t l = 0;
    for (size_t c = 0; c < n6; c++) {
        int any = every;
        for (int q = 0; q < 8 && !any; q++) any = full[8 * c + q] != 0.0;
        pos[c] = any ? (uint32_t)l++ : CUBE_NO_CELL;
    }
    op->l = l;
    op->cells = (uint32_t *)malloc((l ? l : 1) * sizeof(uint32_t));
    op->cols = (uint32_t *)malloc((l ? l : 1) * 8 * sizeof(uint32_t));
    op->k = (double *)malloc((l ? l : 1) * 8 * sizeof(double));
    op->mu = (double *)malloc((l ? l : 1) * sizeof(double));
    op->s = (double *)malloc((l ? l : 1) * sizeof(double));
    op->diag = (double *)malloc((l ? l : 1) * sizeof(double));
    for (size_t c = 0; c < n6; c++) {
        if (pos[c] == CUBE_NO_CELL) continue;
        size_t i = pos[c];
        op->cells[i] = (uint32_t)c;
        for (int q = 0; q < 8; q++) {
            uint32_t m = slots[8 * c + q];
            op->cols[8 * i + q] = m == CUBE_NO_CELL ? CUBE_NO_CELL : pos[m];
            op->k[8 * i + q] = full[8 * c + q];
        }
        op->mu[i] = area[c] / mean_area;
        double sum = -0.0; /* `k.iter().sum()` */
        for (int q = 0; q < 8; q++) sum += op->k[8 * i + q];
        op->s[i] = sum;
        op->diag[i] = op->mu[i] + sum;
    }
    free(pos);
    free(full);
}

/* `Op::row`, CD:79-91. */
static double cube_row(const cube_op *op, size_t i, const double *p) {
    double acc = op->diag[i] * p[i];
    for (int q = 0; q < 8; q++) {
        uint32_t j = op->cols[8 * i + q];
        if (j != CUBE_NO_CELL) acc -= op->k[8 * i + q] * p[j];
    }
    return acc;
}

/* The cube's LOD first guess, CD:316-378: rows of each face (slots E and W), then columns (N, S),
 * each (mu + K) y = mu b divided through by mu, the couplings across a face's edge taken at the
 * old value of the cell beyond it. */
static void cube_lod(size_t n, const uint32_t *slots, const cube_op *op, const double *b, double *x,
                     double *tmp) {
    size_t nn = n * n;
#pragma omp parallel
    {
        double *lo = (double *)malloc(n * sizeof(double)), *di = (double *)malloc(n * sizeof(double));
        double *up = (double *)malloc(n * sizeof(double));
#pragma omp for schedule(static)
        for (long row = 0; row < (long)(6 * n); row++) {
            size_t c0 = (size_t)row * n;
            double *xr = x + c0;
            for (size_t i = 0; i < n; i++) {
                size_t c = c0 + i;
                double ke = op->k[8 * c + 0], kw = op->k[8 * c + 4], mu = op->mu[c];
                di[i] = 1.0 + (ke + kw) / mu;
                double v = b[c];
                lo[i] = 0.0, up[i] = 0.0;
                if (i > 0) lo[i] = -kw / mu;
                else v += kw * b[slots[8 * c + 4]] / mu;
                if (i + 1 < n) up[i] = -ke / mu;
                else v += ke * b[slots[8 * c + 0]] / mu;
                xr[i] = v;
            }
            or_thomas(n, lo, di, up, xr);
        }
#pragma omp for schedule(static)
        for (long col = 0; col < (long)(6 * n); col++) {
            size_t f = (size_t)col / n, i = (size_t)col % n;
            double *tc = tmp + (size_t)col * n;
            for (size_t j = 0; j < n; j++) {
                size_t c = f * nn + j * n + i;
                double kn = op->k[8 * c + 2], ks = op->k[8 * c + 6], mu = op->mu[c];
                di[j] = 1.0 + (kn + ks) / mu;
                double v = x[c];
                lo[j] = 0.0, up[j] = 0.0;
                if (j > 0) lo[j] = -kn / mu;
                else v += kn * x[slots[8 * c + 2]] / mu;
                if (j + 1 < n) up[j] = -ks / mu;
                else v += ks * x[slots[8 * c + 6]] / mu;
                tc[j] = v;
            }
            or_thomas(n, lo, di, up, tc);
        }
        free(lo), free(di), free(up);
    }
#pragma omp parallel for schedule(static)
    for (long row = 0; row < (long)(6 * n); row++) {
        size_t f = (size_t)row / n, j = (size_t)row % n;
        for (size_t i = 0; i < n; i++) x[(size_t)row * n + i] = tmp[f * nn + i * n + j];
    }
}

/* -- the cube's V-cycle (CM) -- */

typedef struct {
    size_t side, cells;
    uint32_t *ptr, *col;         /* neighbour lists (CSR) */
    uint32_t *src_ptr, *src;     /* per entry, the finer entries summed into it (level >= 1) */
    uint32_t *kid_ptr, *kids;    /* per cell, its block of finer cells, ascending (level >= 1) */
    uint32_t *parent;            /* per finer cell, its block's cell here (level >= 1) */
    float *val, *mass, *diag, *r, *z, *t;
} cube_level;

static void cube_level_free(cube_level *L) {
    free(L->ptr), free(L->col), free(L->src_ptr), free(L->src), free(L->kid_ptr), free(L->kids);
    free(L->parent), free(L->val), free(L->mass), free(L->diag), free(L->r), free(L->z), free(L->t);
}

/* `finest`, CM:65-90: the slots in order, none skipped. */
static void cube_finest(size_t n, const uint32_t *slots, cube_level *L) {
    size_t cells = 6 * n * n;
    memset(L, 0, sizeof *L);
    L->side = n, L->cells = cells;
    L->ptr = (uint32_t *)malloc((cells + 1) * sizeof(uint32_t));
    L->col = (uint32_t *)malloc(8 * cells * sizeof(uint32_t));
    uint32_t at = 0;
    L->ptr[0] = 0;
    for (size_t c = 0; c < cells; c++) {
        for (int q = 0; q < 8; q++)
            if (slots[8 * c + q] != CUBE_NO_CELL) L->col[at++] = slots[8 * c + q];
        L->ptr[c + 1] = at;
    }
}

/* `coarsen`, CM:93-165: 2 x 2 blocks of each face; a coarse cell's neighbours ascending, each
 * with the finer entries between the two blocks in (kid, entry) order. */
static void cube_coarsen(const cube_level *f, cube_level *C) {
    size_t nf = f->side, m = (nf + 1) / 2, cells = 6 * m * m;
    memset(C, 0, sizeof *C);
    C->side = m, C->cells = cells;
    C->parent = (uint32_t *)malloc(f->cells * sizeof(uint32_t));
    for (size_t c = 0; c < f->cells; c++) {
        size_t face = c / (nf * nf), rest = c % (nf * nf), i = rest % nf, j = rest / nf;
        C->parent[c] = (uint32_t)(face * m * m + (j / 2) * m + i / 2);
    }
    C->kid_ptr = (uint32_t *)calloc(cells + 1, sizeof(uint32_t));
    for (size_t c = 0; c < f->cells; c++) C->kid_ptr[C->parent[c] + 1]++;
    for (size_t q = 0; q < cells; q++) C->kid_ptr[q + 1] += C->kid_ptr[q];
    C->kids = (uint32_t *)malloc(f->cells * sizeof(uint32_t));
    uint32_t *fill = (uint32_t *)malloc(cells * sizeof(uint32_t));
    memcpy(fill, C->kid_ptr, cells * sizeof(uint32_t));
    for (size_t c = 0; c < f->cells; c++) C->kids[fill[C->parent[c]]++] = (uint32_t)c;
    free(fill);
    size_t ne = f->ptr[f->cells];
    C->ptr = (uint32_t *)malloc((cells + 1) * sizeof(uint32_t));
    C->col = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
    C->src_ptr = (uint32_t *)malloc((ne + 1) * sizeof(uint32_t));
    C->src = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
    uint32_t *pp = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
    uint32_t *pe = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
    size_t ncol = 0, nsrc = 0;
    C->ptr[0] = 0, C->src_ptr[0] = 0;
    for (size_t r = 0; r < cells; r++) {
        size_t np = 0;
        for (uint32_t q = C->kid_ptr[r]; q < C->kid_ptr[r + 1]; q++) {
            uint32_t c = C->kids[q];
            for (uint32_t e = f->ptr[c]; e < f->ptr[c + 1]; e++) {
                uint32_t p = C->parent[f->col[e]];
                if (p != r) pp[np] = p, pe[np] = e, np++;
            }
        }
        /* Stable insertion sort by the coarse neighbour. */
        for (size_t a = 1; a < np; a++) {
            uint32_t kp = pp[a], ke = pe[a];
            size_t b = a;
            while (b > 0 && pp[b - 1] > kp) pp[b] = pp[b - 1], pe[b] = pe[b - 1], b--;
            pp[b] = kp, pe[b] = ke;
        }
        for (size_t a = 0; a < np; a++) {
            if (a == 0 || pp[a] != pp[a - 1]) {
                if (a > 0) C->src_ptr[ncol] = (uint32_t)nsrc;
                C->col[ncol++] = pp[a];
            }
            C->src[nsrc++] = pe[a];
        }
        if (np) C->src_ptr[ncol] = (uint32_t)nsrc;
        C->ptr[r + 1] = (uint32_t)ncol;
    }
    free(pp), free(pe);
}

typedef struct {
    size_t nl;
    cube_level *lv;
} cube_hier;

/* `Hierarchy::new`, CM:170-186: coarsen until a level has at most COARSEST cells or side 1. */
static void cube_hier_new(size_t n, const uint32_t *slots, cube_hier *h) {
    h->lv = (cube_level *)malloc(40 * sizeof(cube_level));
    cube_finest(n, slots, &h->lv[0]);
    h->nl = 1;
    while (!(h->lv[h->nl - 1].cells <= COARSEST || h->lv[h->nl - 1].side <= 1)) {
        cube_coarsen(&h->lv[h->nl - 1], &h->lv[h->nl]);
        h->nl++;
    }
}

static void cube_hier_free(cube_hier *h) {
    for (size_t l = 0; l < h->nl; l++) cube_level_free(&h->lv[l]);
    free(h->lv);
}

/* `prepare`, CM:197-260: the finest level is the operator in f32 (couplings in slot order, mass,
 * diagonal); a coarse entry is SCALE times the sum of its finer entries, a coarse mass the sum of
 * its block's, its diagonal mass + its couplings. */
static void cube_prepare(cube_hier *h, const uint32_t *slots, const cube_op *op) {
    for (size_t l = 0; l < h->nl; l++) {
        cube_level *L = &h->lv[l];
        size_t ne = L->ptr[L->cells];
        L->val = (float *)malloc((ne ? ne : 1) * sizeof(float));
        L->mass = (float *)malloc(L->cells * sizeof(float));
        L->diag = (float *)malloc(L->cells * sizeof(float));
        L->r = (float *)calloc(L->cells, sizeof(float));
        L->z = (float *)calloc(L->cells, sizeof(float));
        L->t = (float *)calloc(L->cells, sizeof(float));
    }
    cube_level *L0 = &h->lv[0];
    size_t at = 0;
    for (size_t c = 0; c < L0->cells; c++)
        for (int q = 0; q < 8; q++)
            if (slots[8 * c + q] != CUBE_NO_CELL) L0->val[at++] = (float)op->k[8 * c + q];
    for (size_t c = 0; c < L0->cells; c++) L0->mass[c] = (float)op->mu[c], L0->diag[c] = (float)op->diag[c];
    for (size_t l = 1; l < h->nl; l++) {
        const cube_level *F = &h->lv[l - 1];
        cube_level *C = &h->lv[l];
        for (size_t e = 0; e < C->ptr[C->cells]; e++) {
            float sum = 0.0f;
            for (uint32_t q = C->src_ptr[e]; q < C->src_ptr[e + 1]; q++) sum += F->val[C->src[q]];
            C->val[e] = SCALE * sum;
        }
        for (size_t r = 0; r < C->cells; r++) {
            float sm = 0.0f, k = 0.0f;
            for (uint32_t q = C->kid_ptr[r]; q < C->kid_ptr[r + 1]; q++) sm += F->mass[C->kids[q]];
            for (uint32_t e = C->ptr[r]; e < C->ptr[r + 1]; e++) k += C->val[e];
            C->mass[r] = sm;
            C->diag[r] = sm + k;
        }
    }
}

/* `sweep`, CM:285-310: z = OMEGA r / diag from zero, else z + OMEGA (r - A z) / diag. */
static void cube_sweep(cube_level *L, int first) {
    if (first) {
        for (size_t i = 0; i < L->cells; i++) L->z[i] = OMEGA * L->r[i] / L->diag[i];
        return;
    }
    for (size_t i = 0; i < L->cells; i++) {
        float az = L->diag[i] * L->z[i];
        for (uint32_t e = L->ptr[i]; e < L->ptr[i + 1]; e++) az -= L->val[e] * L->z[L->col[e]];
        L->t[i] = L->z[i] + OMEGA * (L->r[i] - az) / L->diag[i];
    }
    float *tmp = L->z;
    L->z = L->t, L->t = tmp;
}

/* `vcycle`, CM:312-363. */
static void cube_vcycle(cube_level *lv, size_t nl) {
    cube_level *L = &lv[0];
    if (nl == 1) {
        for (int k = 0; k < COARSE_SWEEPS; k++) cube_sweep(L, k == 0);
        return;
    }
    for (int k = 0; k < NU; k++) cube_sweep(L, k == 0);
    for (size_t i = 0; i < L->cells; i++) {
        float az = L->diag[i] * L->z[i];
        for (uint32_t e = L->ptr[i]; e < L->ptr[i + 1]; e++) az -= L->val[e] * L->z[L->col[e]];
        L->t[i] = L->r[i] - az;
    }
    cube_level *C = &lv[1];
    for (size_t r = 0; r < C->cells; r++) {
        float sum = 0.0f;
        for (uint32_t q = C->kid_ptr[r]; q < C->kid_ptr[r + 1]; q++) sum += L->t[C->kids[q]];
        C->r[r] = sum;
    }
    cube_vcycle(lv + 1, nl - 1);
    for (size_t c = 0; c < L->cells; c++) L->z[c] += C->z[C->parent[c]];
    for (int k = 0; k < NU; k++) cube_sweep(L, 0);
}

/* `precondition`, CM:262-273. */
static void cube_precondition(cube_hier *h, const double *r, double *z) {
    cube_level *L = &h->lv[0];
    for (size_t i = 0; i < L->cells; i++) L->r[i] = (float)r[i];
    cube_vcycle(h->lv, h->nl);
    for (size_t i = 0; i < L->cells; i++) z[i] = (double)L->z[i];
}

/* `ordered_sum`, CUBE:50-58: each chunk of CUBE_CHUNK summed in order, the chunks in order. */
static double cube_dot(size_t l, const double *a, const double *b, const double *div) {
    double total = -0.0;
    for (size_t c0 = 0; c0 < l; c0 += CUBE_CHUNK) {
        size_t c1 = c0 + CUBE_CHUNK < l ? c0 + CUBE_CHUNK : l;
        double part = -0.0;
        for (size_t i = c0; i < c1; i++) part += div ? a[i] * b[i] / div[i] : a[i] * b[i];
        total += part;
    }
    return total;
}

/* `cg`, CD:380-500: CG from x, preconditioned by the diagonal or the V-cycle, to an RMS residual
 * of tol_m over `cells` cells; two passes an iteration. */
static size_t cube_cg(const cube_op *op, cube_hier *h, double *x, const double *rhs, size_t cells,
                      double tol_m, int mg) {
    size_t l = op->l;
    double tol2 = tol_m * tol_m * (double)cells;
    double *r = (double *)malloc(l * sizeof(double)), *po = (double *)calloc(l, sizeof(double));
    double *pn = (double *)calloc(l, sizeof(double)), *ap = (double *)calloc(l, sizeof(double));
    double *z = (double *)calloc(l, sizeof(double));
    size_t nparts = (l + CUBE_CHUNK - 1) / CUBE_CHUNK;
    double *parts = (double *)malloc(2 * (nparts ? nparts : 1) * sizeof(double));
    for (size_t i = 0; i < l; i++) r[i] = rhs[i] - cube_row(op, i, x);
    size_t its = 0;
    double rr0 = cube_dot(l, r, r, NULL);
    if (!(rr0 <= tol2)) {
        double rz;
        if (mg) {
            cube_precondition(h, r, z);
            rz = cube_dot(l, r, z, NULL);
        } else {
            rz = cube_dot(l, r, r, op->diag);
        }
        int have_beta = 0;
        double beta = 0.0;
        for (int it = 0; it < CG_MAX_ITER; it++) {
#pragma omp parallel for schedule(static)
            for (long q = 0; q < (long)nparts; q++) {
                size_t c0 = (size_t)q * CUBE_CHUNK, c1 = c0 + CUBE_CHUNK < l ? c0 + CUBE_CHUNK : l;
                double sum = 0.0;
                for (size_t i = c0; i < c1; i++) {
#define DIR(j) ((have_beta ? (mg ? z[j] : r[j] / op->diag[j]) + beta * po[j] \
                           : (mg ? z[j] : r[j] / op->diag[j])))
                    double pi = DIR(i);
                    double acc = op->diag[i] * pi;
                    for (int s = 0; s < 8; s++) {
                        uint32_t j = op->cols[8 * i + s];
                        if (j != CUBE_NO_CELL) acc -= op->k[8 * i + s] * DIR(j);
                    }
#undef DIR
                    pn[i] = pi;
                    ap[i] = acc;
                    sum += pi * acc;
                }
                parts[q] = sum;
            }
            double pap = -0.0;
            for (size_t q = 0; q < nparts; q++) pap += parts[q];
            double *sw = po;
            po = pn, pn = sw;
            if (isnan(pap) || pap <= 0.0) break;
            double alpha = rz / pap;
#pragma omp parallel for schedule(static)
            for (long q = 0; q < (long)nparts; q++) {
                size_t c0 = (size_t)q * CUBE_CHUNK, c1 = c0 + CUBE_CHUNK < l ? c0 + CUBE_CHUNK : l;
                double s0 = 0.0, s1 = 0.0;
                for (size_t i = c0; i < c1; i++) {
                    x[i] += alpha * po[i];
                    double rv = r[i] - alpha * ap[i];
                    r[i] = rv;
                    s0 += rv * rv;
                    s1 += rv * rv / op->diag[i];
                }
                parts[2 * q] = s0, parts[2 * q + 1] = s1;
            }
            double rr = -0.0, rzj = -0.0;
            for (size_t q = 0; q < nparts; q++) rr += parts[2 * q];
            for (size_t q = 0; q < nparts; q++) rzj += parts[2 * q + 1];
            its++;
            if (rr <= tol2) break;
            double rz_new;
            if (mg) {
                cube_precondition(h, r, z);
                rz_new = cube_dot(l, r, z, NULL);
            } else {
                rz_new = rzj;
            }
            beta = rz_new / rz;
            have_beta = 1;
            rz = rz_new;
        }
    }
    free(r), free(po), free(pn), free(ap), free(z), free(parts);
    return its;
}

/* `solve` / `step_in` / `run`, CD:227-314, on the grid's slots (6 n^2 x 8, CUBE_NO_CELL for none),
 * couplings (6 n^2 x 8), areas and their mean: the operator, the first guess (every cell taking
 * part), CG by the V-cycle (solver 0, tol_m) or the diagonal (solver 1, `step_in`: the cells with
 * a coupling, CG_TOL_M), and the residual's sum put back in proportion to each cell's coupling. */
size_t or_cube_solve(size_t n, const uint32_t *slots, const double *lap, const double *area,
                     double mean_area, double *hgt, const double *kd, double dt, int solver,
                     double tol_m) {
    if (isnan(dt) || dt <= 0.0) return 0;
    size_t n6 = 6 * n * n;
    int mg = solver == 0;
    cube_op op;
    cube_build(n6, slots, lap, area, mean_area, kd, dt, mg, &op);
    size_t l = op.l;
    if (l == 0) {
        cube_op_free(&op);
        return 0;
    }
    double *rhs = (double *)malloc(l * sizeof(double)), *x = (double *)malloc(l * sizeof(double));
    for (size_t i = 0; i < l; i++) {
        double v = hgt[op.cells[i]];
        rhs[i] = op.mu[i] * v;
        x[i] = v;
    }
    if (op.all) {
        double *b = (double *)malloc(l * sizeof(double)), *tmp = (double *)malloc(l * sizeof(double));
        memcpy(b, x, l * sizeof(double));
        cube_lod(n, slots, &op, b, x, tmp);
        free(b), free(tmp);
    }
    cube_hier h = {0, NULL};
    if (mg) {
        cube_hier_new(n, slots, &h);
        cube_prepare(&h, slots, &op);
    }
    size_t its = cube_cg(&op, &h, x, rhs, n6, mg ? tol_m : CG_TOL_M, mg);
    /* The put-back, CD:286-301. */
    double *ap = (double *)malloc(l * sizeof(double));
    for (size_t i = 0; i < l; i++) ap[i] = cube_row(&op, i, x);
    double lost = -0.0, weight = -0.0;
    for (size_t c0 = 0; c0 < l; c0 += CUBE_CHUNK) {
        size_t c1 = c0 + CUBE_CHUNK < l ? c0 + CUBE_CHUNK : l;
        double a = -0.0, w = -0.0;
        for (size_t i = c0; i < c1; i++) a += rhs[i] - ap[i];
        for (size_t i = c0; i < c1; i++) w += op.s[i];
        lost += a, weight += w;
    }
    if (weight > 0.0 && lost != 0.0) {
        double kk = lost / weight;
        for (size_t i = 0; i < l; i++) x[i] += kk * op.s[i] / op.mu[i];
    }
    for (size_t i = 0; i < l; i++) hgt[op.cells[i]] = x[i];
    free(ap), free(rhs), free(x);
    if (mg) cube_hier_free(&h);
    cube_op_free(&op);
    return its;
}
