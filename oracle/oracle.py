"""ctypes wrapper of the CPU oracle (oracle/dtoracle.c). TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` load this
module, and only as the checker. The product (the HIP library behind include/dtdiff.h) never
calls it. Each wrapper names the reference function it restates; dtoracle.c cites file:line.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libdtoracle.so")

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_sz = ctypes.c_size_t
_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    L.or_splitmix_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    L.or_splitmix_next.restype = ctypes.c_uint64
    L.or_splitmix_f64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    L.or_splitmix_f64.restype = ctypes.c_double
    L.or_terrain.argtypes = [_sz, _sz, ctypes.c_uint64, _f64p]
    L.or_fractal.argtypes = [_sz, _sz, ctypes.c_uint64, _f64p]
    L.or_submarine_kd.argtypes = [ctypes.c_double] * 5
    L.or_submarine_kd.restype = ctypes.c_double
    L.or_shelf_kd.argtypes = [_sz, _f64p, ctypes.c_uint64, _f64p]
    L.or_operator.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_double,
                              _f64p, _f64p, _f64p]
    L.or_apply.argtypes = [_sz, _sz, ctypes.c_int, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.or_apply_reference.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p,
                                     ctypes.c_double, _f64p, _f64p]
    L.or_thomas.argtypes = [_sz, _f64p, _f64p, _f64p, _f64p]
    L.or_cyclic.argtypes = [_sz, _f64p, _f64p, _f64p, ctypes.c_double, ctypes.c_double, _f64p]
    L.or_lod.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, _f64p, ctypes.c_double]
    L.or_hier_new.argtypes = [_sz, _sz, ctypes.c_int, _f64p, _f64p, _f64p, ctypes.c_float, _sz,
                              ctypes.c_float]
    L.or_hier_new.restype = ctypes.c_void_p
    L.or_hier_new_opts.argtypes = [_sz, _sz, ctypes.c_int, _f64p, _f64p, _f64p, ctypes.c_float,
                                   _sz, ctypes.c_float, ctypes.c_int, ctypes.c_int]
    L.or_hier_new_opts.restype = ctypes.c_void_p
    L.or_bil_prolong.argtypes = [_sz, _sz, _sz, _sz, ctypes.c_int, _f32p, _f32p]
    L.or_bil_restrict.argtypes = [_sz, _sz, ctypes.c_int, _sz, _sz, ctypes.c_int, _f32p, _f32p]
    L.or_coarse_exact.argtypes = [_sz, _sz, ctypes.c_int, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.or_hier_free.argtypes = [ctypes.c_void_p]
    L.or_hier_levels.argtypes = [ctypes.c_void_p]
    L.or_hier_levels.restype = _sz
    L.or_hier_level.argtypes = [ctypes.c_void_p, _sz, ctypes.POINTER(_sz), ctypes.POINTER(_sz),
                                ctypes.POINTER(ctypes.c_int)] + [ctypes.c_void_p] * 5
    L.or_hier_precondition.argtypes = [ctypes.c_void_p, _f32p, _f64p, _f32p]
    L.or_hier_precondition.restype = ctypes.c_double
    L.or_cg_on.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, _f64p, _f64p,
                           ctypes.c_double, ctypes.c_double, ctypes.c_float, _sz, ctypes.c_float]
    L.or_cg_on.restype = _sz
    L.or_solve.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, _f64p,
                           ctypes.c_double, ctypes.c_double]
    L.or_solve.restype = _sz
    L.or_jacobi_cg.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, _f64p, _f64p,
                               ctypes.c_double, ctypes.c_double]
    L.or_jacobi_cg.restype = _sz
    L.or_reference_cg.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, _f64p,
                                  ctypes.c_double, _f64p]
    L.or_reference_cg.restype = _sz
    L.or_bound.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_double,
                           ctypes.c_double]
    L.or_bound.restype = ctypes.c_double
    L.or_rel_residual.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p,
                                  ctypes.c_double, _f64p, _f64p]
    L.or_rel_residual.restype = ctypes.c_double
    L.or_threads.restype = ctypes.c_int
    L.or_live_of.argtypes = [_sz, _sz, ctypes.c_double, _f64p, _f64p, ctypes.c_double,
                             ctypes.POINTER(ctypes.c_ubyte)]
    L.or_live_of.restype = ctypes.c_int
    L.or_step_in.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _f64p, _f64p,
                             ctypes.c_double]
    L.or_step_in.restype = _sz
    L.or_spread_diffusivity.argtypes = [ctypes.c_double] * 3
    L.or_spread_diffusivity.restype = ctypes.c_double
    _u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    L.or_d8_neighbour.argtypes = [_sz, _sz, ctypes.c_int, _sz, ctypes.c_int]
    L.or_d8_neighbour.restype = ctypes.c_long
    L.or_slope_failure.argtypes = [_sz, _sz, ctypes.c_double, ctypes.c_int, _u8p,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_uint, _f64p, _f64p]
    L.or_slope_failure.restype = ctypes.c_double
    _u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    L.or_cube_solve.argtypes = [_sz, _u32p, _f64p, _f64p, ctypes.c_double, _f64p, _f64p,
                                ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.or_cube_solve.restype = _sz
    _lib = L
    return L


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


# ---- fixtures ------------------------------------------------------------------------------

class SplitMix64:
    """SplitMix64 (crates/deeptime-core/src/sphere/mod.rs:43-63) through the oracle."""

    def __init__(self, seed: int):
        self.state = ctypes.c_uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self) -> int:
        return lib().or_splitmix_next(ctypes.byref(self.state))

    def next_f64(self) -> float:
        return lib().or_splitmix_f64(ctypes.byref(self.state))


def terrain(w, h, seed):
    """`terrain`, multigrid/tests.rs:77-98."""
    z = np.zeros(w * h)
    lib().or_terrain(w, h, seed, z)
    return z


def fractal(w, h, seed):
    """`fractal`, diffusion/parity.rs:340-359."""
    z = np.zeros(w * h)
    lib().or_fractal(w, h, seed, z)
    return z


def submarine_kd(z, d0, wave_base, tidal_range, land):
    """`submarine_kd`, surface/marine.rs:44-59."""
    return lib().or_submarine_kd(z, d0, wave_base, tidal_range, land)


def shelf_kd(z, seed):
    """`shelf_kd`, multigrid/tests.rs:102-110."""
    z = _c(z)
    kd = np.zeros_like(z)
    lib().or_shelf_kd(z.size, z, seed, kd)
    return kd


# ---- operator ------------------------------------------------------------------------------

def operator(w, h, dx, wrap, kd, dt):
    """`Operator::new`, diffusion.rs:532-603 -> (east, south, diag)."""
    e, s, g = np.zeros(w * h), np.zeros(w * h), np.zeros(w * h)
    lib().or_operator(w, h, dx, int(wrap), _c(kd), dt, e, s, g)
    return e, s, g


def apply(w, h, wrap, east, south, diag, x):
    """`Operator::apply`, diffusion.rs:609-663."""
    y = np.zeros(w * h)
    lib().or_apply(w, h, int(wrap), _c(east), _c(south), _c(diag), _c(x), y)
    return y


def apply_reference(w, h, dx, wrap, kd, dt, x):
    """`apply_reference`, multigrid/tests.rs:17-43."""
    y = np.zeros(w * h)
    lib().or_apply_reference(w, h, dx, int(wrap), _c(kd), dt, _c(x), y)
    return y


# ---- first guess ---------------------------------------------------------------------------

def thomas(lo, di, up, b):
    """`thomas`, diffusion.rs:1029-1046."""
    b = _c(b).copy()
    lib().or_thomas(b.size, _c(lo), _c(di), _c(up), b)
    return b


def cyclic(lo, di, up, alpha, beta, b):
    """`cyclic`, diffusion.rs:1050-1071."""
    b = _c(b).copy()
    lib().or_cyclic(b.size, _c(lo), _c(di), _c(up), alpha, beta, b)
    return b


def lod(w, h, dx, wrap, hgt, kd, dt):
    """`lod`, diffusion.rs:370-489 (every row visited)."""
    x = _c(hgt).copy()
    lib().or_lod(w, h, dx, int(wrap), x, _c(kd), dt)
    return x


# ---- multigrid cycle -----------------------------------------------------------------------

class Hierarchy:
    """`Hierarchy`, multigrid.rs:475-497 (f32 levels of the V-cycle)."""

    def __init__(self, w, h, wrap, east, south, diag, omega=0.8, nu=2, scale=0.5, transfer=0,
                 coarsest=0):
        """transfer 1: bilinear prolongation and its transpose as the restriction; coarsest 1:
        the exact f64 solve of the coarsest level (options with no reference counterpart)."""
        self.w, self.h = w, h
        self._keep = (_c(east), _c(south), _c(diag))
        self.ptr = lib().or_hier_new_opts(w, h, int(wrap), *self._keep, omega, nu, scale,
                                          int(transfer), int(coarsest))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_hier_free(self.ptr)
            self.ptr = None

    def levels(self):
        return lib().or_hier_levels(self.ptr)

    def level(self, l):
        """Level l: dict of w, h, wrap and f32 east, south, diag, wd, mass."""
        w, h, wr = _sz(), _sz(), ctypes.c_int()
        lib().or_hier_level(self.ptr, l, ctypes.byref(w), ctypes.byref(h), ctypes.byref(wr),
                            None, None, None, None, None)
        n = w.value * h.value
        arrs = {k: np.zeros(n, np.float32) for k in ("east", "south", "diag", "wd", "mass")}
        lib().or_hier_level(self.ptr, l, None, None, None,
                            *[arrs[k].ctypes.data_as(ctypes.c_void_p)
                              for k in ("east", "south", "diag", "wd", "mass")])
        arrs.update(w=w.value, h=h.value, wrap=bool(wr.value))
        return arrs

    def precondition(self, r32, r64):
        """z = M^-1 r (one V-cycle from zero); returns (z, sum r64*z)."""
        z = np.zeros(self.w * self.h, np.float32)
        rz = lib().or_hier_precondition(self.ptr, _c(r32, np.float32), _c(r64), z)
        return z, rz


def bil_prolong(wf, hf, wc, hc, cwrap, zc):
    """The cell-centred bilinear prolongation P zc onto a wf x hf level (an option of the cycle)."""
    out = np.zeros(wf * hf, np.float32)
    lib().or_bil_prolong(wf, hf, wc, hc, int(cwrap), _c(zc, np.float32), out)
    return out


def bil_restrict(wf, hf, wrapf, wc, hc, cwrap, res):
    """P^T res onto the wc x hc level."""
    out = np.zeros(wc * hc, np.float32)
    lib().or_bil_restrict(wf, hf, int(wrapf), wc, hc, int(cwrap), _c(res, np.float32), out)
    return out


def coarse_exact(w, h, wrap, east, south, diag, r):
    """The exact f64 solve of a level's f32 operator (the cycle's exact-coarsest option)."""
    z = np.zeros(w * h, np.float32)
    lib().or_coarse_exact(w, h, int(wrap), _c(east, np.float32), _c(south, np.float32),
                          _c(diag, np.float32), _c(r, np.float32), z)
    return z


# ---- solves --------------------------------------------------------------------------------

def cg_on(w, h, dx, wrap, x, b, kd, dt, tol_m, omega=0.8, nu=2, scale=0.5):
    """`cg_on`, multigrid.rs:562-763. Returns (x, iterations)."""
    x = _c(x).copy()
    its = lib().or_cg_on(w, h, dx, int(wrap), x, _c(b), _c(kd), dt, tol_m, omega, nu, scale)
    return x, its


def solve(w, h, dx, wrap, hgt, kd, dt, tol_m):
    """`solve` (multigrid), diffusion.rs:77-116. Returns (h1, iterations)."""
    x = _c(hgt).copy()
    its = lib().or_solve(w, h, dx, int(wrap), x, _c(kd), dt, tol_m)
    return x, its


def jacobi_cg(w, h, dx, wrap, x, b, kd, dt, tol_m=1e-6):
    """Jacobi-preconditioned `cg`, diffusion.rs:927-1026 (every row)."""
    x = _c(x).copy()
    its = lib().or_jacobi_cg(w, h, dx, int(wrap), x, _c(b), _c(kd), dt, tol_m)
    return x, its


def jacobi_step(w, h, dx, wrap, hgt, kd, dt):
    """`step` (Jacobi solve with the splitting's guess), diffusion.rs:204-281, every row."""
    b = _c(hgt)
    x = lod(w, h, dx, wrap, b, kd, dt)
    return jacobi_cg(w, h, dx, wrap, x, b, kd, dt, 1e-6)


def live_of(w, h, dx, hgt, kd, dt):
    """`Live::of`, diffusion.rs:308-355: the visited strips of 8 rows (a bool array), or None when
    every row is visited."""
    ns = (h + 7) // 8
    out = (ctypes.c_ubyte * max(ns, 1))()
    on = lib().or_live_of(w, h, dx, _c(hgt), _c(kd), dt, out)
    return np.array(list(out)[:ns], dtype=bool) if on else None


def step_in(w, h, dx, wrap, hgt, kd, dt):
    """`step_in` (the Jacobi solve), diffusion.rs:262-283, skipping the rows `Live` skips.
    Returns (h1, iterations)."""
    x = _c(hgt).copy()
    its = lib().or_step_in(w, h, dx, int(wrap), x, _c(kd), dt)
    return x, its


def spread_diffusivity(z, strength, d0):
    """`spread::diffusivity`, spread.rs:76-84."""
    return lib().or_spread_diffusivity(z, strength, d0)


def reference_cg(w, h, dx, wrap, b, kd, dt):
    """`reference`, multigrid/tests.rs:46-74: plain CG to relative residual 1e-13."""
    x = np.zeros(w * h)
    lib().or_reference_cg(w, h, dx, int(wrap), _c(b), _c(kd), dt, x)
    return x


def bound(w, h, dx, wrap, kd, dt, tol):
    """`bound`, multigrid/tests.rs:157-164."""
    return lib().or_bound(w, h, dx, int(wrap), _c(kd), dt, tol)


def rel_residual(w, h, dx, wrap, kd, dt, b, x):
    """||b - A x|| / ||b|| in f64."""
    return lib().or_rel_residual(w, h, dx, int(wrap), _c(kd), dt, _c(b), _c(x))


def threads() -> int:
    return lib().or_threads()


# ---- submarine slope failure -----------------------------------------------------------------

FAILURE_RATE = 0.05  # surface/marine.rs:231
FAILURE_ITERATIONS = 40  # surface/marine.rs:232
# The reference tests' S_c, `2.5f64.to_radians().tan()` (marine.rs:721): libm's tan, as math's.
CRITICAL_SLOPE = math.tan(2.5 * (math.pi / 180.0))


def d8_neighbour(w, h, wrap, c, k):
    """`Dims::neighbour`, surface/grid.rs:46-60 (-1 for none)."""
    return lib().or_d8_neighbour(w, h, int(wrap), c, k)


def slope_failure(w, h, dx, wrap, sea, sea_level, critical_slope, rate, iterations, hgt, avail):
    """The full-sweep slope failure (`slope_failure_reference`, surface/marine.rs:595-655; the
    active-set `slope_failure`, marine.rs:292-580, equals it bit for bit).  Returns (h, avail,
    metres moved); the inputs are not changed."""
    x, a = _c(hgt).copy(), _c(avail).copy()
    m = lib().or_slope_failure(w, h, dx, int(wrap), _c(sea, np.uint8), sea_level,
                               critical_slope, rate, iterations, x, a)
    return x, a, m


def splitmix_below(rng, n):
    """`SplitMix64::below` (sphere/mod.rs:66-69): Lemire's multiply-shift into [0, n)."""
    return (rng.next_u64() * n) >> 64


_GAMMA = 0x9E3779B97F4A7C15


def _splitmix_f64_block(seed, start, n):
    """Draws start+1 .. start+n of SplitMix64(seed).next_f64 (sphere/mod.rs:52-63), vectorized:
    the state after k draws is seed + k * gamma."""
    k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(_GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 2.0**53)


def failure_world(w, hh, seed, mounds):
    """`failure_world` (surface/marine.rs:661-716): a coast, a sea floor, and either sediment
    everywhere (mounds = 0) or `mounds` sediment mounds on a crystalline floor.  The draws are the
    reference's in its order (one a cell for the heights, then the sediment).  Returns
    (h, avail, sea as uint8)."""
    n = w * hh
    f = _splitmix_f64_block(seed, 0, n)
    i = np.arange(n) % w
    h = np.where(i < w // 6, 5.0 + 50.0 * f,
                 np.where(i < w // 6 + 2, (f - 0.5) * 0.02, -3000.0 + 400.0 * f))
    avail = np.zeros(n)
    if mounds == 0:
        avail = 2000.0 * _splitmix_f64_block(seed, n, n)
        neg = np.flatnonzero(h < 0.0)
        h[neg] += 1500.0 * _splitmix_f64_block(seed, 2 * n, neg.size)
    else:
        rng = SplitMix64((seed + n * _GAMMA) & 0xFFFFFFFFFFFFFFFF)
        for _ in range(mounds):
            ci, cj = splitmix_below(rng, w), splitmix_below(rng, hh)
            r = 1 + splitmix_below(rng, 4)
            for dj in range(2 * r + 1):
                for di in range(2 * r + 1):
                    ii, j = (ci + di + r * (w - 1)) % w, cj + dj
                    if j < r or j - r >= hh:
                        continue
                    c = (j - r) * w + ii
                    if h[c] < -10.0:
                        h[c] = -100.0 - 30.0 * rng.next_f64()
                        avail[c] = 3000.0 * rng.next_f64()
    return h, avail, (h < 0.0).astype(np.uint8)


# ---- cube-grid diffusion ---------------------------------------------------------------------

def cube_solve(g, hgt, kd, dt, solver, tol_m):
    """`surface::cube::diffusion::solve` (cube/diffusion.rs:227-314) on a grid's slots, couplings
    and areas (a `dtdiff.cubegrid.CubeGrid`): solver 0 multigrid to tol_m, 1 `step_in` (Jacobi,
    the cells with a coupling, CG_TOL_M).  Returns (h1, iterations); hgt is not changed."""
    x = _c(hgt).copy()
    its = lib().or_cube_solve(g.n, np.ascontiguousarray(g.slots, dtype=np.uint32), _c(g.lap),
                              _c(g.area), g.mean_area, x, _c(kd), dt, int(solver), tol_m)
    return x, its
