"""deep-time's implicit diffusion solve on MI355X, behind the C ABI of include/dtdiff.h.

The Python surface mirrors the reference's interface for this path
(crates/deeptime-core/src/surface/diffusion.rs:77-97, grid.rs:11-37, params.rs:154):

    ws = Workspace()                      # the session's device slot + the solve's scratch
    d = Dims(w, h, dx, wrap_x=True)
    its = solve(ws, d, h, kd, dt, DiffusionSolver.MULTIGRID, tol_m)   # h: float64, in place

and adds the device-resident entry points the benchmark uses (solve_device, solve_rel_device),
plus the pieces the parity tests compare one by one (operator, apply, lod, Cycle).

There is no CPU fallback: if the HIP library is missing or a call fails, an exception says so.
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdtdiff.so")

DTD_OK = 0
DTD_EINVAL = -1
DTD_ETOOBIG = -2
DTD_EDEVICE = -3
DTD_ESTATE = -4
CYCLE_F32 = 0
CYCLE_F64 = 1
CYCLE_BF16 = 2  # the f32 cycle with its two finest levels' records stored in bf16
GUESS_LOD = 0
GUESS_RHS = 1
TRANSFER_AGGREGATE = 0  # the reference's 2x2 aggregation
TRANSFER_BILINEAR = 1   # bilinear prolongation and its transpose (an option)
COARSEST_SWEEPS = 0     # the reference's 16 sweeps on the coarsest level
COARSEST_EXACT = 1      # the coarsest level solved exactly in f64 (an option)
SUMS_BLOCKS = 0         # the multigrid solve's sums per block (fast)
SUMS_STRIPS = 1         # the reference's order: per 8-row strip, per row for the put-back

# Every symbol include/dtdiff.h declares (the CPU tests check the library exports them all).
SYMBOLS = (
    "dtd_default_opts", "dtd_create", "dtd_destroy", "dtd_last_error", "dtd_solve",
    "dtd_solve_device", "dtd_solve_rel_device", "dtd_operator", "dtd_apply", "dtd_lod",
    "dtd_cycle_new", "dtd_cycle_levels", "dtd_cycle_level", "dtd_cycle_precondition",
    "dtd_time_kernel", "dtd_solve_batch", "dtd_step_device", "dtd_fill_kd",
    "dtd_slope_failure_device", "dtd_slope_failure", "dtd_cube_new", "dtd_cube_free",
    "dtd_cube_solve", "dtd_cube_solve_device", "dtd_nn_conv3x3", "dtd_nn_conv3x3_cat",
    "dtd_nn_conv3x3_bwd_input",
    "dtd_nn_bwd_weight_scratch", "dtd_nn_conv3x3_bwd_weight", "dtd_nn_relu_bwd", "dtd_nn_pool",
    "dtd_nn_up_cat", "dtd_nn_up_cat_bwd", "dtd_nn_input", "dtd_nn_loss", "dtd_nn_guess",
    "dtd_nn_adam", "dtd_nn_to_bf16", "dtd_sync", "dtd_solve_rel_guess_device", "dtd_lod_device",
    "dtd_inject_fault", "dtd_solve_device_noalloc", "dtd_tune", "dtd_tune_plan",
    "dtd_tune_batch", "dtd_nn_block_mean", "dtd_nn_guess_coarse", "dtd_audit",
    "dtd_audit_line",
)
# The stage's slope failure constants (crates/deeptime-core/src/surface/marine.rs:231-232).
FAILURE_RATE = 0.05
FAILURE_ITERATIONS = 40
KERNEL_SWEEP, KERNEL_DIRECTION, KERNEL_UPDATE, KERNEL_RESTRICT = 0, 1, 2, 3
KERNEL_PRE, KERNEL_POST, KERNEL_COPY, KERNEL_STREAM5, KERNEL_BANDCOPY = 4, 5, 6, 7, 8


class DiffusionSolver(enum.IntEnum):
    """`DiffusionSolver`, crates/deeptime-core/src/surface/params.rs:154."""
    MULTIGRID = 0
    JACOBI = 1


@dataclass(frozen=True)
class Dims:
    """`Dims`, crates/deeptime-core/src/surface/grid.rs:11-37: w x h cells of side dx metres,
    wrapping in x when wrap_x; row-major, row 0 the north row."""
    w: int
    h: int
    dx: float
    wrap_x: bool

    def cells(self) -> int:
        return self.w * self.h


class Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("omega", ctypes.c_float), ("nu", ctypes.c_uint),
                ("scale", ctypes.c_float), ("max_iter", ctypes.c_uint),
                ("cycle_precision", ctypes.c_int), ("guess", ctypes.c_int),
                ("transfer", ctypes.c_int), ("coarsest", ctypes.c_int), ("sums", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_size_t), ("restarts", ctypes.c_size_t),
                ("rel_residual", ctypes.c_double), ("ms_setup", ctypes.c_double),
                ("ms_guess", ctypes.c_double), ("ms_cg", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("rows_visited", ctypes.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Grid(ctypes.Structure):
    """One grid of a batched solve (dtd_grid): dtd_solve_device's arguments."""
    _fields_ = [("w", ctypes.c_size_t), ("h", ctypes.c_size_t), ("dx", ctypes.c_double),
                ("wrap_x", ctypes.c_int), ("dt", ctypes.c_double), ("tol_m", ctypes.c_double),
                ("h_dev", ctypes.c_void_p), ("kd_dev", ctypes.c_void_p)]


class Failure(ctypes.Structure):
    """`FailureParams` (surface/marine.rs:254-262) as dtd_failure."""
    _fields_ = [("critical_slope", ctypes.c_double), ("rate", ctypes.c_double),
                ("iterations", ctypes.c_uint)]


class BatchStats(ctypes.Structure):
    _fields_ = [("grids", ctypes.c_size_t), ("cells", ctypes.c_size_t),
                ("iterations_max", ctypes.c_size_t), ("iterations_sum", ctypes.c_size_t),
                ("restarts", ctypes.c_size_t), ("rel_residual_max", ctypes.c_double),
                ("ms_setup", ctypes.c_double), ("ms_guess", ctypes.c_double),
                ("ms_cg", ctypes.c_double), ("ms_total", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DtdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dtdiff error {code}: {msg}")
        self.code = code


_lib = None
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def load(path: str = LIB_PATH):
    """The HIP library (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python deep-time_amd/build.py`")
    L = ctypes.CDLL(path)
    L.dtd_default_opts.argtypes = [ctypes.POINTER(Opts)]
    L.dtd_create.argtypes = [ctypes.POINTER(Opts), ctypes.POINTER(_vp)]
    L.dtd_create.restype = ctypes.c_int
    L.dtd_destroy.argtypes = [_vp]
    L.dtd_last_error.argtypes = [_vp]
    L.dtd_last_error.restype = ctypes.c_char_p
    L.dtd_solve.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _dp, _dp,
                            ctypes.c_double, ctypes.c_int, ctypes.c_double,
                            ctypes.POINTER(_sz)]
    L.dtd_solve_device.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _vp, _vp,
                                   ctypes.c_double, ctypes.c_double, ctypes.POINTER(Stats)]
    L.dtd_solve_rel_device.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _vp, _vp,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(Stats)]
    L.dtd_operator.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _dp,
                               ctypes.c_double, _dp, _dp, _dp]
    L.dtd_apply.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _dp, ctypes.c_double,
                            _dp, _dp]
    L.dtd_lod.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _dp, _dp,
                          ctypes.c_double]
    L.dtd_cycle_new.argtypes = [_vp, _sz, _sz, ctypes.c_int, _dp, _dp, _dp]
    L.dtd_cycle_levels.argtypes = [_vp]
    L.dtd_cycle_levels.restype = _sz
    L.dtd_cycle_level.argtypes = [_vp, _sz, ctypes.POINTER(_sz), ctypes.POINTER(_sz),
                                  ctypes.POINTER(ctypes.c_int), _fp, _fp, _fp, _fp, _fp]
    L.dtd_cycle_precondition.argtypes = [_vp, _fp, _dp, _fp, _dp]
    L.dtd_time_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
    L.dtd_fill_kd.argtypes = [_vp, _sz, _vp, _vp, _vp] + [ctypes.c_double] * 5 + [_vp]
    L.dtd_step_device.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _vp, _vp,
                                  ctypes.c_double, ctypes.POINTER(Stats)]
    L.dtd_solve_batch.argtypes = [_vp, _sz, ctypes.POINTER(Grid), ctypes.c_double,
                                  ctypes.POINTER(_sz), _dp, ctypes.POINTER(BatchStats)]
    L.dtd_slope_failure_device.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _vp,
                                           ctypes.c_double, ctypes.POINTER(Failure), _vp, _vp,
                                           ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_uint)]
    _u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    _u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    _f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.dtd_cube_new.argtypes = [_vp, _sz, _u32p, _f64p, _f64p, ctypes.c_double,
                               ctypes.POINTER(_vp)]
    L.dtd_cube_free.argtypes = [_vp]
    L.dtd_cube_free.restype = None
    L.dtd_cube_solve.argtypes = [_vp, _vp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_double, ctypes.POINTER(_sz)]
    L.dtd_cube_solve_device.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_double, ctypes.POINTER(_sz)]
    L.dtd_slope_failure.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _u8p,
                                    ctypes.c_double, ctypes.POINTER(Failure), _dp, _dp,
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_uint)]
    _i, _d, _f = ctypes.c_int, ctypes.c_double, ctypes.c_float
    L.dtd_nn_conv3x3.argtypes = [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _i]
    L.dtd_nn_conv3x3_cat.argtypes = [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp,
                                     _vp, _i]
    L.dtd_nn_conv3x3_bwd_input.argtypes = [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]
    L.dtd_nn_bwd_weight_scratch.argtypes = [_i, _i, _i, _i, _i, ctypes.POINTER(_sz)]
    L.dtd_nn_bwd_weight_scratch.restype = _sz
    L.dtd_nn_conv3x3_bwd_weight.argtypes = [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _sz, _vp,
                                            _sz, _vp, _vp]
    L.dtd_nn_relu_bwd.argtypes = [_vp, _sz, _vp, _vp, _vp]
    L.dtd_nn_pool.argtypes = [_vp, _i, _i, _i, _i, _vp, _vp]
    L.dtd_nn_up_cat.argtypes = [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp]
    L.dtd_nn_up_cat_bwd.argtypes = [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]
    L.dtd_nn_input.argtypes = [_vp, _sz, _vp, _vp, _vp, _d, _d, _d, _vp]
    L.dtd_nn_loss.argtypes = [_vp, _sz, _vp, _i, _vp, _vp, _d, _vp, _vp,
                              ctypes.POINTER(ctypes.c_double)]
    L.dtd_nn_guess.argtypes = [_vp, _sz, _vp, _vp, _i, _d, _vp]
    L.dtd_lod_device.argtypes = [_vp, _sz, _sz, _d, _i, _vp, _vp, _d, _vp]
    L.dtd_nn_block_mean.argtypes = [_vp, _sz, _sz, _i, _i, _vp, _vp]
    L.dtd_nn_guess_coarse.argtypes = [_vp, _sz, _sz, _i, _vp, _vp, _i, _d, _vp]
    L.dtd_nn_adam.argtypes = [_vp, _sz, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _i, _vp]
    L.dtd_nn_to_bf16.argtypes = [_vp, _sz, _vp, _vp]
    L.dtd_sync.argtypes = [_vp]
    L.dtd_solve_rel_guess_device.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _vp,
                                             _vp, ctypes.c_double, ctypes.c_double, _vp,
                                             ctypes.POINTER(Stats)]
    L.dtd_solve_device_noalloc.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _vp,
                                           _vp, ctypes.c_double, ctypes.c_double,
                                           ctypes.POINTER(Stats)]
    L.dtd_solve_device_noalloc.restype = ctypes.c_int
    L.dtd_tune.argtypes = [_vp, _sz, _sz, ctypes.c_double, ctypes.c_int, _vp, _vp,
                           ctypes.c_double, ctypes.c_int, _dp, _dp]
    L.dtd_tune.restype = ctypes.c_int
    L.dtd_tune_plan.argtypes = [_vp, _sz, _sz, ctypes.c_int, ctypes.c_char_p, _sz]
    L.dtd_tune_plan.restype = ctypes.c_int
    L.dtd_tune_batch.argtypes = [_vp, _sz, ctypes.POINTER(Grid), ctypes.c_double, ctypes.c_int,
                                 _dp, _dp]
    L.dtd_tune_batch.restype = ctypes.c_int
    L.dtd_inject_fault.argtypes = [_vp, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    L.dtd_inject_fault.restype = ctypes.c_int
    L.dtd_audit.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    L.dtd_audit.restype = ctypes.c_int
    L.dtd_audit_line.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t]
    L.dtd_audit_line.restype = ctypes.c_int
    for name in ("dtd_solve", "dtd_solve_device", "dtd_solve_rel_device", "dtd_operator",
                 "dtd_apply", "dtd_lod", "dtd_cycle_new", "dtd_cycle_level",
                 "dtd_cycle_precondition", "dtd_time_kernel"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def default_opts() -> Opts:
    o = Opts()
    load().dtd_default_opts(ctypes.byref(o))
    return o


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _ptr(t):
    """Device address of a DeviceArray (anything with data_ptr()) or an int address."""
    return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())


class Workspace:
    """The solve's kept state (`Workspace`, diffusion.rs:213-260) together with the session's
    device slot (`Ambient::kept`, compute.rs:110-122): a dtdiff context whose device buffers and
    levels stay resident between solves of one grid shape."""

    def __init__(self, device: int = 0, opts: Opts | None = None, **kw):
        L = load()
        o = opts if opts is not None else default_opts()
        o.device = device
        for k, v in kw.items():
            setattr(o, k, v)
        self.opts = o
        self._ctx = _vp()
        rc = L.dtd_create(ctypes.byref(o), ctypes.byref(self._ctx))
        if rc != DTD_OK:
            raise DtdError(rc, "dtd_create failed")

    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            load().dtd_destroy(self._ctx)
            self._ctx = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != DTD_OK:
            raise DtdError(rc, load().dtd_last_error(self._ctx).decode())


def solve(ws: Workspace, d: Dims, h: np.ndarray, kd: np.ndarray, dt: float,
          solver: DiffusionSolver, tol_m: float) -> int:
    """One implicit step (`diffusion::solve`, diffusion.rs:77-97): h (float64, w*h) holds h0 on
    entry and h1 on return; returns the conjugate-gradient iterations."""
    if not (isinstance(h, np.ndarray) and h.dtype == np.float64 and h.flags.c_contiguous):
        raise TypeError("h must be a C-contiguous float64 array (it is solved in place)")
    if h.size != d.cells() or np.size(kd) != d.cells():
        raise ValueError("h and kd must have w*h cells")
    kd, kdp = _f64(kd)
    its = _sz(0)
    ws.check(load().dtd_solve(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), h.ctypes.data_as(_dp),
                              kdp, dt, int(solver), tol_m, ctypes.byref(its)))
    return its.value


def solve_device(ws: Workspace, d: Dims, h_dev, kd_dev, dt: float, tol_m: float) -> Stats:
    """The solve on device arrays (DeviceArray or addresses of w*h doubles), h in place."""
    st = Stats()
    ws.check(load().dtd_solve_device(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), _ptr(h_dev),
                                     _ptr(kd_dev), dt, tol_m, ctypes.byref(st)))
    return st


def solve_device_noalloc(ws: Workspace, d: Dims, h_dev, kd_dev, dt: float, tol_m: float) -> Stats:
    """solve_device with the workspace in the library's fixed scratch array (no allocation)."""
    st = Stats()
    ws.check(load().dtd_solve_device_noalloc(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), _ptr(h_dev),
                                             _ptr(kd_dev), dt, tol_m, ctypes.byref(st)))
    return st


def solve_rel_device(ws: Workspace, d: Dims, h_dev, kd_dev, dt: float, rel_tol: float) -> Stats:
    """The solve to a true f64 relative residual ||h0 - A h1|| / ||h0|| <= rel_tol."""
    st = Stats()
    ws.check(load().dtd_solve_rel_device(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), _ptr(h_dev),
                                         _ptr(kd_dev), dt, rel_tol, ctypes.byref(st)))
    return st


def solve_rel_guess_device(ws: Workspace, d: Dims, h_dev, kd_dev, dt: float, rel_tol: float,
                           x0_dev) -> Stats:
    """solve_rel_device started from the guess x0_dev instead of the splitting."""
    st = Stats()
    ws.check(load().dtd_solve_rel_guess_device(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x),
                                               _ptr(h_dev), _ptr(kd_dev), dt, rel_tol,
                                               _ptr(x0_dev), ctypes.byref(st)))
    return st


def lod_device(ws: Workspace, d: Dims, h_dev, kd_dev, dt: float, out_dev) -> None:
    """The splitting's first guess of h_dev into out_dev (device arrays), on the context's
    stream."""
    ws.check(load().dtd_lod_device(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), _ptr(h_dev),
                                   _ptr(kd_dev), dt, _ptr(out_dev)))


def fill_kd(ws: Workspace, n: int, h_dev, class_dev, diffu_dev, kd_land: float, sea: float,
            d0: float, wave_base: float, tidal_range: float, kd_dev) -> None:
    """The surface stage's kd fill (surface/mod.rs:1982-1992, `submarine_kd`) on device arrays."""
    ws.check(load().dtd_fill_kd(ws._ctx, n, _ptr(h_dev), _ptr(class_dev), _ptr(diffu_dev),
                                kd_land, sea, d0, wave_base, tidal_range, _ptr(kd_dev)))


def slope_failure(ws: Workspace, d: Dims, sea, sea_level: float, h, avail,
                  critical_slope: float, rate: float = FAILURE_RATE,
                  iterations: int = FAILURE_ITERATIONS):
    """`marine::slope_failure` (surface/marine.rs:292-580) on host arrays: h and avail (f64,
    C-contiguous) are updated in place; sea is one 0/1 byte a cell.  Returns (metres moved,
    iterations run)."""
    for a in (h, avail):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
            raise TypeError("h and avail must be C-contiguous float64 arrays (changed in place)")
        if a.size != d.cells() or np.size(sea) != d.cells():
            raise ValueError("sea, h and avail must have w*h cells")
    p = Failure(critical_slope, rate, iterations)
    moved, its = ctypes.c_double(), ctypes.c_uint()
    ws.check(load().dtd_slope_failure(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x),
                                      np.ascontiguousarray(sea, dtype=np.uint8), sea_level,
                                      ctypes.byref(p), h.ctypes.data_as(_dp),
                                      avail.ctypes.data_as(_dp), ctypes.byref(moved),
                                      ctypes.byref(its)))
    return moved.value, its.value


def slope_failure_device(ws: Workspace, d: Dims, sea_dev, sea_level: float, h_dev, avail_dev,
                         critical_slope: float, rate: float = FAILURE_RATE,
                         iterations: int = FAILURE_ITERATIONS):
    """The slope failure on device arrays (sea: one byte a cell).  Returns (moved, iterations)."""
    p = Failure(critical_slope, rate, iterations)
    moved, its = ctypes.c_double(), ctypes.c_uint()
    ws.check(load().dtd_slope_failure_device(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x),
                                             _ptr(sea_dev), sea_level, ctypes.byref(p),
                                             _ptr(h_dev), _ptr(avail_dev), ctypes.byref(moved),
                                             ctypes.byref(its)))
    return moved.value, its.value


class Cube:
    """A cube grid on the device (dtd_cube_new): the slots, couplings and areas `CubeGrid` hands
    the cube's diffusion solve (crates/deeptime-core/src/surface/cube/diffusion.rs), e.g. a
    `dtdiff.cubegrid.CubeGrid`.  solve() is `cube::diffusion::solve` on host arrays (h in place),
    solve_device() on device arrays."""

    def __init__(self, ws: Workspace, grid):
        self.ws, self.n = ws, grid.n
        self._g = _vp()
        ws.check(load().dtd_cube_new(ws._ctx, grid.n,
                                     np.ascontiguousarray(grid.slots, dtype=np.uint32),
                                     np.ascontiguousarray(grid.lap, dtype=np.float64),
                                     np.ascontiguousarray(grid.area, dtype=np.float64),
                                     grid.mean_area, ctypes.byref(self._g)))

    def solve(self, h: np.ndarray, kd, dt: float, solver: DiffusionSolver, tol_m: float) -> int:
        if not (isinstance(h, np.ndarray) and h.dtype == np.float64 and h.flags.c_contiguous):
            raise TypeError("h must be a C-contiguous float64 array (it is solved in place)")
        if h.size != 6 * self.n * self.n or np.size(kd) != h.size:
            raise ValueError("h and kd must have 6 n^2 cells")
        kd, kdp = _f64(kd)
        its = _sz(0)
        self.ws.check(load().dtd_cube_solve(self.ws._ctx, self._g, h.ctypes.data_as(_dp), kdp,
                                            dt, int(solver), tol_m, ctypes.byref(its)))
        return its.value

    def solve_device(self, h_dev, kd_dev, dt: float, solver: DiffusionSolver,
                     tol_m: float) -> int:
        its = _sz(0)
        self.ws.check(load().dtd_cube_solve_device(self.ws._ctx, self._g, _ptr(h_dev),
                                                   _ptr(kd_dev), dt, int(solver), tol_m,
                                                   ctypes.byref(its)))
        return its.value

    def close(self):
        if self._g:
            load().dtd_cube_free(self._g)
            self._g = _vp()


def step_device(ws: Workspace, d: Dims, h_dev, kd_dev, dt: float) -> Stats:
    """`step_in` (the Jacobi-preconditioned solve, diffusion.rs:262-283) on device arrays."""
    st = Stats()
    ws.check(load().dtd_step_device(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), _ptr(h_dev),
                                    _ptr(kd_dev), dt, ctypes.byref(st)))
    return st


class BatchGrids:
    """A batch's dtd_grid records built once, for callers that solve the same grids (the same
    device arrays) again and again: filling 512 records from Python costs about a millisecond,
    which a C caller's loop would not."""

    def __init__(self, items):
        self.n = len(items)
        self.arr = (Grid * max(self.n, 1))()
        for k, (d, h_dev, kd_dev, dt, tol_m) in enumerate(items):
            self.arr[k] = Grid(d.w, d.h, d.dx, int(d.wrap_x), dt, tol_m, _ptr(h_dev).value,
                               _ptr(kd_dev).value)


def solve_batch(ws: Workspace, items, rel_tol: float = 0.0):
    """Many grids at once (dtd_solve_batch).  items: (Dims, h_dev, kd_dev, dt, tol_m) per grid,
    device arrays solved in place, or a BatchGrids built from such a list.  rel_tol > 0 takes
    every grid to that true relative residual (tol_m ignored).  Returns (iterations per grid,
    relative residual per grid, BatchStats)."""
    grids = items if isinstance(items, BatchGrids) else BatchGrids(items)
    n, arr = grids.n, grids.arr
    its = (ctypes.c_size_t * max(n, 1))()
    rels = (ctypes.c_double * max(n, 1))()
    st = BatchStats()
    ws.check(load().dtd_solve_batch(ws._ctx, n, arr, rel_tol, its, rels, ctypes.byref(st)))
    return list(its)[:n], list(rels)[:n], st


def operator(ws: Workspace, d: Dims, kd, dt):
    """`Operator::new` (diffusion.rs:532-603) on the device -> (east, south, diag)."""
    n = d.cells()
    kd, kdp = _f64(kd)
    e, s, g = np.zeros(n), np.zeros(n), np.zeros(n)
    ws.check(load().dtd_operator(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), kdp, dt,
                                 e.ctypes.data_as(_dp), s.ctypes.data_as(_dp),
                                 g.ctypes.data_as(_dp)))
    return e, s, g


def apply(ws: Workspace, d: Dims, kd, dt, x):
    """(I + dt L) x on the device (`Operator::apply`, diffusion.rs:652-663)."""
    kd, kdp = _f64(kd)
    x, xp = _f64(x)
    y = np.zeros(d.cells())
    ws.check(load().dtd_apply(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), kdp, dt, xp,
                              y.ctypes.data_as(_dp)))
    return y


def lod(ws: Workspace, d: Dims, h, kd, dt):
    """The splitting's first guess (`lod`, diffusion.rs:370-489) on the device."""
    kd, kdp = _f64(kd)
    x = np.array(h, dtype=np.float64, copy=True)
    ws.check(load().dtd_lod(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), x.ctypes.data_as(_dp), kdp,
                            dt))
    return x


class Cycle:
    """The device V-cycle of one operator (`DeviceCycle`, multigrid/gpu.rs:562-714)."""

    def __init__(self, ws: Workspace, d: Dims, east, south, diag):
        self.ws, self.d = ws, d
        e, ep = _f64(east)
        s, sp = _f64(south)
        g, gp = _f64(diag)
        ws.check(load().dtd_cycle_new(ws._ctx, d.w, d.h, int(d.wrap_x), ep, sp, gp))

    def levels(self) -> int:
        return load().dtd_cycle_levels(self.ws._ctx)

    def level(self, l: int):
        w, h, wr = _sz(), _sz(), ctypes.c_int()
        L = load()
        self.ws.check(L.dtd_cycle_level(self.ws._ctx, l, ctypes.byref(w), ctypes.byref(h),
                                        ctypes.byref(wr), None, None, None, None, None))
        n = w.value * h.value
        out = {k: np.zeros(n, np.float32) for k in ("east", "south", "diag", "wd", "mass")}
        self.ws.check(L.dtd_cycle_level(self.ws._ctx, l, None, None, None,
                                        *[out[k].ctypes.data_as(_fp)
                                          for k in ("east", "south", "diag", "wd", "mass")]))
        out.update(w=w.value, h=h.value, wrap=bool(wr.value))
        return out

    def precondition(self, r32, r64):
        """z = M^-1 r by one V-cycle from zero; returns (z, sum r64 * z)."""
        r32, r32p = _f32(r32)
        r64, r64p = _f64(r64)
        z = np.zeros(self.d.cells(), np.float32)
        rz = ctypes.c_double()
        self.ws.check(load().dtd_cycle_precondition(self.ws._ctx, r32p, r64p,
                                                    z.ctypes.data_as(_fp), ctypes.byref(rz)))
        return z, rz.value


def tune(ws: Workspace, d: Dims, h_dev, kd_dev, dt: float, reps: int = 5):
    """dtd_tune: searches the band heights and the bottom size for d's shape on the solve of
    (h_dev, kd_dev, dt) and keeps the plan in the workspace; returns (ms with the built-in rules,
    ms with the plan kept, the plan as a dict)."""
    import json

    a, b = ctypes.c_double(), ctypes.c_double()
    ws.check(load().dtd_tune(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), _ptr(h_dev), _ptr(kd_dev),
                             dt, reps, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value, tune_plan(ws, d)


def tune_batch(ws: Workspace, items, rel_tol: float = 1e-10, reps: int = 3):
    """dtd_tune_batch over the grids of `items` (as for solve_batch); their h arrays keep h0.
    Returns (ms with the built-in band heights, ms with the kept ones)."""
    n = len(items)
    arr = (Grid * max(n, 1))()
    for k, (d, h_dev, kd_dev, dt, tol_m) in enumerate(items):
        arr[k] = Grid(d.w, d.h, d.dx, int(d.wrap_x), dt, tol_m, _ptr(h_dev).value,
                      _ptr(kd_dev).value)
    a, b = ctypes.c_double(), ctypes.c_double()
    ws.check(load().dtd_tune_batch(ws._ctx, n, arr, rel_tol, reps, ctypes.byref(a),
                                   ctypes.byref(b)))
    return a.value, b.value


def tune_plan(ws: Workspace, d: Dims) -> dict:
    """The plan dtd_tune keeps for d's shape ({} when none)."""
    import json

    buf = ctypes.create_string_buffer(4096)
    ws.check(load().dtd_tune_plan(ws._ctx, d.w, d.h, int(d.wrap_x), buf, 4096))
    return json.loads(buf.value.decode())


def forget_plan(ws: Workspace, d: Dims) -> None:
    """Drops the tuned plan of d's shape (the built-in rules from then on)."""
    ws.check(load().dtd_tune(ws._ctx, d.w, d.h, d.dx, int(d.wrap_x), None, None, 1.0, 0, None,
                             None))


def inject_fault(ws: Workspace, after: int) -> int:
    """Arms the simulated device error of dtd_inject_fault (after < 0 disarms it); returns the
    number of checked HIP calls the context has made so far."""
    calls = ctypes.c_long()
    ws.check(load().dtd_inject_fault(ws._ctx, after, ctypes.byref(calls)))
    return calls.value


def audit(ws: Workspace, on: bool = True, reset: bool = False) -> None:
    """Switches the opt-in audit of the multigrid solve (`diffusion::audit`, diffusion.rs:118-199;
    on from the start when DEEPTIME_DIFFUSION_AUDIT is set): each solve's answer is compared with
    the Jacobi solve's and with a solve to 1e-9 m RMS.  reset clears the sums."""
    ws.check(load().dtd_audit(ws._ctx, int(on), int(reset)))


def audit_line(ws: Workspace) -> str | None:
    """`audit::audit_line` (diffusion.rs:179-199): the sums so far, or None before any audited
    solve."""
    buf = ctypes.create_string_buffer(1024)
    rc = load().dtd_audit_line(ws._ctx, buf, len(buf))
    if rc == DTD_ESTATE:
        return None
    ws.check(rc)
    return buf.value.decode()


def time_kernel(ws: Workspace, kernel: int, reps: int = 20):
    """Average duration (ms, HIP events on the context's stream) and compulsory bytes per launch
    of one of the solve's kernels over the resident data of the last solve."""
    ms, b = ctypes.c_double(), ctypes.c_double()
    ws.check(load().dtd_time_kernel(ws._ctx, kernel, reps, ctypes.byref(ms), ctypes.byref(b)))
    return ms.value, b.value
