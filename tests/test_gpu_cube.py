"""The cube-grid diffusion solve (surface/cube/diffusion.rs, cube/multigrid.rs) on the GPU.

Through the C ABI (dtd_cube_new / dtd_cube_solve) against the oracle on the same grid: every pass
forms each value by the reference's expression in its order and every sum in the reference's
order (4096-cell chunks in cell order, then the chunks in order), so the iterations are the
oracle's and the heights are its bits, for the multigrid solve (face-wise first guess, the cube's
f32 V-cycle) and the Jacobi `step_in` (the cells with a coupling).
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import cubegrid as C
from dtdiff.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

R = 6.371e6


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace()
    yield w
    w.close()


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def fields(g, seed):
    rng = np.random.default_rng(seed)
    h = 1000.0 * np.sin(3 * g.unit[:, 0] + seed) * np.cos(4 * g.unit[:, 1]) + \
        50.0 * rng.standard_normal(g.cells)
    kd = 1.0e3 * np.exp(0.8 * rng.standard_normal(g.cells))
    return h, kd


@pytest.mark.parametrize("n", [2, 3, 8, 24, 33])
@pytest.mark.parametrize("solver", [dtdiff.DiffusionSolver.MULTIGRID,
                                    dtdiff.DiffusionSolver.JACOBI])
def test_cube_solve_matches_the_oracle_bit_for_bit(ws, n, solver):
    g = C.cube_grid(n, R)
    cube = dtdiff.Cube(ws, g)
    try:
        for seed, ratio in ((1, 5.0), (2, 80.0)):
            print(f"cube n {n} solver {int(solver)} seed {seed}", flush=True)
            h, kd = fields(g, seed)
            dt = ratio * g.mean_area / 1.0e3
            xo, its_o = O.cube_solve(g, h, kd, dt, int(solver), 1e-6)
            x = h.copy()
            its = cube.solve(x, kd, dt, solver, 1e-6)
            assert its == its_o and its > 0
            assert same_bits(x, xo)
    finally:
        cube.close()


def test_cube_jacobi_on_a_patch_and_the_device_entry(ws):
    """step_in on a patch of diffusivity (a subset of the cells takes part); the device entry
    gives the host entry's bits."""
    g = C.cube_grid(32, R)
    cube = dtdiff.Cube(ws, g)
    try:
        h, kd = fields(g, 5)
        kd = np.where(g.unit[:, 2] > 0.5, kd, 0.0)
        dt = 20.0 * g.mean_area / 1.0e3
        xo, its_o = O.cube_solve(g, h, kd, dt, 1, 1e-6)
        x = h.copy()
        its = cube.solve(x, kd, dt, dtdiff.DiffusionSolver.JACOBI, 1e-6)
        assert its == its_o and same_bits(x, xo)
        hd, kdd = DeviceArray.from_numpy(h), DeviceArray.from_numpy(kd)
        its_d = cube.solve_device(hd, kdd, dt, dtdiff.DiffusionSolver.JACOBI, 1e-6)
        synchronize()
        assert its_d == its and same_bits(hd.numpy(), x)
        # No diffusivity anywhere: nothing takes part, nothing changes.
        x = h.copy()
        assert cube.solve(x, np.zeros(g.cells), dt, dtdiff.DiffusionSolver.JACOBI, 1e-6) == 0
        assert same_bits(x, h)
    finally:
        cube.close()


def test_cube_harmonic_on_the_device(ws):
    """The multigrid solve of the harmonic x y at n = 64: the oracle's bits and iterations."""
    g = C.cube_grid(64, R)
    cube = dtdiff.Cube(ws, g)
    try:
        y = 1000.0 * g.unit[:, 0] * g.unit[:, 1]
        d, lam = 1.0e3, 6.0
        dt = 0.2 * R * R / (d * lam)
        kd = np.full(g.cells, d)
        xo, its_o = O.cube_solve(g, y, kd, dt, 0, 1e-9)
        x = y.copy()
        its = cube.solve(x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, 1e-9)
        assert its == its_o and same_bits(x, xo)
    finally:
        cube.close()


def test_cube_arguments(ws):
    g = C.cube_grid(4, R)
    bad = g.slots.copy()
    bad[5, 0] = C.NO_CELL
    with pytest.raises(dtdiff.DtdError):
        dtdiff.Cube(ws, C.CubeGrid(4, R, bad, g.lap, g.area, g.mean_area, g.unit))
    cube = dtdiff.Cube(ws, g)
    try:
        h, kd = fields(g, 3)
        x = h.copy()
        assert cube.solve(x, kd, 0.0, dtdiff.DiffusionSolver.MULTIGRID, 1e-6) == 0
        assert same_bits(x, h)
        with pytest.raises(dtdiff.DtdError):
            cube.solve(x, kd, 1.0, 7, 1e-6)
        # The multigrid tolerance must be finite and positive, as for the planar solve.
        for bad_tol in (0.0, -1e-6, float("inf"), float("nan")):
            with pytest.raises(dtdiff.DtdError, match="tol_m"):
                cube.solve(x, kd, 1.0, dtdiff.DiffusionSolver.MULTIGRID, bad_tol)
        assert same_bits(x, h)
    finally:
        cube.close()


@pytest.mark.parametrize("solver", [dtdiff.DiffusionSolver.MULTIGRID,
                                    dtdiff.DiffusionSolver.JACOBI])
def test_cube_solve_on_a_relabelled_table(ws, solver):
    """Two inner cells of different faces swap labels: the table is no longer the structured one,
    so every cell reads its neighbours from the table, and the solve still gives the oracle's
    bits on the same (relabelled) grid."""
    import dataclasses

    g0 = C.cube_grid(8, R)
    a, b = 0 * 64 + 3 * 8 + 4, 3 * 64 + 5 * 8 + 2
    perm = np.arange(g0.cells)
    perm[a], perm[b] = b, a
    slots = np.full_like(g0.slots, C.NO_CELL)
    has = g0.slots != C.NO_CELL
    slots[perm] = np.where(has, perm[np.where(has, g0.slots, 0)], C.NO_CELL)
    lap, area = np.empty_like(g0.lap), np.empty_like(g0.area)
    lap[perm], area[perm] = g0.lap, g0.area
    g = dataclasses.replace(g0, slots=slots.astype(np.uint32), lap=lap, area=area)
    cube = dtdiff.Cube(ws, g)
    try:
        h0, kd0 = fields(g0, 3)
        h, kd = np.empty_like(h0), np.empty_like(kd0)
        h[perm], kd[perm] = h0, kd0
        dt = 20.0 * g.mean_area / 1.0e3
        xo, its_o = O.cube_solve(g, h, kd, dt, int(solver), 1e-6)
        x = h.copy()
        its = cube.solve(x, kd, dt, solver, 1e-6)
        assert its == its_o and its > 0
        assert same_bits(x, xo)
    finally:
        cube.close()
