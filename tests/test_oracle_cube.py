"""The cube-grid inputs (dtdiff.cubegrid) and the oracle's cube-grid solve
(surface/cube/diffusion.rs + cube/multigrid.rs), pinned without the reference.

The reference holds no golden vectors for the cube solve.  What pins it:
- the grid: neighbours mutual, seven at the 24 corner cells and eight elsewhere, none the cell
  itself, symmetric couplings, areas summing to 4 pi R^2 (sphere/cube.rs's own checks);
- the mathematics: one backward-Euler step of the spherical harmonic x y (l = 2) decays by
  1 / (1 + dt D l (l + 1) / R^2); the error falls at second order with n (the reference's
  `diffusion_converges_on_a_harmonic`, cube/tests.rs:398-433, asks < 0.3 of the last);
- the volume sum a h kept to 1e-12 over three steps by both solvers (cube/tests.rs:343-395);
- the same cone on three faces the grid's symmetry maps onto one another gives the same peak,
  spread and volume (cube/tests.rs:370-395 turns its cone the same way);
- a constant field stays constant; cells with no coupling keep their bits (CD:31-35);
- the multigrid and Jacobi solves agree, the multigrid in far fewer iterations.
"""
import numpy as np
import pytest

import oracle as O
from dtdiff import cubegrid as C

R = 6.371e6


@pytest.fixture(scope="module")
def g24():
    return C.cube_grid(24, R)


@pytest.mark.parametrize("n", [2, 3, 8, 33])
def test_grid_topology_and_metric(n):
    g = C.cube_grid(n, R)
    s = g.slots.astype(np.int64)
    has = g.slots != C.NO_CELL
    idx = np.arange(g.cells)
    counts = np.bincount(has.sum(1), minlength=9)
    assert counts[7] == 24 and counts[8] == g.cells - 24
    nb = np.where(has, s, 0)
    back = g.slots[nb] == idx[:, None, None].astype(np.uint32)
    assert np.all(back.any(2) | ~has)  # mutual
    k_back = np.argmax(back, axis=2)
    assert np.all((g.lap[nb, k_back] == g.lap) | ~has)  # symmetric, bit for bit
    assert not np.any((s == idx[:, None]) & has)
    assert abs(g.area.sum() / (4 * np.pi * R * R) - 1) < 1e-13
    assert np.all(g.lap[:, ::2][has[:, ::2]] > 0)  # side couplings positive


def test_harmonic_step_converges_at_second_order():
    last = np.inf
    for n in (16, 32, 64):
        g = C.cube_grid(n, R)
        y = 1000.0 * g.unit[:, 0] * g.unit[:, 1]
        d, lam = 1.0e3, 6.0
        dt = 0.2 * R * R / (d * lam)
        x, _ = O.cube_solve(g, y, np.full(g.cells, d), dt, 0, 1e-9)
        exact = y / (1.0 + dt * d * lam / (R * R))
        err = np.sqrt(np.sum(g.area * (x - exact) ** 2) / np.sum(g.area * exact ** 2))
        assert err < 0.3 * last
        last = err
    assert last < 1e-4


def cone_steps(g, centre, solver):
    h = C.cone(g, centre, np.radians(25.0), 3000.0)
    kd = np.full(g.cells, 1.0e3)
    dt = 5.0 * g.mean_area / 1.0e3
    v0 = np.sum(g.area * h)
    its = 0
    for _ in range(3):
        h, k = O.cube_solve(g, h, kd, dt, solver, 1e-3)
        its += k
    v1 = np.sum(g.area * h)
    spread = np.sum(g.area * h * np.arccos(np.clip(g.unit @ centre, -1, 1)) ** 2) / \
        np.sum(g.area * h)
    return h.max(), spread, v0, v1, its


@pytest.mark.parametrize("solver", [0, 1])
def test_cone_keeps_its_volume_and_is_the_same_turned(g24, solver):
    out = [cone_steps(g24, np.array(c, dtype=float), solver)
           for c in ((1, 0, 0), (0, 1, 0), (-1, 0, 0))]
    for peak, spread, v0, v1, its in out:
        assert abs(v1 - v0) <= 1e-12 * abs(v0) and its > 0
    peaks = np.array([o[0] for o in out])
    spreads = np.array([o[1] for o in out])
    assert np.ptp(peaks) <= 1e-6 * peaks.mean() and np.ptp(spreads) <= 1e-6 * spreads.mean()


def test_multigrid_and_jacobi_agree(g24):
    h = C.cone(g24, np.array([0.3, 0.5, 0.8]), np.radians(40.0), 2000.0)
    kd = 1.0e3 * (1.0 + g24.unit[:, 2] ** 2)
    dt = 20.0 * g24.mean_area / 1.0e3
    xm, im = O.cube_solve(g24, h, kd, dt, 0, 1e-9)
    xj, ij = O.cube_solve(g24, h, kd, dt, 1, 1e-9)
    assert np.abs(xm - xj).max() <= 1e-5 and 0 < im < ij / 2


def test_constant_stays_and_uncoupled_cells_keep_their_bits(g24):
    h = np.full(g24.cells, 123.25)
    x, _ = O.cube_solve(g24, h, np.full(g24.cells, 50.0), 3.0e9, 0, 1e-9)
    assert np.abs(x - 123.25).max() <= 1e-10
    # Jacobi on a patch of diffusivity: the cells with no coupling are left as they were.
    h = 1000.0 * np.sin(7 * g24.unit[:, 0]) * np.cos(5 * g24.unit[:, 1])
    kd = np.where(g24.unit[:, 2] > 0.6, 1.0e3, 0.0)
    x, its = O.cube_solve(g24, h, kd, 5.0 * g24.mean_area / 1.0e3, 1, 1e-9)
    has = g24.slots != C.NO_CELL
    nb = np.where(has, g24.slots, 0).astype(np.int64)
    coupled = (kd > 0) | ((kd[nb] > 0) & has).any(1)
    assert its > 0 and np.array_equal(x[~coupled], h[~coupled])
    assert abs(np.sum(g24.area * x) - np.sum(g24.area * h)) <= 1e-12 * np.sum(g24.area * abs(h))
