"""The oracle's submarine slope failure (surface/marine.rs:254-655), pinned without the reference.

The reference holds no golden vectors for it.  What pins the restatement:
- the D8 neighbour rule of grid.rs:21-60 and projection.rs:350-368 on the shapes where it bites
  (one and two cells wide, rows at the poles, a grid that does not wrap);
- a closed form: two wet cells, one above the other's critical drop, move exactly
  min(avail, rate * excess) in one iteration;
- the reference's own tests, restated: a sediment cliff relaxes to the critical slope with its
  mass kept, a crystalline step and land stay put (marine.rs:785-836); few cells move after the
  first sweep of a mound world (marine.rs:767-783); the worlds of the bit-for-bit test
  (marine.rs:719-765) move sediment in at least 20 cases;
- invariants a plausible mistake breaks: matter kept to round-off, sediment never negative, dry
  cells never changed, and the mirror image of a non-wrapping world giving the mirrored answer
  (a swapped diagonal or a wrong sign of an offset breaks it).
"""
import math

import numpy as np
import pytest

import oracle as O

CASES = [(96, 64, True, 3), (96, 64, True, 0), (160, 48, False, 6), (1, 40, True, 1),
         (2, 40, True, 1), (40, 1, True, 2), (33, 17, True, 40)]


def run(w, hh, wrap, h, avail, sea, iterations, dx=4800.0, rate=O.FAILURE_RATE,
        crit=O.CRITICAL_SLOPE):
    return O.slope_failure(w, hh, dx, wrap, sea, 0.0, crit, rate, iterations, h, avail)


def test_neighbour_rule():
    # E, NE, N, NW, W, SW, S, SE; north is row j - 1.
    assert [O.d8_neighbour(5, 4, True, 6, k) for k in range(8)] == [7, 2, 1, 0, 5, 10, 11, 12]
    # x wraps on a wrapping grid, not otherwise; rows clamp.
    assert O.d8_neighbour(5, 4, True, 9, 0) == 5 and O.d8_neighbour(5, 4, False, 9, 0) == -1
    assert O.d8_neighbour(5, 4, True, 5, 4) == 9 and O.d8_neighbour(5, 4, True, 5, 3) == 4
    assert O.d8_neighbour(5, 4, True, 2, 2) == -1 and O.d8_neighbour(5, 4, True, 17, 6) == -1
    # One cell wide: east and west wrap to the cell itself, which is skipped; the diagonals wrap
    # to the cells above and below, which count again.
    assert [O.d8_neighbour(1, 3, True, 1, k) for k in range(8)] == [-1, 0, 0, 0, -1, 2, 2, 2]
    # Two cells wide: east and west are the same other cell.
    assert O.d8_neighbour(2, 1, True, 0, 0) == 1 and O.d8_neighbour(2, 1, True, 0, 4) == 1


@pytest.mark.parametrize("avail0", [1.0, 1e6])
def test_two_cells_closed_form(avail0):
    """A 2 x 1 strip (no wrap): cell 0 lies d above cell 1, d > S_c dx.  Cell 0 would give
    rate (d - S_c dx) and has avail0; it gives min(avail0, that), and cell 1 gains it."""
    dx, crit, rate = 1000.0, 0.05, 0.05
    h0 = np.array([-100.0, -300.0])
    a0 = np.array([avail0, 0.0])
    x, a, moved = run(2, 1, False, h0, a0, np.ones(2, np.uint8), 1, dx=dx, rate=rate, crit=crit)
    give = min(avail0, rate * (200.0 - crit * dx))
    assert math.isclose(h0[0] - x[0], give, rel_tol=1e-14)
    assert math.isclose(x[1] - h0[1], give, rel_tol=1e-14)
    assert math.isclose(moved, give, rel_tol=1e-14)
    assert math.isclose(a[0], avail0 - give, rel_tol=1e-12, abs_tol=1e-12) and a[1] == give


def test_a_sediment_cliff_relaxes_to_the_critical_slope():
    """marine.rs:785-836: a 64 x 3 strip, land, a shelf on 3 km of sediment, a deep floor."""
    w, hh = 64, 3
    i = np.tile(np.arange(w), hh)
    h = np.where(i < 4, 50.0, np.where(i < 20, -100.0, -3000.0))
    avail = np.where((i >= 4) & (i < 20), 3000.0, 0.0)
    sea = (h < 0.0).astype(np.uint8)
    s_c = O.CRITICAL_SLOPE
    x, _, moved = O.slope_failure(w, hh, 1000.0, False, sea, 0.0, s_c, 0.05, 30000, h, avail)
    assert abs(h.sum() - x.sum()) < 1e-6 * abs(h.sum())
    row = x[w:2 * w]
    assert row[0] == 50.0
    steepest = max(abs(row[k] - row[k + 1]) / 1000.0 for k in range(4, w - 1))
    assert steepest < 1.2 * s_c
    assert row[24] > -2900.0 and moved > 0.0
    # A step with no sediment stays a step.
    h2 = np.where(i >= 20, -3000.0, -100.0)
    x2, _, m2 = O.slope_failure(w, hh, 1000.0, False, np.ones(w * hh, np.uint8), 0.0, s_c, 0.05,
                                30000, h2, np.zeros(w * hh))
    assert np.array_equal(x2, h2) and m2 == 0.0


def test_few_cells_move_after_the_first_sweep():
    """marine.rs:767-783."""
    h, a, sea = O.failure_world(96, 64, 2, 3)
    x, _, _ = run(96, 64, True, h, a, sea, 1)
    movers = int((x != h).sum())
    assert 0 < movers <= 96 * 64 // 16


@pytest.mark.parametrize("w,hh,wrap,mounds", CASES)
def test_worlds_keep_matter_and_move_only_wet_sediment(w, hh, wrap, mounds):
    moving = 0
    for seed in (1, 2, 3):
        for its in (O.FAILURE_ITERATIONS, 400):
            h, a, sea = O.failure_world(w, hh, seed, mounds)
            x, av, moved = run(w, hh, wrap, h, a, sea, its)
            assert abs(x.sum() - h.sum()) <= 1e-9 * max(moved, 1.0)
            assert np.all(av >= 0.0)
            dry = sea == 0
            assert np.array_equal(x[dry], h[dry]) and np.array_equal(av[dry], a[dry])
            # What a cell lost is at most the sediment it had.
            assert np.all(h - x <= a + 1e-9)
            moving += moved > 0.0
    assert moving >= 1 or w * hh < 100


def test_enough_cases_move():
    """The reference's bit-for-bit test needs at least 20 of its cases to move sediment."""
    cases = 0
    for w, hh, wrap, mounds in CASES:
        for seed in (1, 2, 3):
            for its in (O.FAILURE_ITERATIONS, 400):
                h, a, sea = O.failure_world(w, hh, seed, mounds)
                cases += run(w, hh, wrap, h, a, sea, its)[2] > 0.0
    assert cases >= 20


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("axis", [0, 1])
def test_mirror_image(seed, axis):
    """x -> w - 1 - x on a non-wrapping world swaps E/W, NE/NW and SE/SW; y -> h - 1 - y swaps
    N/S, NE/SE and NW/SW.  The answer mirrors to rounding (the neighbour sums are taken in another
    order)."""
    w, hh = 160, 48
    h, a, sea = O.failure_world(w, hh, seed, 0)
    x, av, m = run(w, hh, False, h, a, sea, 40)
    flip = lambda v: np.flip(v.reshape(hh, w), axis=axis).ravel().copy()
    xf, avf, mf = run(w, hh, False, flip(h), flip(a), flip(sea), 40)
    assert m > 0.0 and math.isclose(m, mf, rel_tol=1e-9)
    assert np.abs(flip(xf) - x).max() <= 1e-6
    assert np.abs(flip(avf) - av).max() <= 1e-6
