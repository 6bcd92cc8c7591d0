"""Parity of the HIP path (through the C ABI) with the CPU oracle, on the GPU.

Per kernel, the device does each cell's arithmetic in the reference's order, so the operator,
the first guess, the cycle's levels and the cycle's correction are compared bit for bit.  The
solve's sums are taken in another (fixed) order than the reference's strips, so whole solves are
compared as the reference compares its device cycle (multigrid/tests.rs:445-493): within the
tolerance bound of the exact solution, within twice the bound of the oracle's solve, in at most
two more iterations; and, at a tight tolerance, to a relative L2 difference of 1e-10.
"""
import os

import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (7, 1), (2, 2), (2, 5), (3, 5), (5, 3), (17, 9), (33, 33), (64, 40),
         (129, 67), (200, 96)]
SHAPES = [(w, h, wrap) for (w, h) in SIZES for wrap in (True, False)]


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace()
    yield w
    w.close()


def rms(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)))


def case(w, h, seed=3):
    b = O.terrain(w, h, seed * 7 + w)
    return b, O.shelf_kd(b, seed)


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("w,h,wrap", SHAPES)
def test_operator_and_apply_bit_for_bit(ws, w, h, wrap):
    print(f"operator {w}x{h} wrap={wrap}", flush=True)
    b, kd = case(w, h)
    d = dtdiff.Dims(w, h, 2400.0, wrap)
    e, s, g = dtdiff.operator(ws, d, kd, 2e6)
    oe, os_, og = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    assert same_bits(e, oe) and same_bits(s, os_) and same_bits(g, og)
    y = dtdiff.apply(ws, d, kd, 2e6, b)
    assert same_bits(y, O.apply(w, h, wrap, oe, os_, og, b))


def parallel_lines(w, h):
    """Whether the device solves the first guess's lines by the partition method (widths and
    heights of 16 * 2^k cells, 64 to 16384); other shapes take one recurrence per line."""
    ok = lambda n: n % 16 == 0 and 4 <= n // 16 <= 1024 and (n // 16) & (n // 16 - 1) == 0
    return ok(w) and ok(h)


@pytest.mark.parametrize("w,h,wrap", SHAPES + [(300, 129, True), (1000, 24, False)])
def test_first_guess_bit_for_bit(ws, w, h, wrap):
    """One Thomas (or Sherman-Morrison) recurrence per line: the reference's, bit for bit."""
    print(f"lod {w}x{h} wrap={wrap}", flush=True)
    assert not parallel_lines(w, h)
    b, kd = case(w, h)
    d = dtdiff.Dims(w, h, 2400.0, wrap)
    assert same_bits(dtdiff.lod(ws, d, b, kd, 2e6), O.lod(w, h, 2400.0, wrap, b, kd, 2e6))


@pytest.mark.parametrize("w,h,wrap,gen", [(64, 64, True, "shelf"), (128, 64, False, "shelf"),
                                          (1024, 512, True, "shelf"), (256, 2048, False, "shelf"),
                                          (2048, 1024, True, "lognormal"), (8192, 64, True, "lognormal"),
                                          (64, 4096, False, "lognormal"), (512, 4096, True, "lognormal"),
                                          (128, 8192, True, "lognormal")])
def test_parallel_first_guess_matches_the_recurrences(ws, w, h, wrap, gen):
    """The partition + cyclic-reduction line solver solves the same systems: equal to the
    oracle's recurrences to 1e-12 of the field's largest value (lines of 64 to 8192 cells: rows
    by cyclic reduction in LDS, columns in place by their segments' reduced systems)."""
    print(f"parallel lod {w}x{h} wrap={wrap} {gen}", flush=True)
    assert parallel_lines(w, h)
    if gen == "shelf":
        b, kd = case(w, h)
        dx, dt = 2400.0, 2e6
    else:
        _, b, kd, dt = F.case_lognormal(w=w, h=h)
        dx = 1.0
    d = dtdiff.Dims(w, h, dx, wrap)
    got = dtdiff.lod(ws, d, b, kd, dt)
    want = O.lod(w, h, dx, wrap, b, kd, dt)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


# Widths that are multiples of 64 take the fused, LDS-tiled smoothing kernels on those levels.
FUSED = [(64, 40, False), (128, 67, True), (256, 33, False), (192, 130, True), (1024, 512, True),
         (2048, 1023, False)]


@pytest.mark.parametrize("w,h,wrap", SHAPES + [(300, 129, True), (513, 257, False)] + FUSED)
def test_cycle_levels_and_correction_bit_for_bit(ws, w, h, wrap):
    """The device cycle against the CPU cycle (multigrid/tests.rs:381-440), here bit for bit."""
    print(f"cycle {w}x{h} wrap={wrap}", flush=True)
    b, kd = case(w, h, 13)
    d = dtdiff.Dims(w, h, 2400.0, wrap)
    e, s, g = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    hy = O.Hierarchy(w, h, wrap, e, s, g)
    cy = dtdiff.Cycle(ws, d, e, s, g)
    assert cy.levels() == hy.levels()
    for l in range(hy.levels()):
        a, o = cy.level(l), hy.level(l)
        assert (a["w"], a["h"], a["wrap"]) == (o["w"], o["h"], o["wrap"])
        for k in ("east", "south", "diag", "wd", "mass"):
            assert same_bits(a[k], o[k]), (l, k)
    for rnd in range(3):
        r = F.splitmix_f64(17 + w, w * h, start=rnd * w * h) - 0.5
        z, rz = cy.precondition(r.astype(np.float32), r)
        zo, rzo = hy.precondition(r.astype(np.float32), r)
        assert same_bits(z, zo)
        assert abs(rz - rzo) <= 1e-12 * abs(rzo)
        own = float(r @ z.astype(np.float64))
        assert abs(rz - own) <= 1e-9 * max(abs(own), 1e-30)


def test_the_gpu_solve_meets_its_tolerance_and_keeps_the_mass(ws):
    """multigrid/tests.rs:445-493 over the reference's SIZES, seeds and (dt, tol) pairs."""
    for (w, h) in SIZES:
        for wrap in (True, False):
            for seed in (1, 2):
                b = O.terrain(w, h, seed * 7 + w)
                kd = O.shelf_kd(b, seed)
                d = dtdiff.Dims(w, h, 2400.0, wrap)
                for dt, tol in [(2.0e6, 1.0e-6), (2.0e6, 1.0e-3), (2.0e4, 1.0e-6)]:
                    label = f"{w}x{h} wrap {wrap} seed {seed} dt {dt} tol {tol}"
                    exact = O.reference_cg(w, h, 2400.0, wrap, b, kd, dt)
                    x = b.copy()
                    its = dtdiff.solve(ws, d, x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, tol)
                    assert its < 400, label
                    lim = O.bound(w, h, 2400.0, wrap, kd, dt, tol)
                    assert rms(x, exact) <= lim, label
                    scale = max(np.abs(b).sum(), 1.0)
                    assert abs(b.sum() - x.sum()) <= 1e-12 * scale, label
                    xo, its_o = O.solve(w, h, 2400.0, wrap, b, kd, dt, tol)
                    assert rms(x, xo) <= 2.0 * lim, label
                    assert its <= its_o + 2, label
        print(f"solve {w}x{h} ok", flush=True)


@pytest.mark.parametrize("w,h,gen", [(512, 256, "constant"), (1024, 512, "lognormal"), (300, 129, "lognormal")])
def test_tight_solve_matches_the_oracle_to_1e10(ws, w, h, gen):
    """The parity bar: relative L2 between the GPU FP64 answer and the oracle's <= 1e-10 on
    identical seeded inputs, both solved to a relative residual of 1e-12."""
    print(f"tight {gen} {w}x{h}", flush=True)
    d, h0, kd, dt = (F.case_constant if gen == "constant" else F.case_lognormal)(w=w, h=h)
    tol = 1e-12 * np.linalg.norm(h0) / np.sqrt(h0.size)
    x = h0.copy()
    its = dtdiff.solve(ws, d, x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, tol)
    xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    assert abs(its - its_o) <= 2
    assert O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x) <= 2e-12


def test_solve_rel_device_meets_the_true_relative_residual(ws):
    from dtdiff.hipmem import DeviceArray
    d, h0, kd, dt = F.case_lognormal(w=2048, h=1024)
    hd = DeviceArray.from_numpy(h0)
    kdd = DeviceArray.from_numpy(kd)
    st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10)
    x = hd.numpy()
    rel = O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x)
    assert rel <= 1e-10 and abs(st.rel_residual - rel) <= 1e-3 * rel + 1e-16
    assert abs(h0.sum() - x.sum()) <= 1e-12 * np.abs(h0).sum()
    # No new extremes (the scheme is monotone for this M-matrix).
    assert x.max() <= h0.max() + 1e-6 and x.min() >= h0.min() - 1e-6


def test_empty_grid_and_bad_arguments(ws):
    d = dtdiff.Dims(0, 5, 1.0, True)
    x = np.zeros(0)
    assert dtdiff.solve(ws, d, x, np.zeros(0), 1.0, dtdiff.DiffusionSolver.MULTIGRID, 1e-6) == 0
    d = dtdiff.Dims(4, 4, 1.0, True)
    x = np.ones(16)
    for tol in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(dtdiff.DtdError) as e:
            dtdiff.solve(ws, d, x, np.ones(16), 1.0, dtdiff.DiffusionSolver.MULTIGRID, tol)
        assert e.value.code == dtdiff.DTD_EINVAL
    with pytest.raises(dtdiff.DtdError) as e:
        dtdiff.solve(ws, d, x, np.ones(16), 1.0, 7, 1e-6)  # no such solver
    assert e.value.code == dtdiff.DTD_EINVAL
    assert (x == 1).all()  # untouched on error
    # A constant field is already the answer (A 1 = 1): no iteration, nothing moves.
    its = dtdiff.solve(ws, d, x, np.ones(16), 1.0, dtdiff.DiffusionSolver.MULTIGRID, 1e-6)
    assert its == 0 and np.allclose(x, 1.0, rtol=0, atol=1e-14)


def test_results_are_repeatable_and_independent_of_workspace_history(ws):
    out = []
    for (w, h) in [(200, 96), (64, 40), (200, 96)]:
        b, kd = case(w, h, 5)
        d = dtdiff.Dims(w, h, 2400.0, True)
        x = b.copy()
        its = dtdiff.solve(ws, d, x, kd, 2e6, dtdiff.DiffusionSolver.MULTIGRID, 1e-7)
        out.append((x, its))
    assert same_bits(out[0][0], out[2][0]) and out[0][1] == out[2][1]
    fresh = dtdiff.Workspace()
    b, kd = case(200, 96, 5)
    x = b.copy()
    dtdiff.solve(fresh, dtdiff.Dims(200, 96, 2400.0, True), x, kd, 2e6,
                 dtdiff.DiffusionSolver.MULTIGRID, 1e-7)
    assert same_bits(x, out[0][0])


def test_graph_and_direct_launches_agree(ws, monkeypatch):
    b, kd = case(129, 67, 4)
    d = dtdiff.Dims(129, 67, 2400.0, True)
    x1 = b.copy()
    dtdiff.solve(ws, d, x1, kd, 2e6, dtdiff.DiffusionSolver.MULTIGRID, 1e-8)
    monkeypatch.setenv("DTD_NO_GRAPH", "1")
    direct = dtdiff.Workspace()
    x2 = b.copy()
    dtdiff.solve(direct, d, x2, kd, 2e6, dtdiff.DiffusionSolver.MULTIGRID, 1e-8)
    assert same_bits(x1, x2)


@pytest.mark.parametrize("prec,guess", [(dtdiff.CYCLE_F64, dtdiff.GUESS_LOD),
                                        (dtdiff.CYCLE_F32, dtdiff.GUESS_RHS)])
def test_cycle_variants_reach_the_same_answer(prec, guess):
    wsv = dtdiff.Workspace(cycle_precision=prec, guess=guess)
    d, h0, kd, dt = F.case_lognormal(w=512, h=256)
    tol = 1e-11 * np.linalg.norm(h0) / np.sqrt(h0.size)
    x = h0.copy()
    its = dtdiff.solve(wsv, d, x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, tol)
    xo, _ = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
    assert its < 400
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


@pytest.mark.parametrize("w,h,wrap", [(200, 96, True), (129, 67, False), (64, 40, True)])
def test_unfused_cycle_with_three_sweeps_bit_for_bit(w, h, wrap):
    """nu = 3 takes the one-sweep-per-launch kernels; the cycle still equals the oracle's."""
    wsn = dtdiff.Workspace(nu=3)
    b, kd = case(w, h, 21)
    e, s, g = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    hy = O.Hierarchy(w, h, wrap, e, s, g, 0.8, 3, 0.5)
    cy = dtdiff.Cycle(wsn, dtdiff.Dims(w, h, 2400.0, wrap), e, s, g)
    r = F.splitmix_f64(5 + w, w * h) - 0.5
    z, rz = cy.precondition(r.astype(np.float32), r)
    zo, rzo = hy.precondition(r.astype(np.float32), r)
    assert same_bits(z, zo) and abs(rz - rzo) <= 1e-12 * abs(rzo)


@pytest.mark.parametrize("w,h", [(1, 1024), (4, 2048), (8, 4096), (1, 65536), (13, 3000)])
@pytest.mark.parametrize("wrap", [True, False])
def test_tall_narrow_grids_match_the_oracle(ws, w, h, wrap):
    """A grid a few cells wide launches many more band blocks (60 columns by 4 rows a wave) than
    256-cell blocks; each block leaves a partial sum, and the buffer is sized for the larger
    count.  The solve still meets its tolerance and lies within the bound of the oracle's."""
    print(f"narrow {w}x{h} wrap={wrap}", flush=True)
    b, kd = case(w, h, 9)
    d = dtdiff.Dims(w, h, 2400.0, wrap)
    for dt, tol in [(2.0e6, 1.0e-6), (2.0e4, 1.0e-6)]:
        x = b.copy()
        its = dtdiff.solve(ws, d, x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, tol)
        xo, its_o = O.solve(w, h, 2400.0, wrap, b, kd, dt, tol)
        lim = O.bound(w, h, 2400.0, wrap, kd, dt, tol)
        assert its < 400 and its <= its_o + 2, (dt, its, its_o)
        assert rms(x, xo) <= 2.0 * lim
        assert abs(b.sum() - x.sum()) <= 1e-12 * max(np.abs(b).sum(), 1.0)
    # From h0 itself as the first guess (the splitting is exact on one-column grids), so the
    # conjugate gradients iterate through every streamed pass.
    wsr = dtdiff.Workspace(guess=dtdiff.GUESS_RHS)
    try:
        x = b.copy()
        its = dtdiff.solve(wsr, d, x, kd, 2.0e6, dtdiff.DiffusionSolver.MULTIGRID, 1e-9)
        exact = O.reference_cg(w, h, 2400.0, wrap, b, kd, 2.0e6)
        assert 1 <= its < 400
        assert rms(x, exact) <= O.bound(w, h, 2400.0, wrap, kd, 2.0e6, 1e-9)
    finally:
        wsr.close()


def test_a_cycle_is_refused_once_the_workspace_moved_on(ws):
    """dtd_cycle_precondition works on the levels dtd_cycle_new built.  A solve of another shape
    (new buffers) or of the same shape (the levels then hold the solve's operator) ends that
    cycle: precondition is refused with DTD_ESTATE instead of reading or writing arrays of the
    wrong size or using the wrong operator."""
    for other in [(200, 96), (64, 40)]:
        w, h = 64, 40
        b, kd = case(w, h, 2)
        e, s, g = O.operator(w, h, 2400.0, True, kd, 2e6)
        cy = dtdiff.Cycle(ws, dtdiff.Dims(w, h, 2400.0, True), e, s, g)
        r = F.splitmix_f64(3, w * h) - 0.5
        z, _ = cy.precondition(r.astype(np.float32), r)
        bo, kdo = case(*other, 4)
        x = bo.copy()
        dtdiff.solve(ws, dtdiff.Dims(*other, 2400.0, True), x, kdo, 2e6,
                     dtdiff.DiffusionSolver.MULTIGRID, 1e-6)
        assert cy.levels() == 0
        with pytest.raises(dtdiff.DtdError) as err:
            cy.precondition(r.astype(np.float32), r)
        assert err.value.code == dtdiff.DTD_ESTATE
        # Building the cycle again restores it, with the same bits as before.
        cy = dtdiff.Cycle(ws, dtdiff.Dims(w, h, 2400.0, True), e, s, g)
        z2, _ = cy.precondition(r.astype(np.float32), r)
        assert same_bits(z, z2)
