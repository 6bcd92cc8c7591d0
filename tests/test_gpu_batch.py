"""The batched solve (dtd_solve_batch) through the C ABI, on the GPU.

Each grid of a batch is the reference's one-grid solve (diffusion.rs:77-116), so each is checked
against the oracle's solve of the same grid as the single-grid solves are: to the relative target
in the true residual, to the oracle's answer within the tolerance bound (the operator's eigenvalues
are at least 1, so ||x - x*|| <= ||r||), in about as many iterations.  Beyond that: a grid's answer
is the same bits whatever else is in the batch and on every repeat, the iteration cap holds per
grid, and bad arguments are refused.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F
from dtdiff.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace()
    yield w
    w.close()


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def shelf_case(w, h, wrap, seed):
    b = O.terrain(w, h, seed * 7 + w)
    return dtdiff.Dims(w, h, 2400.0, wrap), b, O.shelf_kd(b, seed), 2e6


# Mixed shapes: odd and even widths (both lane widths of the smoothers), wrapping and closed,
# single-level grids (<= 64 cells), one- and two-column rings, and the regional workload's kind.
MIXED = [
    ("shelf", 17, 9, True), ("shelf", 8, 8, False), ("shelf", 1, 7, True), ("shelf", 2, 5, True),
    ("shelf", 33, 33, False), ("shelf", 129, 67, True), ("lognormal", 300, 129, True),
    ("lognormal", 128, 64, False), ("constant", 96, 160, True), ("lognormal", 1024, 96, True),
]


def make(items, seed=5):
    out = []
    for k, (gen, w, h, wrap) in enumerate(items):
        if gen == "shelf":
            d, b, kd, dt = shelf_case(w, h, wrap, seed + k)
        elif gen == "constant":
            d, b, kd, dt = F.case_constant(seed=seed + k, w=w, h=h)
            d = dtdiff.Dims(w, h, d.dx, wrap)
        else:
            d, b, kd, dt = F.case_lognormal(seed=seed + k, w=w, h=h, ratio=20.0 + 10 * k)
            d = dtdiff.Dims(w, h, d.dx, wrap)
        out.append((d, b, kd, dt))
    return out


def run(ws, grids, rel_tol=0.0, tols=None):
    dev = [(DeviceArray.from_numpy(b), DeviceArray.from_numpy(kd)) for (_, b, kd, _) in grids]
    items = [(d, hd, kdd, dt, (tols[k] if tols else 0.0))
             for k, ((d, _, _, dt), (hd, kdd)) in enumerate(zip(grids, dev))]
    its, rels, st = dtdiff.solve_batch(ws, items, rel_tol)
    synchronize()
    return [hd.numpy() for hd, _ in dev], its, rels, st


def test_batch_meets_the_target_and_matches_the_oracle(ws):
    print("batch mixed rel 1e-10", flush=True)
    grids = make(MIXED)
    xs, its, rels, st = run(ws, grids, 1e-10)
    assert st.grids == len(grids) and st.cells == sum(d.cells() for d, *_ in grids)
    for (d, b, kd, dt), x, k, rel in zip(grids, xs, its, rels):
        true_rel = O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, b, x)
        assert true_rel <= 1e-10, (d, true_rel)
        assert abs(true_rel - rel) <= 1e-3 * rel + 1e-15
        tol = 1e-10 * np.linalg.norm(b) / np.sqrt(b.size)
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, b, kd, dt, tol)
        # Both within ||r|| <= 1e-10 ||b|| of the exact solution.
        assert np.linalg.norm(x - xo) <= 3e-10 * np.linalg.norm(b), d
        assert k <= its_o + 2 + (st.restarts and 8), (d, k, its_o)
        assert abs(x.sum() - b.sum()) <= 1e-9 * np.abs(b).sum()  # the put-back keeps the mass
    assert st.rel_residual_max <= 1e-10


def test_batch_tolerance_mode_stops_like_the_oracle(ws):
    """rel_tol = 0: grid k stops at its own RMS tol_m, as dtd_solve would (iterations within
    one of the oracle's, and the answer within twice the tolerance bound of the oracle's)."""
    print("batch mixed tol_m", flush=True)
    grids = make(MIXED, seed=11)
    tols = [1e-6 * np.linalg.norm(b) / np.sqrt(b.size) * (1 + k) for k, (_, b, _, _) in
            enumerate(grids)]
    xs, its, rels, st = run(ws, grids, 0.0, tols)
    assert st.restarts == 0
    for (d, b, kd, dt), x, k, tol in zip(grids, xs, its, tols):
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, b, kd, dt, tol)
        assert abs(k - its_o) <= 1, (d, k, its_o)
        bound = O.bound(d.w, d.h, d.dx, d.wrap_x, kd, dt, tol)
        assert np.sqrt(np.mean((x - xo) ** 2)) <= 2 * bound, d


def test_batch_first_guess_is_the_splitting(ws):
    """With a tolerance the first guess already meets, no iteration runs and the answer is the
    splitting's guess plus the put-back, as in the oracle's solve (its put-back sums are taken in
    strips, so the two differ by rounding only)."""
    print("batch first guess", flush=True)
    grids = make(MIXED, seed=61)
    xs, its, _, _ = run(ws, grids, 0.0, [1e30] * len(grids))
    assert its == [0] * len(grids)
    for (d, b, kd, dt), x in zip(grids, xs):
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, b, kd, dt, 1e30)
        assert its_o == 0
        assert np.abs(x - xo).max() <= 1e-13 * np.abs(xo).max(), d


# Widths that are multiples of 16 from 32 to 1024 (the regional workload's kind): the first
# guess's rows go by the one-wave partition method instead of the recurrences.
SEGMENTED = [
    ("lognormal", 32, 64, True), ("lognormal", 96, 32, False), ("constant", 160, 96, True),
    ("lognormal", 512, 64, True), ("lognormal", 1024, 40, True), ("shelf", 48, 33, False),
    ("lognormal", 272, 128, True),
]


def test_batch_rows_by_partition_give_the_splitting(ws):
    """The rows solved by the partition method in one wave (a ring's reduced system by
    Sherman-Morrison) give the splitting's guess to rounding: with a tolerance the first guess
    meets, each grid's answer is the oracle's (the recurrences) to 1e-12."""
    print("batch rows by partition", flush=True)
    grids = make(SEGMENTED, seed=71)
    xs, its, _, _ = run(ws, grids, 0.0, [1e30] * len(grids))
    assert its == [0] * len(grids)
    for (d, b, kd, dt), x in zip(grids, xs):
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, b, kd, dt, 1e30)
        assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max(), d


def test_batch_rows_by_partition_solve_to_the_target(ws):
    print("batch rows by partition, solved", flush=True)
    grids = make(SEGMENTED, seed=73)
    xs, its, rels, _ = run(ws, grids, 1e-10)
    for (d, b, kd, dt), x, k, rel in zip(grids, xs, its, rels):
        assert O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, b, x) <= 1e-10, d
        tol = 1e-10 * np.linalg.norm(b) / np.sqrt(b.size)
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, b, kd, dt, tol)
        assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo), d


def test_a_grid_answer_does_not_depend_on_the_batch(ws):
    print("batch independence", flush=True)
    grids = make(MIXED, seed=21)
    xs, its, _, _ = run(ws, grids, 1e-10)
    for k in (5, 6, 9):
        alone, its_alone, _, _ = run(ws, [grids[k]], 1e-10)
        assert same_bits(alone[0], xs[k]) and its_alone[0] == its[k]
    # Another order, and a repeat of the first batch after it (plans rebuilt and reused).
    order = list(reversed(range(len(grids))))
    xr, itr, _, _ = run(ws, [grids[k] for k in order], 1e-10)
    for j, k in enumerate(order):
        assert same_bits(xr[j], xs[k]) and itr[j] == its[k]
    again, its2, _, _ = run(ws, grids, 1e-10)
    assert all(same_bits(a, b) for a, b in zip(again, xs)) and its2 == its


def test_batch_iteration_cap_per_grid():
    print("batch cap", flush=True)
    w = dtdiff.Workspace(max_iter=3)
    try:
        grids = make(MIXED[4:], seed=31)
        tols = [1e-14] * len(grids)
        _, its, _, _ = run(w, grids, 0.0, tols)
        assert its == [3] * len(grids)
    finally:
        w.close()


def test_batch_without_graphs_is_the_same(ws, monkeypatch):
    """The loop launched kernel by kernel gives the captured graphs' bits."""
    print("batch no graph", flush=True)
    grids = make(MIXED[5:], seed=41)
    xs, its, _, _ = run(ws, grids, 1e-10)
    monkeypatch.setenv("DTD_NO_GRAPH", "1")
    w2 = dtdiff.Workspace()
    try:
        xn, itn, _, _ = run(w2, grids, 1e-10)
    finally:
        w2.close()
    assert all(same_bits(a, b) for a, b in zip(xn, xs)) and itn == its


def test_batch_rhs_guess(ws):
    """The context's first guess b itself (no splitting) reaches the same target."""
    print("batch rhs guess", flush=True)
    w = dtdiff.Workspace(guess=dtdiff.GUESS_RHS)
    try:
        grids = make(MIXED[5:], seed=51)
        xs, its, rels, st = run(w, grids, 1e-10)
        for (d, b, kd, dt), x in zip(grids, xs):
            assert O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, b, x) <= 1e-10
    finally:
        w.close()


def test_batch_arguments(ws):
    print("batch arguments", flush=True)
    d, b, kd, dt = shelf_case(17, 9, True, 3)
    hd, kdd = DeviceArray.from_numpy(b), DeviceArray.from_numpy(kd)
    # No grids; an empty grid is skipped.
    its, rels, st = dtdiff.solve_batch(ws, [], 1e-10)
    assert its == [] and st.grids == 0
    empty = dtdiff.Dims(0, 5, 1.0, True)
    its, rels, st = dtdiff.solve_batch(ws, [(empty, hd, kdd, dt, 1.0), (d, hd, kdd, dt, 1.0)],
                                       1e-10)
    assert its[0] == 0 and its[1] > 0 and st.grids == 1
    with pytest.raises(dtdiff.DtdError) as e:
        dtdiff.solve_batch(ws, [(d, 0, kdd, dt, 1.0)], 1e-10)
    assert e.value.code == -1
    with pytest.raises(dtdiff.DtdError) as e:
        dtdiff.solve_batch(ws, [(d, hd, kdd, dt, 0.0)], 0.0)  # tol_m must be positive
    assert e.value.code == -1
    w3 = dtdiff.Workspace(nu=3)
    try:
        with pytest.raises(dtdiff.DtdError) as e:
            dtdiff.solve_batch(w3, [(d, hd, kdd, dt, 1.0)], 1e-10)
        assert e.value.code == -1
    finally:
        w3.close()
