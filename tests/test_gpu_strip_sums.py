"""The reference's summation order on the GPU (DTD_SUMS_STRIPS): every sum of the multigrid
solve taken per strip of 8 rows in row-major order, the strips in order (multigrid.rs:596-723,
314-318), the put-back's per row (multigrid.rs:740-755), and the splitting's own recurrences as
the first guess.  With it the device's solve is the reference's bit for bit: the same iteration
count and the same h1 as the oracle (which restates the reference's loop operation by
operation), on the reference's test sizes with both boundary kinds, its seeds and (dt, tol)
pairs (multigrid/tests.rs:166-179, 445-493), and on the benchmark's fields at small sizes.  The
fast default path sums per block and agrees to ~1e-15 (tests/test_gpu_parity.py).
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (7, 1), (2, 2), (2, 5), (3, 5), (5, 3), (17, 9), (33, 33), (64, 40),
         (129, 67), (200, 96)]


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace(sums=dtdiff.SUMS_STRIPS)
    yield w
    w.close()


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("w,h", SIZES)
def test_the_reference_sizes_solve_bit_for_bit(ws, w, h):
    for wrap in (True, False):
        for seed in (1, 2):
            b = O.terrain(w, h, seed * 7 + w)
            kd = O.shelf_kd(b, seed)
            d = dtdiff.Dims(w, h, 2400.0, wrap)
            for dt, tol in [(2.0e6, 1.0e-6), (2.0e6, 1.0e-3), (2.0e4, 1.0e-6)]:
                label = f"{w}x{h} wrap {wrap} seed {seed} dt {dt} tol {tol}"
                print(label, flush=True)
                x = b.copy()
                its = dtdiff.solve(ws, d, x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, tol)
                xo, its_o = O.solve(w, h, 2400.0, wrap, b, kd, dt, tol)
                assert its == its_o, label
                assert same_bits(x, xo), label


@pytest.mark.parametrize("w,h,gen,rel", [(512, 256, "constant", 1e-10),
                                         (320, 160, "lognormal", 1e-10),
                                         (256, 128, "lognormal", 1e-12),
                                         (250, 77, "lognormal", 1e-8)])
def test_benchmark_fields_solve_bit_for_bit(ws, w, h, gen, rel):
    print(f"strips {gen} {w}x{h} rel {rel}", flush=True)
    d, h0, kd, dt = (F.case_constant if gen == "constant" else F.case_lognormal)(w=w, h=h)
    tol = rel * np.linalg.norm(h0) / np.sqrt(h0.size)
    x = h0.copy()
    its = dtdiff.solve(ws, d, x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, tol)
    xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
    assert its == its_o
    assert same_bits(x, xo)


def test_the_block_order_differs_only_by_rounding():
    """The default order on the same solve: the same answer to rounding, not the same bits
    (which is what the strip order is for)."""
    d, h0, kd, dt = F.case_lognormal(w=320, h=160)
    tol = 1e-10 * np.linalg.norm(h0) / np.sqrt(h0.size)
    xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
    fast = dtdiff.Workspace()
    try:
        x = h0.copy()
        its = dtdiff.solve(fast, d, x, kd, dt, dtdiff.DiffusionSolver.MULTIGRID, tol)
    finally:
        fast.close()
    assert abs(its - its_o) <= 2
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


def test_the_strip_order_is_refused_where_it_is_not_built():
    for kw in ({"cycle_precision": dtdiff.CYCLE_F64}, {"cycle_precision": dtdiff.CYCLE_BF16},
               {"transfer": dtdiff.TRANSFER_BILINEAR}, {"coarsest": dtdiff.COARSEST_EXACT},
               {"sums": 2}):
        with pytest.raises(dtdiff.DtdError):
            dtdiff.Workspace(sums=kw.pop("sums", dtdiff.SUMS_STRIPS), **kw)
