"""The cycle's options (no reference counterpart) on the GPU, through the C ABI: the bilinear
transfers (prolongation and its transpose as the restriction) and the exact f64 coarsest solve.

Each option does its cells' arithmetic in the oracle's order, so the cycle's correction is the
oracle's bit for bit (tests/test_oracle_cycle_options.py pins the oracle's options against dense
constructions).  Whole solves reach the same true residual and the same answer as the reference
cycle's, since the system's solution is unique.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F
from dtdiff.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

OPTIONS = [(1, 0), (0, 1), (1, 1)]
SHAPES = [(1, 7, True), (2, 5, True), (3, 5, False), (17, 9, True), (33, 33, False),
          (64, 40, True), (129, 67, True), (200, 96, False), (300, 129, True), (513, 257, False),
          (1024, 512, True), (2048, 1023, False)]


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("transfer,coarsest", OPTIONS)
@pytest.mark.parametrize("w,h,wrap", SHAPES)
def test_option_cycle_bit_for_bit(w, h, wrap, transfer, coarsest):
    print(f"options {transfer}/{coarsest} cycle {w}x{h} wrap={wrap}", flush=True)
    b = O.terrain(w, h, 13 * 7 + w)
    kd = O.shelf_kd(b, 13)
    e, s, g = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    ws = dtdiff.Workspace(transfer=transfer, coarsest=coarsest)
    try:
        cy = dtdiff.Cycle(ws, dtdiff.Dims(w, h, 2400.0, wrap), e, s, g)
        hy = O.Hierarchy(w, h, wrap, e, s, g, transfer=transfer, coarsest=coarsest)
        assert cy.levels() == hy.levels()
        for rnd in range(2):
            r = F.splitmix_f64(17 + w, w * h, start=rnd * w * h) - 0.5
            z, rz = cy.precondition(r.astype(np.float32), r)
            zo, rzo = hy.precondition(r.astype(np.float32), r)
            assert same_bits(z, zo), rnd
            assert abs(rz - rzo) <= 1e-12 * abs(rzo)
    finally:
        ws.close()


@pytest.mark.parametrize("transfer,coarsest", OPTIONS)
@pytest.mark.parametrize("prec", [dtdiff.CYCLE_F32, dtdiff.CYCLE_F64])
def test_option_solves_reach_the_reference_answer(transfer, coarsest, prec):
    for (w, h, gen) in [(1024, 512, "lognormal"), (300, 129, "lognormal"), (512, 256, "constant"),
                        (8, 7, "lognormal")]:
        print(f"options {transfer}/{coarsest} prec {prec} solve {w}x{h}", flush=True)
        d, h0, kd, dt = (F.case_lognormal if gen == "lognormal" else F.case_constant)(w=w, h=h)
        ws = dtdiff.Workspace(transfer=transfer, coarsest=coarsest, cycle_precision=prec)
        try:
            hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
            st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10)
            x = hd.numpy()
        finally:
            ws.close()
        assert O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x) <= 1e-10
        tol = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(h0.size)
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
        assert np.linalg.norm(x - xo) <= 2e-10 * np.linalg.norm(h0)
        assert abs(x.sum() - h0.sum()) <= 1e-9 * np.abs(h0).sum()
        print(f"  {st.iterations} iterations (reference cycle {its_o})", flush=True)


def test_option_combinations_that_are_refused():
    with pytest.raises(dtdiff.DtdError):
        dtdiff.Workspace(transfer=dtdiff.TRANSFER_BILINEAR, cycle_precision=dtdiff.CYCLE_BF16)
    with pytest.raises(dtdiff.DtdError):
        dtdiff.Workspace(transfer=7)
    with pytest.raises(dtdiff.DtdError):
        dtdiff.Workspace(coarsest=3)
    ws = dtdiff.Workspace(coarsest=dtdiff.COARSEST_EXACT)
    try:
        d, h0, kd, dt = F.case_lognormal(w=64, h=32)
        hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
        with pytest.raises(dtdiff.DtdError, match="reference's f32 cycle"):
            dtdiff.solve_batch(ws, [(d, hd, kdd, dt, 0.0)], 1e-10)
    finally:
        ws.close()
