"""The finest level's cycle records formed from kd inside the fused smoothers, against the stored
f32 records (DTD_NO_KD=1 at workspace creation): the same answer bit for bit, in the same
iterations to the same residual.

With kd-formed records the pre-smoothing forms each cell's faces 0.5 (K_a + K_b) dt / dx^2 and
A p from kd and the direction p itself, and the post-smoothing forms the faces again, instead of
reading (east, south, diag) records and the direction pass's stored A p.  The division by dx^2 is
a reciprocal with two fma corrections there (correctly rounded); a dx whose square is not a power
of two exercises it, and a power of two (1, 4, 1/8) the exact multiply by 1 / dx^2 taken instead.  Shapes: periodic and closed, widths that are not a multiple of a band (124
columns), odd widths (one-column lanes), heights that are not a multiple of a band's rows.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F
from dtdiff.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

CASES = [(1024, 512, True, 2400.0), (1000, 300, False, 3.7), (317, 129, True, 2400.0),
         (250, 77, False, 1.0), (2048, 1030, True, 0.3), (512, 256, True, 4.0),
         (300, 100, False, 0.125)]


def solve(monkeypatch, d, h0, kd, dt, kd_records):
    monkeypatch.setenv("DTD_NO_KD", "0" if kd_records else "1")
    ws = dtdiff.Workspace()
    try:
        hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
        st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10)
        return hd.numpy(), st
    finally:
        ws.close()


@pytest.mark.parametrize("w,h,wrap,dx", CASES)
def test_kd_formed_records_match_stored_records(monkeypatch, w, h, wrap, dx):
    print(f"kd records {w}x{h} wrap={wrap} dx={dx}", flush=True)
    d = dtdiff.Dims(w, h, dx, wrap)
    kd = F.lognormal_k(w, h, w + h)
    h0 = F.spectral_noise(w, h, w * 3 + h)
    dt = 40.0 * dx * dx / float(kd.mean())
    xk, sk = solve(monkeypatch, d, h0, kd, dt, True)
    xs, ss = solve(monkeypatch, d, h0, kd, dt, False)
    assert np.array_equal(xk.view(np.uint64), xs.view(np.uint64))
    assert (sk.iterations, sk.restarts) == (ss.iterations, ss.restarts)
    assert sk.rel_residual == ss.rel_residual
    assert O.rel_residual(w, h, dx, wrap, kd, dt, h0, xk) <= 1e-10
