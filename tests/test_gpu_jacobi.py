"""The Jacobi-preconditioned solve (`step_in`, diffusion.rs:262-283) on the GPU against the oracle.

The device takes the conjugate gradients' sums in the reference's order (per row in ascending
column order, then the rows in order), so the iteration counts are the oracle's exactly.  Only the
put-back's two sums are taken in another order, so the answers agree to rounding of that one step
(and bit for bit where the put-back does nothing).  The grids are the reference's parity grids
(diffusion/parity.rs): rough and stiff, fractal ranges whose rows the reference skips, blobs,
awkward shapes, the iteration cap and a NaN.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff.hipmem import DeviceArray, synchronize

from test_oracle_jacobi import BLOBS, DT, DX, blobs, ranges, rough, same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace()
    yield w
    w.close()


def gpu_step(ws, w, h, wrap, hgt, kd, dt=DT):
    x = np.array(hgt, dtype=np.float64)
    its = dtdiff.solve(ws, dtdiff.Dims(w, h, DX, wrap), x, kd, dt,
                       dtdiff.DiffusionSolver.JACOBI, 1.0)
    return x, its


def close(x, xo):
    """Equal to rounding of the put-back's step: 1e-12 of the field's largest value."""
    fin = np.isfinite(xo)
    return (np.array_equal(fin, np.isfinite(x))
            and np.abs(x[fin] - xo[fin]).max(initial=0.0) <= 1e-12 * np.abs(xo[fin]).max(initial=1.0))


CASES = [("rough", 160, 96, 1, True), ("rough", 160, 96, 3, True), ("rough", 96, 40, 9, True),
         ("rough", 33, 17, 11, False), ("rough", 7, 9, 11, True), ("rough", 2, 5, 11, True),
         ("rough", 1, 9, 11, True), ("rough", 9, 1, 11, True), ("rough", 1, 1, 11, True),
         ("ranges", 160, 96, 1, True), ("ranges", 96, 61, 4, False), ("ranges", 64, 41, 5, True)]


@pytest.mark.parametrize("gen,w,h,seed,wrap", CASES)
def test_jacobi_step_matches_the_oracle(ws, gen, w, h, seed, wrap):
    print(f"jacobi {gen} {w}x{h} seed {seed} wrap {wrap}", flush=True)
    hgt, kd = rough(w, h, seed) if gen == "rough" else ranges(w, h, seed, 2.0)
    x, its = gpu_step(ws, w, h, wrap, hgt, kd)
    xo, its_o = O.step_in(w, h, DX, wrap, hgt, kd, DT)
    assert its == its_o
    assert close(x, xo)
    if gen == "rough" and w * h > 1000:
        assert its > 5


@pytest.mark.parametrize("k", range(len(BLOBS)))
def test_jacobi_step_on_blobs(ws, k):
    w, h, wrap, on = BLOBS[k]
    print(f"jacobi blob {k} {w}x{h}", flush=True)
    hgt, kd = blobs(w, h, on)
    x, its = gpu_step(ws, w, h, wrap, hgt, kd)
    xo, its_o = O.step_in(w, h, DX, wrap, hgt, kd, DT)
    assert its == its_o and close(x, xo)


def test_jacobi_iteration_cap_and_several_steps(ws):
    print("jacobi cap", flush=True)
    hgt, kd = rough(128, 64, 8)
    x, its = gpu_step(ws, 128, 64, True, hgt, kd, 2.0e8)
    xo, its_o = O.step_in(128, 64, DX, True, hgt, kd, 2.0e8)
    assert its == its_o and its >= 50 and close(x, xo)
    # Three steps in a row, each from the last answer (the oracle's own answer fed to both, so the
    # put-back's rounding does not compound).
    hgt, kd = ranges(160, 96, 2, 0.5)
    for _ in range(3):
        x, its = gpu_step(ws, 160, 96, True, hgt, kd)
        xo, its_o = O.step_in(160, 96, DX, True, hgt, kd, DT)
        assert its == its_o and close(x, xo)
        hgt = xo


def test_jacobi_nothing_to_diffuse_is_exact(ws):
    """kd = 0: no iteration, no put-back; the answer is h itself, bit for bit."""
    hgt, kd = rough(64, 32, 5)
    kd[:] = 0.0
    x, its = gpu_step(ws, 64, 32, True, hgt, kd)
    xo, its_o = O.step_in(64, 32, DX, True, hgt, kd, DT)
    assert its == 0 and its_o == 0 and same_bits(x, xo) and same_bits(x, hgt)


def test_jacobi_nan_spreads_as_in_the_every_row_solve(ws):
    hgt, kd = blobs(64, 48, BLOBS[0][3])
    hgt[20 * 64 + 11] = np.nan
    x, its = gpu_step(ws, 64, 48, True, hgt, kd)
    xo, its_o = O.step_in(64, 48, DX, True, hgt, kd, DT)
    assert its == its_o
    assert np.array_equal(np.isnan(x), np.isnan(xo)) and np.isnan(x[3 * 64])


def test_jacobi_device_entry_and_graphs(ws, monkeypatch):
    """dtd_step_device gives dtd_solve's bits; so does the loop launched kernel by kernel."""
    print("jacobi device/graph", flush=True)
    hgt, kd = rough(96, 40, 9)
    x, its = gpu_step(ws, 96, 40, True, hgt, kd)
    hd, kdd = DeviceArray.from_numpy(hgt), DeviceArray.from_numpy(kd)
    st = dtdiff.step_device(ws, dtdiff.Dims(96, 40, DX, True), hd, kdd, DT)
    synchronize()
    assert st.iterations == its and same_bits(hd.numpy(), x)
    monkeypatch.setenv("DTD_NO_GRAPH", "1")
    w2 = dtdiff.Workspace()
    try:
        xn, itn = gpu_step(w2, 96, 40, True, hgt, kd)
    finally:
        w2.close()
    assert itn == its and same_bits(xn, x)


def test_jacobi_large_grid(ws):
    """A 2048 x 1024 shore-like grid: iterations equal to the oracle's, answer within rounding."""
    print("jacobi 2048x1024", flush=True)
    w, h = 2048, 1024
    hgt = O.terrain(w, h, 17)
    kd = O.shelf_kd(hgt, 17)
    x, its = gpu_step(ws, w, h, True, hgt, kd, 2e6)
    xo, its_o = O.step_in(w, h, DX, True, hgt, kd, 2e6)
    assert its == its_o and close(x, xo)


def visited_rows(w, h, hgt, kd):
    """The rows of the strips the oracle's `Live::of` visits (every row when it skips none)."""
    lv = O.live_of(w, h, DX, hgt, kd, DT)
    if lv is None:
        return h
    return sum(min(8, h - 8 * s) for s in range(len(lv)) if lv[s])


def test_jacobi_loop_skips_the_rows_live_skips(monkeypatch):
    """The device loop visits the strips `Live::of` visits (diffusion.rs:298-357) and gives the
    bits of the loop that visits every row (DTD_NO_LIVE=1), iterations included; a NaN in a
    visited row sends it back over every row, as the reference does."""
    print("jacobi live strips", flush=True)
    cases = [("ranges", 160, 96, 1, True, 2.0), ("ranges", 96, 61, 4, False, 0.5),
             ("ranges", 64, 41, 5, True, 2.0)]
    cases += [("blob", *BLOBS[k][:2], k, BLOBS[k][2], 0.0) for k in range(len(BLOBS))]
    cases += [("nan", 64, 48, 0, True, 0.0)]
    live = dtdiff.Workspace()
    monkeypatch.setenv("DTD_NO_LIVE", "1")
    every = dtdiff.Workspace()
    skipped = 0
    try:
        for gen, w, h, seed, wrap, pct in cases:
            if gen == "ranges":
                hgt, kd = ranges(w, h, seed, pct)
            else:
                hgt, kd = blobs(w, h, BLOBS[seed][3])
                if gen == "nan":
                    hgt[20 * 64 + 11] = np.nan
            d = dtdiff.Dims(w, h, DX, wrap)
            out = []
            for wsx in (live, every):
                hd, kdd = DeviceArray.from_numpy(hgt), DeviceArray.from_numpy(kd)
                st = dtdiff.step_device(wsx, d, hd, kdd, DT)
                synchronize()
                out.append((hd.numpy(), st))
            (xl, sl), (xe, se) = out
            assert sl.iterations == se.iterations, (gen, w, h, seed)
            assert same_bits(xl, xe), (gen, w, h, seed)
            assert se.rows_visited == h
            want = h if gen == "nan" else visited_rows(w, h, hgt, kd)
            assert sl.rows_visited == want, (gen, w, h, seed, sl.rows_visited, want)
            skipped += h - sl.rows_visited
    finally:
        live.close()
        every.close()
    assert skipped > 0
