"""The submarine slope failure (surface/marine.rs:292-580) on the GPU against the oracle.

The device's heights and sediment are the oracle's (the reference's full sweep, marine.rs:595-655)
bit for bit on the worlds of the reference's own bit-for-bit test (marine.rs:719-765): dense and
sparse sediment, a grid that does not wrap, grids one and two cells wide, one row; at the stage's
40 iterations and at 400.  The metres moved are a sum over the movers taken in another order (the
reference adds one mover at a time), so they agree to the rounding bound of a sum of that many
non-negative terms: (terms) * 2^-52 relative.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff.hipmem import DeviceArray, synchronize

from test_oracle_failure import CASES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace()
    yield w
    w.close()


def moved_close(m, mo, terms):
    return abs(m - mo) <= max(terms, 1) * 2.0**-52 * mo


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def gpu_failure(ws, w, hh, wrap, h, a, sea, its, dx=4800.0, crit=O.CRITICAL_SLOPE,
                rate=O.FAILURE_RATE):
    x, av = h.copy(), a.copy()
    moved, n = dtdiff.slope_failure(ws, dtdiff.Dims(w, hh, dx, wrap), sea, 0.0, x, av, crit,
                                    rate, its)
    return x, av, moved, n


@pytest.mark.parametrize("w,hh,wrap,mounds", CASES)
def test_failure_matches_the_oracle_bit_for_bit(ws, w, hh, wrap, mounds):
    for seed in (1, 2, 3):
        for its in (O.FAILURE_ITERATIONS, 400):
            print(f"failure {w}x{hh} wrap {wrap} mounds {mounds} seed {seed} its {its}",
                  flush=True)
            h, a, sea = O.failure_world(w, hh, seed, mounds)
            xo, ao, mo = O.slope_failure(w, hh, 4800.0, wrap, sea, 0.0, O.CRITICAL_SLOPE,
                                         O.FAILURE_RATE, its, h, a)
            x, av, m, n = gpu_failure(ws, w, hh, wrap, h, a, sea, its)
            assert same_bits(x, xo) and same_bits(av, ao)
            assert moved_close(m, mo, its * w * hh)
            assert 1 <= n <= its


def test_cliff_relaxes_on_the_device_as_in_the_oracle(ws):
    """marine.rs:785-836's strip, 30000 iterations: the oracle's bits."""
    print("failure cliff", flush=True)
    w, hh = 64, 3
    i = np.tile(np.arange(w), hh)
    h = np.where(i < 4, 50.0, np.where(i < 20, -100.0, -3000.0))
    avail = np.where((i >= 4) & (i < 20), 3000.0, 0.0)
    sea = (h < 0.0).astype(np.uint8)
    xo, ao, mo = O.slope_failure(w, hh, 1000.0, False, sea, 0.0, O.CRITICAL_SLOPE, 0.05, 30000,
                                 h, avail)
    x, av, m, n = gpu_failure(ws, w, hh, False, h, avail, sea, 30000, dx=1000.0, rate=0.05)
    assert same_bits(x, xo) and same_bits(av, ao) and moved_close(m, mo, n * w * hh)


def test_device_entry_and_a_larger_world(ws):
    """dtd_slope_failure_device on resident arrays gives the host entry's bits; a 1024 x 512 world
    with sediment everywhere matches the oracle."""
    print("failure 1024x512", flush=True)
    w, hh = 1024, 512
    h, a, sea = O.failure_world(w, hh, 7, 0)
    xo, ao, mo = O.slope_failure(w, hh, 4800.0, True, sea, 0.0, O.CRITICAL_SLOPE,
                                 O.FAILURE_RATE, O.FAILURE_ITERATIONS, h, a)
    hd, ad, sd = DeviceArray.from_numpy(h), DeviceArray.from_numpy(a), DeviceArray.from_numpy(sea)
    m, n = dtdiff.slope_failure_device(ws, dtdiff.Dims(w, hh, 4800.0, True), sd, 0.0, hd, ad,
                                       O.CRITICAL_SLOPE)
    synchronize()
    assert same_bits(hd.numpy(), xo) and same_bits(ad.numpy(), ao)
    assert moved_close(m, mo, n * w * hh) and n == O.FAILURE_ITERATIONS


def test_failure_arguments(ws):
    d = dtdiff.Dims(4, 3, 100.0, True)
    h, a, sea = np.full(12, -50.0), np.ones(12), np.ones(12, np.uint8)
    # No iterations and an empty grid change nothing.
    assert gpu_failure(ws, 4, 3, True, h, a, sea, 0)[2:] == (0.0, 0)
    assert dtdiff.slope_failure(ws, dtdiff.Dims(0, 3, 100.0, True), np.zeros(0, np.uint8), 0.0,
                                np.zeros(0), np.zeros(0), 0.1) == (0.0, 0)
    with pytest.raises(dtdiff.DtdError):
        dtdiff.slope_failure(ws, d, sea, 0.0, h.copy(), a.copy(), 0.1, rate=float("nan"))
    # A flat sea floor moves nothing: one iteration, which ends the loop.
    x, av, m, n = gpu_failure(ws, 4, 3, True, h, a, sea, 40)
    assert same_bits(x, h) and m == 0.0 and n == 1
