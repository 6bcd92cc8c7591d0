"""The surface stage's diffusivity fill before its solve (surface/mod.rs:1982-1992) on the GPU.

kd = `submarine_kd` (marine.rs:44-59) of the water depth, blended to the rock class's land value
ashore.  The device's tanh and exp are not glibc's: where they are used the oracle's values are
matched to 1e-13 relative (an ulp of tanh near -1 is amplified in 1 + tanh, and so in the marine
share of a coastal cell), and bit for bit where they are not (no tidal range, no wave base).  A fill
followed by the solve on the device matches the oracle's fill and solve.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

DIFFU = np.array([1.0, 0.3, 2.7, 0.5, 1.9, 0.05])


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace()
    yield w
    w.close()


def terrain(w, h, seed, sea_pct=60):
    z = O.fractal(w, h, seed)
    sea = float(np.sort(z)[int(z.size * sea_pct / 100)])
    cls = (np.arange(z.size) * 7919 + seed) % len(DIFFU)
    return z, sea, cls.astype(np.uint8)


def oracle_fill(z, sea, cls, kd_land, d0, wb, tr):
    return np.array([O.submarine_kd(sea - v, d0, wb, tr, kd_land * DIFFU[c])
                     for v, c in zip(z, cls)])


def gpu_fill(ws, z, sea, cls, kd_land, d0, wb, tr):
    hd, cd, dd = DeviceArray.from_numpy(z), DeviceArray.from_numpy(cls), DeviceArray.from_numpy(DIFFU)
    kd = DeviceArray(z.size)
    dtdiff.fill_kd(ws, z.size, hd, cd, dd, kd_land, sea, d0, wb, tr, kd)
    synchronize()
    return kd.numpy()


@pytest.mark.parametrize("d0,wb,tr", [(300.0, 100.0, 2.0), (300.0, 100.0, 0.0),
                                      (300.0, 0.0, 2.0), (300.0, 0.0, 0.0), (5.0, 30.0, 10.0)])
def test_fill_matches_the_oracle(ws, d0, wb, tr):
    print(f"fill d0={d0} wb={wb} tr={tr}", flush=True)
    z, sea, cls = terrain(192, 96, 7)
    got = gpu_fill(ws, z, sea, cls, 2.0, d0, wb, tr)
    want = oracle_fill(z, sea, cls, 2.0, d0, wb, tr)
    if tr == 0.0 and wb == 0.0:
        # No tanh, no exp: every operation is IEEE-rounded the same way.
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    else:
        assert np.all(np.abs(got - want) <= 1e-13 * np.abs(want) + 1e-300)
    # The blend's limits: deep sea -> d0 exp(...), high land -> kd_land * diffu[class].
    land = z > sea + 50 * max(tr, 1.0)
    assert np.allclose(got[land], 2.0 * DIFFU[cls[land]], rtol=1e-9, atol=0)


def test_fill_then_solve_matches_the_oracle(ws):
    """The stage's step on the device: fill, then the multigrid solve (the oracle fills and solves
    the same fields); the answers agree to the solve's tolerance."""
    w, h = 256, 128
    z, sea, cls = terrain(w, h, 11, sea_pct=78)
    z = z.astype(np.float64)
    d0, wb, tr, dt, dx = 300.0, 100.0, 2.0, 2.0e6, 4800.0
    hd, cd, dd = DeviceArray.from_numpy(z), DeviceArray.from_numpy(cls), DeviceArray.from_numpy(DIFFU)
    kdd = DeviceArray(z.size)
    dtdiff.fill_kd(ws, z.size, hd, cd, dd, 2.0, sea, d0, wb, tr, kdd)
    d = dtdiff.Dims(w, h, dx, True)
    st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-12)
    synchronize()
    x = hd.numpy()
    kd_o = oracle_fill(z, sea, cls, 2.0, d0, wb, tr)
    tol = 1e-12 * np.linalg.norm(z) / np.sqrt(z.size)
    xo, _ = O.solve(w, h, dx, True, z, kd_o, dt, tol)
    assert st.rel_residual <= 1e-12
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
