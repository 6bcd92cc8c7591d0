"""The allocation-free solve (dtd_solve_device_noalloc) through the C ABI, against the allocating
entry point it varies (dtd_solve_device): the same kernels on the same inputs, so the same bits
and iteration counts, on calls in a row whose shapes shrink (the fixed array keeps the last
call's contents), with no device memory taken, and a grid too large for the array refused.
"""
import ctypes

import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F
from dtdiff.hipmem import DeviceArray, hip, synchronize

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def free_device_bytes():
    h = hip()
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    h.hipMemGetInfo.restype = ctypes.c_int
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def lognormal(w, h, wrap, seed, ratio):
    d, h0, kd, dt = F.case_lognormal(seed=seed, w=w, h=h, ratio=ratio)
    return dtdiff.Dims(w, h, d.dx, wrap), h0, kd, dt


# Shrinking calls: a large grid, then smaller ones of other shapes (odd widths, a closed grid).
CASES = [(1024, 512, True, 3, 50.0), (300, 129, False, 4, 20.0), (64, 40, True, 5, 120.0)]


def test_noalloc_gives_the_allocating_entry_points_bits_on_calls_in_a_row():
    ref = dtdiff.Workspace()
    ws = dtdiff.Workspace()
    try:
        wants = []
        for (w, h, wrap, seed, ratio) in CASES:
            d, h0, kd, dt = lognormal(w, h, wrap, seed, ratio)
            tol = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(h0.size)
            hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
            st = dtdiff.solve_device(ref, d, hd, kdd, dt, tol)
            wants.append((d, h0, kd, dt, tol, hd.numpy(), st.iterations))
        for rnd in range(2):
            for (d, h0, kd, dt, tol, want, its) in wants:
                print(f"noalloc {d.w}x{d.h} wrap={d.wrap_x} round {rnd}", flush=True)
                hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
                synchronize()
                before = free_device_bytes()
                st = dtdiff.solve_device_noalloc(ws, d, hd, kdd, dt, tol)
                synchronize()
                after = free_device_bytes()
                x = hd.numpy()
                assert same_bits(x, want) and st.iterations == its, (d, st.iterations, its)
                if rnd == 1:
                    assert after == before  # nothing allocated on the device
                rel = O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x)
                assert rel <= 1e-9
    finally:
        ws.close()
        ref.close()


def test_noalloc_refuses_a_grid_too_large_and_keeps_the_contexts_own_workspace():
    ws = dtdiff.Workspace()
    try:
        d, h0, kd, dt = lognormal(512, 256, True, 6, 50.0)
        tol = 1e-9 * float(np.linalg.norm(h0)) / np.sqrt(h0.size)
        hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
        dtdiff.solve_device(ws, d, hd, kdd, dt, tol)
        want = hd.numpy()
        big = dtdiff.Dims(2048, 1024, 1.0, True)
        hb = DeviceArray.from_numpy(np.arange(big.cells(), dtype=np.float64))
        kb = DeviceArray.from_numpy(np.ones(big.cells()))
        with pytest.raises(dtdiff.DtdError) as err:
            dtdiff.solve_device_noalloc(ws, big, hb, kb, 1.0, 1e-6)
        assert err.value.code == dtdiff.DTD_ETOOBIG and "scratch" in str(err.value)
        assert same_bits(hb.numpy(), np.arange(big.cells(), dtype=np.float64))
        # A no-allocation call of another shape leaves the context's workspace as it was.
        ds, h0s, kds, dts = lognormal(200, 96, False, 7, 30.0)
        hs, ks = DeviceArray.from_numpy(h0s), DeviceArray.from_numpy(kds)
        dtdiff.solve_device_noalloc(ws, ds, hs, ks, dts, 1e-6)
        hd = DeviceArray.from_numpy(h0)
        c0 = dtdiff.inject_fault(ws, -1)
        dtdiff.solve_device(ws, d, hd, kdd, dt, tol)
        assert same_bits(hd.numpy(), want)
        x = hd.numpy()
        hd = DeviceArray.from_numpy(h0)
        n_again = dtdiff.inject_fault(ws, -1)
        dtdiff.solve_device(ws, d, hd, kdd, dt, tol)
        # No rebuild of the shape: the same number of checked HIP calls as a kept-shape solve.
        assert n_again - c0 == dtdiff.inject_fault(ws, -1) - n_again
        assert same_bits(hd.numpy(), x)
    finally:
        ws.close()
