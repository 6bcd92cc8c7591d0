"""The failure seam of the device solve, through the C ABI.

The reference's `Pre` (crates/deeptime-core/src/surface/diffusion/multigrid.rs:385-472) turns a
device error into a CPU fallback and, on the next solve, drops the device state and builds it
again (multigrid.rs:443-455).  Here there is no fallback inside the library: the entry point
returns DTD_EDEVICE with a message, the caller's arrays keep their values (the caller falls back),
and the next call builds the context's buffers, levels and graphs again and solves exactly as a
fresh context does.

Device errors are simulated on the host: dtd_inject_fault skips one HIP call and reports it as
failed, so nothing runs on the GPU that a good solve would not run.  The tests step the failure
through every checked HIP call of a solve.
"""
import ctypes

import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F
from dtdiff.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

MG = dtdiff.DiffusionSolver.MULTIGRID


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def case(w, h, seed):
    b = O.terrain(w, h, seed * 7 + w)
    return b, O.shelf_kd(b, seed)


def calls_of(ws, f):
    """Checked HIP calls f makes on the context."""
    c0 = dtdiff.inject_fault(ws, -1)
    f()
    return dtdiff.inject_fault(ws, -1) - c0


def fails_at(ws, k, f):
    """Runs f with the failure armed k calls ahead; returns the error, or None if f finished
    first."""
    dtdiff.inject_fault(ws, k)
    try:
        f()
    except dtdiff.DtdError as e:
        return e
    finally:
        dtdiff.inject_fault(ws, -1)
    return None


def test_each_failing_call_of_a_host_solve_leaves_h_and_the_next_call_recovers():
    w, h, wrap, dt, tol = 130, 66, True, 2.0e6, 1e-7
    b, kd = case(w, h, 3)
    d = dtdiff.Dims(w, h, 2400.0, wrap)
    fresh = dtdiff.Workspace()
    want = b.copy()
    its_want = dtdiff.solve(fresh, d, want, kd, dt, MG, tol)
    fresh.close()
    ws = dtdiff.Workspace()
    other = dtdiff.Dims(64, 40, 2400.0, True)
    bo, kdo = case(64, 40, 5)

    def to_other():
        xo = bo.copy()
        dtdiff.solve(ws, other, xo, kdo, dt, MG, tol)  # the context then holds another shape

    x = b.copy()
    dtdiff.solve(ws, d, x, kd, dt, MG, tol)
    to_other()
    x = b.copy()
    n_fresh = calls_of(ws, lambda: dtdiff.solve(ws, d, x, kd, dt, MG, tol))  # builds the shape
    x = b.copy()
    n_kept = calls_of(ws, lambda: dtdiff.solve(ws, d, x, kd, dt, MG, tol))   # shape kept
    assert same_bits(x, want) and n_kept < n_fresh
    print(f"{n_fresh} checked calls on a new shape, {n_kept} on a kept one", flush=True)
    # Failures in a solve on the kept shape, then in one that builds the shape again (the
    # allocations, the levels, the graph capture).
    for rebuild in (False, True):
        n = n_fresh if rebuild else n_kept
        for k in range(n):
            if rebuild:
                to_other()
            x = b.copy()
            err = fails_at(ws, k, lambda: dtdiff.solve(ws, d, x, kd, dt, MG, tol))
            assert err is not None, (rebuild, k)
            assert err.code == dtdiff.DTD_EDEVICE and "injected failure" in str(err), (k, err)
            assert same_bits(x, b), (rebuild, k)  # h keeps h0
            x = b.copy()
            its = dtdiff.solve(ws, d, x, kd, dt, MG, tol)
            assert its == its_want and same_bits(x, want), (rebuild, k)
    ws.close()


def test_each_failing_call_of_a_device_solve_leaves_h_dev_and_recovers():
    d, h0, kd, dt = F.case_lognormal(w=256, h=128)
    ws = dtdiff.Workspace()
    kdd = DeviceArray.from_numpy(kd)
    hd = DeviceArray.from_numpy(h0)
    st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10)
    want = hd.numpy()
    hd = DeviceArray.from_numpy(h0)
    n = calls_of(ws, lambda: dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10))
    assert same_bits(hd.numpy(), want)
    # The last checked call is the wait after the answer's copy into h_dev was queued; every
    # failure before it leaves h_dev as it was.
    for k in range(n - 1):
        hd = DeviceArray.from_numpy(h0)
        err = fails_at(ws, k, lambda: dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10))
        assert err is not None and err.code == dtdiff.DTD_EDEVICE, k
        assert same_bits(hd.numpy(), h0), k
        st2 = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10)
        assert same_bits(hd.numpy(), want) and st2.iterations == st.iterations, k
    ws.close()


def test_failures_in_the_jacobi_and_batched_solves_recover():
    ws = dtdiff.Workspace()
    w, h = 96, 48
    b, kd = case(w, h, 7)
    d = dtdiff.Dims(w, h, 2400.0, True)
    want = b.copy()
    dtdiff.solve(ws, d, want, kd, 2e6, dtdiff.DiffusionSolver.JACOBI, 1e-6)
    x = b.copy()
    n = calls_of(ws, lambda: dtdiff.solve(ws, d, x, kd, 2e6, dtdiff.DiffusionSolver.JACOBI, 1e-6))
    for k in range(0, n, 3):
        x = b.copy()
        err = fails_at(ws, k, lambda: dtdiff.solve(ws, d, x, kd, 2e6,
                                                   dtdiff.DiffusionSolver.JACOBI, 1e-6))
        assert err is not None and err.code == dtdiff.DTD_EDEVICE and same_bits(x, b), k
        x = b.copy()
        dtdiff.solve(ws, d, x, kd, 2e6, dtdiff.DiffusionSolver.JACOBI, 1e-6)
        assert same_bits(x, want), k
    # A batch of three grids.
    shapes = [(64, 64), (96, 32), (128, 64)]
    fields = [F.case_lognormal(w=a, h=c, seed=11 + a) for (a, c) in shapes]

    def run():
        hs = [DeviceArray.from_numpy(h0) for (_, h0, _, _) in fields]
        ks = [DeviceArray.from_numpy(k_) for (_, _, k_, _) in fields]
        items = [(dd, hh, kk, dt_, 0.0) for (dd, _, _, dt_), hh, kk in zip(fields, hs, ks)]
        return hs, dtdiff.solve_batch(ws, items, 1e-10)

    hs, _ = run()
    want = [a.numpy() for a in hs]
    n = calls_of(ws, run)
    for k in range(0, n, 5):
        dtdiff.inject_fault(ws, k)
        try:
            run()
            raised = False
        except dtdiff.DtdError as e:
            raised = e.code == dtdiff.DTD_EDEVICE
        finally:
            dtdiff.inject_fault(ws, -1)
        assert raised, k
        hs, _ = run()
        assert all(same_bits(a.numpy(), wv) for a, wv in zip(hs, want)), k
    ws.close()


def test_a_grid_past_the_cell_limit_is_refused_before_any_allocation():
    """2^31 - 2^16 cells is the limit (32-bit cell indices); a larger grid is refused with
    DTD_ETOOBIG before any device memory is touched, and the resident shape stays."""
    ws = dtdiff.Workspace()
    b, kd = case(64, 40, 2)
    d = dtdiff.Dims(64, 40, 2400.0, True)
    x = b.copy()
    dtdiff.solve(ws, d, x, kd, 2e6, MG, 1e-6)
    x = b.copy()
    n_kept = calls_of(ws, lambda: dtdiff.solve(ws, d, x, kd, 2e6, MG, 1e-6))
    L = dtdiff.load()
    tiny = np.zeros(4)
    p = tiny.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    its = ctypes.c_size_t(0)
    for (w, h) in [(65536, 32768), (46341, 46341), (2**31, 1), (1, 2**40)]:
        c0 = dtdiff.inject_fault(ws, -1)
        rc = L.dtd_solve(ws._ctx, w, h, 1.0, 1, p, p, 1.0, 0, 1e-6, ctypes.byref(its))
        assert rc == dtdiff.DTD_ETOOBIG and "2^31" in L.dtd_last_error(ws._ctx).decode()
        # Only the device selection ran: no allocation, no launch.
        assert dtdiff.inject_fault(ws, -1) - c0 <= 1
        rc = L.dtd_solve_rel_device(ws._ctx, w, h, 1.0, 1, None, None, 1.0, 1e-10, None)
        assert rc == dtdiff.DTD_ETOOBIG
    assert (tiny == 0).all()
    # The resident shape was kept: solving it again makes a kept shape's calls only.
    x = b.copy()
    assert calls_of(ws, lambda: dtdiff.solve(ws, d, x, kd, 2e6, MG, 1e-6)) == n_kept
    ws.close()
