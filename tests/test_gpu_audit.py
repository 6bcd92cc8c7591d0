"""The opt-in audit of the multigrid solve (`diffusion::audit`,
crates/deeptime-core/src/surface/diffusion.rs:118-199) through the C ABI, on the GPU.

With the audit on, every multigrid solve of dtd_solve and dtd_solve_device is followed by the
Jacobi solve and by a solve to 1e-9 m RMS of the same step, and the context sums the three
differences.  Checked against the same comparison made here from the device's answers and the
oracle's Jacobi solve (`step_in`; the device's Jacobi solve takes its iterations and matches it to
1e-12 of the field) and the oracle's tight multigrid solve (the device's and the oracle's tight
solves both lie within their 1e-9 m tolerance of the exact solution, so they agree far below the
differences audited).  Also: the line's fields and number format are the reference's, the
caller's answer is the one it gets without the audit, and switching off and resetting work.
"""
import re

import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

MG = dtdiff.DiffusionSolver.MULTIGRID
TOL = 1e-6
FIELDS = ["solves", "its_multigrid", "its_jacobi"] + [
    f"{p}_{q}_m" for p in ("new_vs_old", "old_err", "new_err")
    for q in ("max", "rms", "worst_rms")]
CASES = [(160, 96, True, 1), (97, 64, False, 2), (256, 128, True, 3)]


def parse(line):
    assert line.startswith("diffusion_audit ")
    kv = dict(f.split("=") for f in line.split()[1:])
    assert list(kv) == FIELDS
    # Rust's `{:.3e}`: three decimals, an exponent with no sign padding or leading zeros.
    for k in FIELDS[3:]:
        assert re.fullmatch(r"\d\.\d{3}e(-?[1-9]\d*|0)", kv[k]), (k, kv[k])
    return {k: float(v) for k, v in kv.items()}


def case(w, h, seed):
    b = O.terrain(w, h, seed * 7 + w)
    return b, O.shelf_kd(b, seed)


def test_the_audit_line_is_the_comparison_of_the_three_solves():
    ws = dtdiff.Workspace()
    try:
        dtdiff.audit(ws, True, reset=True)
        assert dtdiff.audit_line(ws) is None
        its_new, its_old, cells = [], [], 0
        mx, worst, sq = np.zeros(3), np.zeros(3), np.zeros(3)
        for k, (w, h, wrap, seed) in enumerate(CASES):
            print(f"audit {w}x{h} wrap={wrap}", flush=True)
            b, kd = case(w, h, seed)
            d = dtdiff.Dims(w, h, 2400.0, wrap)
            if k % 2 == 0:
                x = b.copy()
                its = dtdiff.solve(ws, d, x, kd, 2e6, MG, TOL)
            else:
                hd, kdd = DeviceArray.from_numpy(b), DeviceArray.from_numpy(kd)
                its = dtdiff.solve_device(ws, d, hd, kdd, 2e6, TOL).iterations
                x = hd.numpy()
            old, io = O.step_in(w, h, 2400.0, wrap, b, kd, 2e6)
            exact, _ = O.solve(w, h, 2400.0, wrap, b, kd, 2e6, 1e-9)
            its_new.append(its), its_old.append(io)
            cells += w * h
            for q, (a, c) in enumerate([(x, old), (old, exact), (x, exact)]):
                e = np.abs(a - c)
                mx[q] = max(mx[q], e.max())
                worst[q] = max(worst[q], np.sqrt(np.mean(e * e)))
                sq[q] += float(np.sum(e * e))
        got = parse(dtdiff.audit_line(ws))
        assert got["solves"] == len(CASES)
        assert got["its_multigrid"] == pytest.approx(np.mean(its_new), abs=0.051)
        assert got["its_jacobi"] == pytest.approx(np.mean(its_old), abs=0.051)
        want = {}
        for q, p in enumerate(("new_vs_old", "old_err", "new_err")):
            want[f"{p}_max_m"] = mx[q]
            want[f"{p}_rms_m"] = np.sqrt(sq[q] / cells)
            want[f"{p}_worst_rms_m"] = worst[q]
        for k, v in want.items():
            print(f"{k}: audit {got[k]:.4e}, here {v:.4e}", flush=True)
            # Three printed digits, and the tight solves' and Jacobi answers' own differences.
            assert abs(got[k] - v) <= 2e-3 * v + 1e-7, k
            assert v > 1e-8, k  # the comparisons measure something
        # Off: solves are not audited; reset clears the sums.
        dtdiff.audit(ws, False)
        b, kd = case(64, 32, 9)
        dtdiff.solve(ws, dtdiff.Dims(64, 32, 2400.0, True), b.copy(), kd, 2e6, MG, TOL)
        assert parse(dtdiff.audit_line(ws))["solves"] == len(CASES)
        dtdiff.audit(ws, False, reset=True)
        assert dtdiff.audit_line(ws) is None
    finally:
        ws.close()


def test_the_audited_solve_returns_the_unaudited_answer():
    w, h, wrap = 200, 96, True
    b, kd = case(w, h, 4)
    d = dtdiff.Dims(w, h, 2400.0, wrap)
    out = []
    for on in (False, True):
        ws = dtdiff.Workspace()
        try:
            dtdiff.audit(ws, on, reset=True)
            x = b.copy()
            its = dtdiff.solve(ws, d, x, kd, 2e6, MG, TOL)
            hd, kdd = DeviceArray.from_numpy(b), DeviceArray.from_numpy(kd)
            st = dtdiff.solve_device(ws, d, hd, kdd, 2e6, TOL)
            out.append((x, its, hd.numpy(), st.iterations))
            assert (dtdiff.audit_line(ws) is not None) == on
        finally:
            ws.close()
    (x0, i0, y0, j0), (x1, i1, y1, j1) = out
    assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64)) and i0 == i1
    assert np.array_equal(y0.view(np.uint64), y1.view(np.uint64)) and j0 == j1
