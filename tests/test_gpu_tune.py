"""The launch tuner (dtd_tune) through the C ABI, on the GPU.

The tuner changes only how the solve is cut into launches (band heights, the bottom's size), never
a cell's arithmetic: the cycle's correction keeps the oracle's bits under a tuned plan, and a tuned
solve meets the same true-residual target and matches the oracle's answer.  The plan is kept per
shape, re-applied when the workspace takes the shape again, and kept only if it is not slower.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F
from dtdiff.hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("w,h", [(1024, 512), (300, 129)])
def test_tuned_plan_keeps_the_answer_and_the_cycle(w, h):
    print(f"tune {w}x{h}", flush=True)
    d, h0, kd, dt = F.case_lognormal(w=w, h=h, seed=8)
    ws = dtdiff.Workspace()
    try:
        hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
        ms_rule, ms_tuned, plan = dtdiff.tune(ws, d, hd, kdd, dt, reps=3)
        print(f"rule {ms_rule:.3f} ms, tuned {ms_tuned:.3f} ms, plan {plan}", flush=True)
        assert same_bits(hd.numpy(), h0)  # only read
        assert 0 < ms_tuned <= ms_rule
        assert set(plan) >= {"levels", "direction", "residual", "finish1", "finish2", "bottom"}
        for pre, post in plan["levels"]:
            assert pre % 4 == 0 and post % 4 == 0 and 0 <= pre <= 64 and 0 <= post <= 64
        assert plan["bottom"] in (-1, 256, 512, 1024, 2048)
        st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, 1e-10)
        x = hd.numpy()
        assert O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x) <= 1e-10
        tol = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(h0.size)
        xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
        assert st.iterations <= its_o + 2
        # The cycle under the plan (dtd_cycle_new takes the shape's plan): the oracle's bits.
        e, s, g = O.operator(w, h, d.dx, d.wrap_x, kd, dt)
        cy = dtdiff.Cycle(ws, d, e, s, g)
        hy = O.Hierarchy(w, h, d.wrap_x, e, s, g)
        r = F.splitmix_f64(3 + w, w * h) - 0.5
        z, rz = cy.precondition(r.astype(np.float32), r)
        zo, rzo = hy.precondition(r.astype(np.float32), r)
        assert same_bits(z, zo) and abs(rz - rzo) <= 1e-12 * abs(rzo)
        # The no-allocation solve follows the plan too: the same bits as dtd_solve_device.
        if d.cells() <= 600_000:
            ha, hb = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(h0)
            dtdiff.solve_device(ws, d, ha, kdd, dt, tol)
            dtdiff.solve_device_noalloc(ws, d, hb, kdd, dt, tol)
            assert same_bits(ha.numpy(), hb.numpy())
    finally:
        ws.close()


def test_plans_are_kept_per_shape_and_can_be_forgotten():
    ws = dtdiff.Workspace()
    try:
        d1, a0, ka, dta = F.case_lognormal(w=256, h=128, seed=9)
        d2, b0, kb, dtb = F.case_constant(w=128, h=64, seed=9)
        ha, kad = DeviceArray.from_numpy(a0), DeviceArray.from_numpy(ka)
        hb, kbd = DeviceArray.from_numpy(b0), DeviceArray.from_numpy(kb)
        _, _, plan = dtdiff.tune(ws, d1, ha, kad, dta, reps=2)
        assert dtdiff.tune_plan(ws, d2) == {}
        dtdiff.solve_rel_device(ws, d2, hb, kbd, dtb, 1e-10)  # another shape in between
        assert dtdiff.tune_plan(ws, d1) == plan
        st = dtdiff.solve_rel_device(ws, d1, ha, kad, dta, 1e-10)
        assert st.rel_residual <= 1e-10
        dtdiff.forget_plan(ws, d1)
        assert dtdiff.tune_plan(ws, d1) == {}
        ha = DeviceArray.from_numpy(a0)
        assert dtdiff.solve_rel_device(ws, d1, ha, kad, dta, 1e-10).rel_residual <= 1e-10
    finally:
        ws.close()


def test_batch_policy_is_tuned_kept_and_keeps_the_answers():
    """A batch of mixed shapes: the tuner leaves the grids' h0 in place, keeps a policy for the
    shape list, and the tuned batch still meets the target and matches the oracle's solves."""
    shapes = [(512, 256), (96, 64), (1024, 96), (64, 8), (300, 129)]
    grids = [F.case_lognormal(w=a, h=b, seed=20 + k, ratio=10.0 + 40 * k)
             for k, (a, b) in enumerate(shapes)]
    ws = dtdiff.Workspace()
    try:
        dev = [(DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)) for (_, h0, kd, _) in grids]
        items = [(d, hd, kdd, dt, 0.0) for (d, _, _, dt), (hd, kdd) in zip(grids, dev)]
        ms_rule, ms_tuned = dtdiff.tune_batch(ws, items, 1e-10, reps=2)
        print(f"batch rule {ms_rule:.3f} ms tuned {ms_tuned:.3f} ms", flush=True)
        assert 0 < ms_tuned <= ms_rule
        for (_, h0, _, _), (hd, _) in zip(grids, dev):
            assert same_bits(hd.numpy(), h0)
        its, rels, st = dtdiff.solve_batch(ws, items, 1e-10)
        for (d, h0, kd, dt), (hd, _) in zip(grids, dev):
            x = hd.numpy()
            assert O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x) <= 1e-10
            tol = 1e-10 * float(np.linalg.norm(h0)) / np.sqrt(h0.size)
            xo, _ = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
            assert np.linalg.norm(x - xo) <= 3e-10 * np.linalg.norm(h0)
    finally:
        ws.close()
