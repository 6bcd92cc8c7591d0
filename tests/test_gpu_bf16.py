"""The bf16-storage variant of the cycle (DTD_CYCLE_BF16) on the GPU.

The two finest levels' smoothers read their operator records (east, south, diag, omega / diag)
rounded to bf16; everything else is the f32 cycle, and the conjugate gradients stay f64.  The
variant is a different preconditioner, not a different answer: the solve still reaches a true f64
relative residual of 1e-10 and lies within the tolerance bound of the oracle's solve.  What the
rounding costs is stated: the cycle's correction z against the f32 cycle's, and the iterations.
"""
import numpy as np
import pytest

import dtdiff
import oracle as O
from dtdiff import fields as F
from dtdiff.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def solve_rel(cycle, d, h0, kd, dt, rel):
    ws = dtdiff.Workspace(cycle_precision=cycle)
    try:
        hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
        st = dtdiff.solve_rel_device(ws, d, hd, kdd, dt, rel)
        synchronize()
        return hd.numpy(), st
    finally:
        ws.close()


@pytest.mark.parametrize("w,h,gen", [(1024, 512, "lognormal"), (512, 256, "constant"),
                                     (300, 129, "lognormal"), (2048, 1024, "lognormal")])
def test_bf16_variant_reaches_the_target_and_the_oracle(w, h, gen):
    print(f"bf16 {gen} {w}x{h}", flush=True)
    d, h0, kd, dt = (F.case_constant if gen == "constant" else F.case_lognormal)(w=w, h=h)
    x16, st16 = solve_rel(dtdiff.CYCLE_BF16, d, h0, kd, dt, 1e-10)
    x32, st32 = solve_rel(dtdiff.CYCLE_F32, d, h0, kd, dt, 1e-10)
    assert O.rel_residual(d.w, d.h, d.dx, d.wrap_x, kd, dt, h0, x16) <= 1e-10
    tol = 1e-10 * np.linalg.norm(h0) / np.sqrt(h0.size)
    xo, its_o = O.solve(d.w, d.h, d.dx, d.wrap_x, h0, kd, dt, tol)
    # Both within ||r|| <= 1e-10 ||b|| of the exact solution (eigenvalues of A are >= 1).
    assert np.linalg.norm(x16 - xo) <= 3e-10 * np.linalg.norm(h0)
    assert st16.iterations <= st32.iterations + 3
    print(f"  iterations bf16 {st16.iterations} f32 {st32.iterations} oracle {its_o}; "
          f"rel L2 bf16 vs oracle {np.linalg.norm(x16 - xo) / np.linalg.norm(xo):.2e}", flush=True)


def test_bf16_cycle_correction_against_the_f32_cycle():
    """One V-cycle (dtd_cycle_precondition) on the same residual: the bf16 records move z by
    about bf16's relative precision (2^-8), no more."""
    w, h = 256, 128
    d, h0, kd, dt = F.case_lognormal(w=w, h=h)
    e, s, g = O.operator(w, h, d.dx, True, kd, dt)
    rng = np.random.default_rng(3)
    r64 = rng.standard_normal(w * h)
    r32 = r64.astype(np.float32)
    out = {}
    for cyc in (dtdiff.CYCLE_F32, dtdiff.CYCLE_BF16):
        ws = dtdiff.Workspace(cycle_precision=cyc)
        try:
            c = dtdiff.Cycle(ws, dtdiff.Dims(w, h, d.dx, True), e, s, g)
            out[cyc] = c.precondition(r32, r64)
        finally:
            ws.close()
    z32, rz32 = out[dtdiff.CYCLE_F32]
    z16, rz16 = out[dtdiff.CYCLE_BF16]
    rel = np.linalg.norm(z16.astype(np.float64) - z32) / np.linalg.norm(z32)
    print(f"  ||z_bf16 - z_f32|| / ||z_f32|| = {rel:.2e}; r.z {rz16:.6e} vs {rz32:.6e}", flush=True)
    assert 0.0 < rel <= 2e-2
    assert rz16 > 0.0 and abs(rz16 - rz32) <= 2e-2 * rz32
