"""Pins of the oracle's cycle options (no reference counterpart; CPU only).

- The bilinear prolongation against P built from the cells' centres: a fine cell takes the two
  nearest coarse centres along each axis with the linear-interpolation weights 3/4 and 1/4; past
  a closed edge the missing centre is the cell's own (zero slope), on a periodic axis it wraps.
  The restriction is checked against the exact transpose of that matrix.
- The exact coarsest solve against a dense f64 solve of the level's matrix.
- The cycle with either option, in dense matrix form, and its symmetry and positivity (the
  conjugate gradients need a symmetric positive definite preconditioner).
"""
import numpy as np
import pytest

import oracle as O
from test_oracle import _jacobi, _level_matrix


def axis_weights(nf, nc, wrap):
    """(nf x nc) linear-interpolation weights from the cell centres: fine cell i at i + 1/2, coarse
    cell I at 2 I + 1 (fine units)."""
    W = np.zeros((nf, nc))
    for i in range(nf):
        x = i + 0.5
        near = i // 2
        side = near + 1 if x > 2 * near + 1 else near - 1
        if wrap:
            side %= nc
        elif side < 0 or side >= nc:
            side = near
        dist = abs(x - (2 * near + 1))  # 1/2: the weights are 1 - 1/4 and 1/4
        W[i, near] += 1.0 - dist / 2.0
        W[i, side] += dist / 2.0
    return W


def dense_P(wf, hf, wc, hc, cwrap):
    return np.kron(axis_weights(hf, hc, False), axis_weights(wf, wc, cwrap))


SHAPES = [(w, h, wrap) for w in (1, 2, 3, 4, 5, 8, 9, 16) for h in (1, 2, 3, 6, 7)
          for wrap in (True, False)]


@pytest.mark.parametrize("wf,hf,wrap", SHAPES)
def test_bilinear_transfers_are_the_centre_interpolation_and_its_transpose(wf, hf, wrap):
    wrap = wrap and wf > 1
    wc, hc = (wf + 1) // 2, (hf + 1) // 2
    cwrap = wrap and wc > 1
    P = dense_P(wf, hf, wc, hc, cwrap)
    assert np.allclose(P.sum(axis=1), 1.0)  # constants kept
    rng = np.random.default_rng(wf * 100 + hf)
    zc = rng.standard_normal(wc * hc).astype(np.float32)
    got = O.bil_prolong(wf, hf, wc, hc, cwrap, zc)
    assert np.allclose(got, P @ zc.astype(np.float64), rtol=0, atol=1e-6 * np.abs(zc).max())
    res = rng.standard_normal(wf * hf).astype(np.float32)
    rc = O.bil_restrict(wf, hf, wrap, wc, hc, cwrap, res)
    assert np.allclose(rc, P.T @ res.astype(np.float64), rtol=0, atol=4e-6 * np.abs(res).sum())
    # A constant is prolonged exactly.
    assert (O.bil_prolong(wf, hf, wc, hc, cwrap, np.full(wc * hc, 3.0, np.float32)) == 3.0).all()


def test_bilinear_prolongation_is_linear_inside_a_closed_level():
    """Away from the edges a field linear in x and y is reproduced (the interpolation's order)."""
    wf, hf, wc, hc = 16, 12, 8, 6
    I, J = np.meshgrid(np.arange(wc), np.arange(hc))
    zc = (0.5 * (2 * I + 1) - 0.25 * (2 * J + 1)).astype(np.float32).ravel()
    got = O.bil_prolong(wf, hf, wc, hc, False, zc).reshape(hf, wf)
    i, j = np.meshgrid(np.arange(wf) + 0.5, np.arange(hf) + 0.5)
    want = 0.5 * i - 0.25 * j
    assert np.allclose(got[1:-1, 1:-1], want[1:-1, 1:-1], rtol=0, atol=1e-5)


def level_of(w, h, wrap, seed):
    b = O.terrain(w, h, seed)
    kd = O.shelf_kd(b, seed)
    e, s, g = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    return e, s, g


@pytest.mark.parametrize("w,h,wrap", [(8, 8, True), (7, 5, False), (2, 2, True), (1, 9, False),
                                      (4, 16, True)])
def test_exact_coarsest_solve_is_a_dense_solve(w, h, wrap):
    e, s, g = level_of(w, h, wrap, 5)
    hy = O.Hierarchy(w, h, wrap, e, s, g)
    L = hy.level(0)
    A = _level_matrix(L)
    r = np.random.default_rng(w + h).standard_normal(w * h).astype(np.float32)
    z = O.coarse_exact(w, h, L["wrap"], L["east"], L["south"], L["diag"], r)
    want = np.linalg.solve(A, r.astype(np.float64))
    assert np.allclose(z, want, rtol=2e-7, atol=1e-7 * np.abs(want).max())


@pytest.mark.parametrize("transfer,coarsest", [(1, 0), (0, 1), (1, 1)])
def test_two_level_cycle_with_the_options_in_matrix_form(transfer, coarsest):
    w, h = 10, 12  # one coarse level of 5 x 6 = 30 cells
    e, s, g = level_of(w, h, True, 6)
    hy = O.Hierarchy(w, h, True, e, s, g, transfer=transfer, coarsest=coarsest)
    assert hy.levels() == 2
    A, Ac = _level_matrix(hy.level(0)), _level_matrix(hy.level(1))
    P = dense_P(w, h, 5, 6, True) if transfer else np.kron(axis_weights(h, 6, False) > 0.5,
                                                           axis_weights(w, 5, True) > 0.5) * 1.0
    r = np.random.default_rng(2).standard_normal(w * h)
    r32 = r.astype(np.float32).astype(np.float64)
    om = 0.8
    z = _jacobi(A, r32, om, om * r32 / np.diag(A), 1)
    rc = P.T @ (r32 - A @ z)
    zc = np.linalg.solve(Ac, rc) if coarsest else _jacobi(Ac, rc, om, om * rc / np.diag(Ac), 15)
    z = _jacobi(A, r32, om, z + P @ zc, 2)
    got, rz = hy.precondition(r.astype(np.float32), r)
    assert np.allclose(got, z, rtol=0, atol=2e-5 * np.abs(z).max())
    assert abs(rz - float(r @ got.astype(np.float64))) <= 1e-12 * abs(rz)


@pytest.mark.parametrize("transfer,coarsest", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_the_option_cycles_are_symmetric_positive_definite(transfer, coarsest):
    """M^-1 applied to the unit vectors of a three-level grid: symmetric to f32 rounding, with
    positive eigenvalues (the reference's SPD test of its cycle, multigrid/tests.rs:242-289)."""
    w, h = 12, 10
    e, s, g = level_of(w, h, True, 7)
    hy = O.Hierarchy(w, h, True, e, s, g, transfer=transfer, coarsest=coarsest)
    assert hy.levels() >= 2
    n = w * h
    M = np.zeros((n, n))
    for k in range(n):
        u = np.zeros(n)
        u[k] = 1.0
        M[:, k] = hy.precondition(u.astype(np.float32), u)[0]
    assert np.abs(M - M.T).max() <= 1e-5 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0.0
