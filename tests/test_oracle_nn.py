"""The learned guess's network oracle (oracle/nn_oracle.py), pinned by mathematics on the CPU.

The network has no counterpart in the reference, so nothing of the reference pins it.  What does:
- the 3x3 convolution against its definition written as loops (periodic x, zero rows past the
  ends);
- each backward pass as the exact adjoint of its forward pass, <A x, y> = <x, A^T y>, for the
  convolution's input, pooling and upsampling, and the weights' gradient as the derivative of
  the (linear in the weights) convolution;
- the whole network's gradient against central finite differences in f64;
- bf16 rounding against hand-picked values (ties to even).
"""
import numpy as np
import pytest

import nn_oracle as N


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38],
                 dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, -2.5, 2.9980e38], dtype=np.float32)
    got = N.bf16(x)
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[5] / 3.0e38 - 1) < 2 ** -8


def test_conv_matches_its_definition():
    rng = np.random.default_rng(1)
    n, h, w, ci, co = 2, 3, 5, 2, 3
    x = rng.standard_normal((n, h, w, ci))
    wt = rng.standard_normal((co, 9, ci))
    b = rng.standard_normal(co)
    want = np.zeros((n, h, w, co))
    for i in range(n):
        for y in range(h):
            for xx in range(w):
                for o in range(co):
                    s = b[o]
                    for t in range(9):
                        dy, dx = t // 3 - 1, t % 3 - 1
                        yy = y + dy
                        if 0 <= yy < h:
                            s += wt[o, t] @ x[i, yy, (xx + dx) % w]
                    want[i, y, xx, o] = s
    assert np.allclose(N.conv3x3(x, wt, b), want, rtol=0, atol=1e-12)


def test_backward_passes_are_adjoints():
    rng = np.random.default_rng(2)
    n, h, w, ci, co = 2, 6, 8, 3, 4
    x = rng.standard_normal((n, h, w, ci))
    wt = rng.standard_normal((co, 9, ci))
    dz = rng.standard_normal((n, h, w, co))
    lhs = np.sum(N.conv3x3(x, wt) * dz)
    assert abs(lhs - np.sum(x * N.conv3x3_bwd_input(dz, wt))) < 1e-10 * abs(lhs)
    # The convolution is linear in its weights: <conv(x; e), dz> = <dw, e> for any e.
    e = rng.standard_normal(wt.shape)
    dw, db = N.conv3x3_bwd_weight(x, dz)
    assert abs(np.sum(N.conv3x3(x, e) * dz) - np.sum(dw * e)) < 1e-10 * np.abs(dw * e).sum()
    assert np.allclose(db, dz.reshape(-1, co).sum(0))
    y = rng.standard_normal((n, h // 2, w // 2, ci))
    assert abs(np.sum(N.pool(x) * y) - np.sum(x * N.pool_bwd(y))) < 1e-12 * np.abs(x).sum()
    d = rng.standard_normal((n, h, w, ci))
    assert abs(np.sum(N.up(y) * d) - np.sum(y * N.up_bwd(d))) < 1e-12 * np.abs(d).sum()


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    _, npar = N.offsets()
    flat = rng.standard_normal(npar) * 0.05
    x0 = rng.standard_normal((1, 8, 16, 16))
    x0[..., 3:] = 0.0
    t = rng.standard_normal((1, 8, 16))
    out, A = N.forward(N.unpack(flat), x0, bf=False)
    L0, dy = N.loss(out, t, bf=False)
    g = N.flat_grad(N.backward(N.unpack(flat), A, dy, bf=False))
    assert np.abs(g).max() > 0
    # Every layer's weights and bias along a random direction; steps of 1e-7 stay clear of the
    # ReLUs' kinks (a step of 1e-6 already crosses some at these sizes).
    offs, _ = N.offsets()
    eps = 1e-7
    for name, ci, co, _ in N.LAYERS:
        a, b = offs[name]
        v = np.zeros(npar)
        v[a:b + co] = rng.standard_normal(b + co - a)
        lp, _ = N.loss(N.forward(N.unpack(flat + eps * v), x0, bf=False)[0], t, bf=False)
        lm, _ = N.loss(N.forward(N.unpack(flat - eps * v), x0, bf=False)[0], t, bf=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g @ v) <= 1e-6 * abs(fd), (name, fd, g @ v)


def test_input_channels():
    h = np.array([[1000.0, -500.0]])
    kd = np.array([[1.0, np.e]])
    x = N.input_channels(h, kd, 10.0, 1.0, 1000.0)
    assert np.allclose(x[..., 0], [[1.0, -0.5]])
    assert np.allclose(x[..., 1], [[0.0, 1.0]])
    assert np.allclose(x[..., 2], 0.25 * np.log(10.0 * kd))
    assert not x[..., 3:].any()
    x = N.input_channels(h, kd, 10.0, 1.0, 1000.0, base=np.array([[250.0, 0.0]]))
    assert np.allclose(x[..., 3], [[0.25, 0.0]]) and not x[..., 4:].any()


def test_coarse_transfers_reproduce_constants_and_linear_fields():
    """The learned coarse correction's block means and its interpolation back: a linear field's
    block mean is its value at the block's centre; interpolating the centres' values of a linear
    field gives the field itself between the first and last centres (in x away from the wrap, in
    y inside the outer centres), a constant everywhere, and f = 1 is the identity."""
    w, h, f = 64, 32, 8
    j, i = np.meshgrid(np.arange(h) + 0.5, np.arange(w) + 0.5, indexing="ij")
    lin = (0.3 * i - 0.7 * j + 2.0).reshape(-1)
    bm = N.block_mean(lin, w, h, f)
    J, I = np.meshgrid(np.arange(h // f) * f + f / 2, np.arange(w // f) * f + f / 2, indexing="ij")
    assert np.allclose(bm, (0.3 * I - 0.7 * J + 2.0).reshape(-1), rtol=0, atol=1e-12)
    assert np.allclose(N.block_mean(np.exp(lin / 10), w, h, f, True),
                       np.exp((0.3 * I - 0.7 * J + 2.0).reshape(-1) / 10), rtol=1e-12)
    back = N.interp_centres(bm, w, h, f).reshape(h, w)
    inner = (slice(f // 2, h - f // 2), slice(f // 2, w - f // 2))
    assert np.allclose(back[inner], lin.reshape(h, w)[inner], rtol=0, atol=1e-12)
    assert np.allclose(N.interp_centres(np.full((h // f) * (w // f), 4.0), w, h, f), 4.0)
    x = np.random.default_rng(3).standard_normal(w * h)
    assert np.array_equal(N.interp_centres(x, w, h, 1), x)
    # Periodic in x: the columns past the last centre blend the last and the first centres.
    c = np.zeros((h // f, w // f))
    c[:, 0] = 1.0
    row = N.interp_centres(c.reshape(-1), w, h, f).reshape(h, w)[h // 2]
    # Column w-1 lies 1/2 - 1/(2f) of the way from the last centre to the first (wrapped), column
    # 0 as far short of the first centre.
    assert abs(row[w - 1] - (0.5 - 0.5 / f)) < 1e-12 and abs(row[0] - (0.5 + 0.5 / f)) < 1e-12
