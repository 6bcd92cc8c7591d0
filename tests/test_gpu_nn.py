"""The learned guess's network kernels (include/dtdiff.h `dtd_nn_*`, deep-time_amd/csrc/nn.hpp) on
the GPU against the numpy oracle (oracle/nn_oracle.py) on the same bf16 inputs.

Tolerances: the device accumulates in f32 where the oracle sums in f64, so an f32 output agrees
to ~1e-6 relative; a bf16 output is the oracle's value rounded to bf16 except where the two sums
straddle a rounding boundary, which moves it by one bf16 step (2^-8 relative).  Through the whole
network such steps compound, so the network's output and gradient are compared in relative L2
(a wrong index, tap, sign or skip term gives O(1) errors there).
"""
import ctypes

import numpy as np
import pytest

import dtdiff
import nn_oracle as N
from dtdiff import learned as LN
from dtdiff.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    w = dtdiff.Workspace()
    yield w
    w.close()


def to_bits(a):
    return (N.bf16(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


def dev_bf16(a):
    return DeviceArray.from_numpy(to_bits(a).reshape(-1))


def get_bf16(d, shape):
    synchronize()
    return from_bits(d.numpy()).reshape(shape)


def get_f32(d, shape):
    synchronize()
    return d.numpy().reshape(shape)


def p(a):
    return ctypes.c_void_p(a.data_ptr())


def close_bf16(got, want, frac=0.98):
    """got equals bf16(want) but for one-step rounding flips on a small share of elements."""
    want_b = N.bf16(want)
    tol = 2.0 ** -7 * np.abs(want_b) + 1e-30
    assert np.all(np.abs(got - want_b) <= tol + 1e-6 * np.abs(want).max()), \
        np.abs(got - want_b).max()
    assert np.mean(got == want_b) >= frac


@pytest.mark.parametrize("ci,co,relu,w", [(16, 16, 0, 32), (16, 32, 1, 32), (48, 64, 1, 32),
                                          (32, 80, 0, 32), (96, 32, 1, 32), (32, 64, 1, 128),
                                          (16, 16, 0, 64), (384, 128, 1, 192),
                                          (96, 32, 1, 256), (64, 64, 0, 128),
                                          (64, 128, 1, 256), (32, 16, 0, 256), (64, 48, 0, 128),
                                          (32, 32, 1, 512), (32, 16, 0, 512), (96, 32, 1, 512),
                                          (16, 32, 1, 128), (16, 32, 1, 256), (16, 16, 0, 128),
                                          (16, 64, 0, 256), (48, 32, 1, 256)])
def test_conv_forward(ws, ci, co, relu, w):
    """w a multiple of 128 takes the LDS-staged kernel (32-channel stages for cin a multiple of 32,
    else 16-channel stages), other multiples of 64 the four-tile kernel, the rest the one-tile
    kernel."""
    rng = np.random.default_rng(ci * 7 + co + w)
    n, h = 2, 6
    x = N.bf16(rng.standard_normal((n, h, w, ci)))
    wt = N.bf16(rng.standard_normal((co, 9, ci)) / np.sqrt(9 * ci))
    b = rng.standard_normal(co).astype(np.float32)
    dx, dw, db = dev_bf16(x), dev_bf16(wt), DeviceArray.from_numpy(b)
    y = DeviceArray(n * h * w * co, np.uint16)
    y32 = DeviceArray(n * h * w * 2, np.float32)
    L = dtdiff.load()
    ws.check(L.dtd_nn_conv3x3(ws._ctx, n, h, w, ci, co, p(dx), p(dw), p(db), relu, p(y), p(y32),
                              2))
    z = N.conv3x3(x, wt, b)
    want = np.maximum(z, 0) if relu else z
    close_bf16(get_bf16(y, (n, h, w, co)), want)
    got32 = get_f32(y32, (n, h, w, 2)).copy()
    assert np.allclose(got32, want[..., :2], rtol=1e-5, atol=1e-5)
    # Only the f32 channels: the same values, no bf16 output.
    y32b = DeviceArray(n * h * w * 2, np.float32)
    ws.check(L.dtd_nn_conv3x3(ws._ctx, n, h, w, ci, co, p(dx), p(dw), p(db), relu, None,
                              p(y32b), 2))
    assert np.array_equal(get_f32(y32b, (n, h, w, 2)), got32)


@pytest.mark.parametrize("ci,co,w,h", [(16, 32, 48, 8), (64, 48, 48, 8), (32, 16, 48, 8),
                                        (64, 128, 128, 8), (32, 32, 48, 3)])
def test_conv_backward(ws, ci, co, w, h):
    """Pixel counts that are multiples of 32 take the 32-pixel weight-gradient steps, others
    (1 x 3 x 48) the 16-pixel steps."""
    rng = np.random.default_rng(ci + 3 * co + w)
    n = 2 if h == 8 else 1
    x = N.bf16(rng.standard_normal((n, h, w, ci)))
    wt = N.bf16(rng.standard_normal((co, 9, ci)) / np.sqrt(9 * ci))
    dz = N.bf16(rng.standard_normal((n, h, w, co)))
    L = dtdiff.load()
    d_x, d_w, d_dz = dev_bf16(x), dev_bf16(wt), dev_bf16(dz)
    flip = DeviceArray(ci * 9 * co, np.uint16)
    d_in = DeviceArray(n * h * w * ci, np.uint16)
    ws.check(L.dtd_nn_conv3x3_bwd_input(ws._ctx, n, h, w, ci, co, p(d_dz), p(d_w), p(flip),
                                        p(d_in)))
    close_bf16(get_bf16(d_in, (n, h, w, ci)), N.conv3x3_bwd_input(dz, wt))
    f = ctypes.c_size_t()
    npl = L.dtd_nn_bwd_weight_scratch(n, h, w, ci, co, ctypes.byref(f))
    planes, part = DeviceArray(npl, np.uint16), DeviceArray(f.value, np.float32)
    gw, gb = DeviceArray(co * 9 * ci, np.float32), DeviceArray(co, np.float32)
    ws.check(L.dtd_nn_conv3x3_bwd_weight(ws._ctx, n, h, w, ci, co, p(d_x), p(d_dz), p(planes),
                                         npl, p(part), f.value, p(gw), p(gb)))
    want_w, want_b = N.conv3x3_bwd_weight(x, dz)
    got_w = get_f32(gw, (co, 9, ci))
    assert np.abs(got_w - want_w).max() <= 1e-5 * np.abs(want_w).max()
    assert np.allclose(get_f32(gb, (co,)), want_b, rtol=1e-5, atol=1e-4)
    # Scratch below the stated size is refused.
    rc = L.dtd_nn_conv3x3_bwd_weight(ws._ctx, n, h, w, ci, co, p(d_x), p(d_dz), p(planes),
                                     npl - 1, p(part), f.value, p(gw), p(gb))
    assert rc != 0


def test_pool_up_cat_and_relu(ws):
    rng = np.random.default_rng(5)
    n, h, w, c, cs = 2, 8, 64, 32, 16
    L = dtdiff.load()
    x = N.bf16(rng.standard_normal((n, h, w, c)))
    dx = dev_bf16(x)
    y = DeviceArray(n * h // 2 * w // 2 * c, np.uint16)
    ws.check(L.dtd_nn_pool(ws._ctx, n, h, w, c, p(dx), p(y)))
    close_bf16(get_bf16(y, (n, h // 2, w // 2, c)), N.pool(x), frac=0.999)
    u = N.bf16(rng.standard_normal((n, h // 2, w // 2, c)))
    skip = N.bf16(rng.standard_normal((n, h, w, cs)))
    du_, ds_ = dev_bf16(u), dev_bf16(skip)
    cat = DeviceArray(n * h * w * (c + cs), np.uint16)
    ws.check(L.dtd_nn_up_cat(ws._ctx, n, h, w, c, cs, p(du_), p(ds_), p(cat)))
    assert np.array_equal(get_bf16(cat, (n, h, w, c + cs)),
                          np.concatenate([N.up(u), skip], axis=3).astype(np.float32))
    dcat = N.bf16(rng.standard_normal((n, h, w, c + cs)))
    dpool = N.bf16(rng.standard_normal((n, h // 2, w // 2, cs)))
    d_dcat, d_dpool = dev_bf16(dcat), dev_bf16(dpool)
    g_u = DeviceArray(n * h // 2 * w // 2 * c, np.uint16)
    g_s = DeviceArray(n * h * w * cs, np.uint16)
    ws.check(L.dtd_nn_up_cat_bwd(ws._ctx, n, h, w, c, cs, p(d_dcat), p(d_dpool), p(g_u),
                                 p(g_s)))
    close_bf16(get_bf16(g_u, (n, h // 2, w // 2, c)), N.up_bwd(dcat[..., :c]), frac=0.999)
    close_bf16(get_bf16(g_s, (n, h, w, cs)), dcat[..., c:] + N.pool_bwd(dpool), frac=0.999)
    yv = N.bf16(rng.standard_normal(4096))
    dyv = N.bf16(rng.standard_normal(4096))
    d_y, d_dy, d_dz = dev_bf16(yv), dev_bf16(dyv), DeviceArray(4096, np.uint16)
    ws.check(L.dtd_nn_relu_bwd(ws._ctx, 4096, p(d_y), p(d_dy), p(d_dz)))
    assert np.array_equal(get_bf16(d_dz, (4096,)), np.where(yv > 0, dyv, 0).astype(np.float32))


@pytest.mark.parametrize("cu,cs,co,w,h", [(256, 128, 128, 128, 4), (128, 64, 64, 256, 6),
                                           (64, 32, 32, 256, 4), (64, 32, 32, 128, 2),
                                           (64, 32, 32, 512, 2)])
def test_conv_on_the_concatenation_in_place(ws, cu, cs, co, w, h):
    """dtd_nn_conv3x3_cat equals dtd_nn_up_cat then dtd_nn_conv3x3 bit for bit (the same sums in
    the same order), and both match the oracle."""
    rng = np.random.default_rng(cu + cs + co + w + h)
    n, ci = 2, cu + cs
    u = N.bf16(rng.standard_normal((n, h // 2, w // 2, cu)))
    skip = N.bf16(rng.standard_normal((n, h, w, cs)))
    wt = N.bf16(rng.standard_normal((co, 9, ci)) / np.sqrt(9 * ci))
    b = rng.standard_normal(co).astype(np.float32)
    du_, ds_, dw, db = dev_bf16(u), dev_bf16(skip), dev_bf16(wt), DeviceArray.from_numpy(b)
    L = dtdiff.load()
    cat = DeviceArray(n * h * w * ci, np.uint16)
    y0, y1 = (DeviceArray(n * h * w * co, np.uint16) for _ in range(2))
    ws.check(L.dtd_nn_up_cat(ws._ctx, n, h, w, cu, cs, p(du_), p(ds_), p(cat)))
    ws.check(L.dtd_nn_conv3x3(ws._ctx, n, h, w, ci, co, p(cat), p(dw), p(db), 1, p(y0), None, 0))
    ws.check(L.dtd_nn_conv3x3_cat(ws._ctx, n, h, w, cu, cs, co, p(du_), p(ds_), p(dw), p(db), 1,
                                  p(y1), None, 0))
    synchronize()
    assert np.array_equal(y0.numpy(), y1.numpy())
    want = np.maximum(N.conv3x3(np.concatenate([N.up(u), skip], axis=3), wt, b), 0)
    close_bf16(get_bf16(y1, (n, h, w, co)), want)
    # Rows that are not a multiple of 128 are refused.
    assert L.dtd_nn_conv3x3_cat(ws._ctx, n, h, 64, cu, cs, co, p(du_), p(ds_), p(dw), p(db), 1,
                                p(y1), None, 0) != 0


def test_inference_forward_equals_the_training_forward(ws):
    """The guess's forward (concatenations read in place) gives the training forward's output."""
    rng = np.random.default_rng(8)
    params = LN.init_params(3)
    offs, _ = LN.offsets()
    a, b = offs["out"]
    params[a:b] = (rng.standard_normal(b - a) * 0.05).astype(np.float32)  # a live output layer
    d = dtdiff.Dims(512, 16, 1.0e4, True)
    hh, kd, base = (DeviceArray.from_numpy(v) for v in (
        rng.standard_normal(d.w * d.h) * 100.0, np.exp(rng.standard_normal(d.w * d.h)),
        rng.standard_normal(d.w * d.h) * 100.0))
    net = LN.UNet(ws, params)
    net.set_input(d, hh, kd, base, 1.0e3)
    net.forward(train=True)
    want = get_f32(net.B["out32"], (-1,)).copy()
    assert np.abs(want).max() > 0
    net.forward(train=False)
    assert np.array_equal(get_f32(net.B["out32"], (-1,)), want)


def test_input_channels_match_the_oracle(ws):
    """dtd_nn_input's 16 channels (two 16-byte stores a pixel) against the oracle's, rounded to
    bf16; channels 4-15 are zero."""
    rng = np.random.default_rng(9)
    n = 3000
    h = rng.standard_normal(n) * 800.0
    kd = np.exp(rng.standard_normal(n))
    kd[:5] = 0.0  # clamped to 1e-30 before the logarithms
    base = h + rng.standard_normal(n) * 30.0
    L = dtdiff.load()
    x = DeviceArray(n * 16, np.uint16)
    hd, kdd, bd = (DeviceArray.from_numpy(v) for v in (h, kd, base))
    ws.check(L.dtd_nn_input(ws._ctx, n, p(hd), p(kdd), p(bd), 20.0, 3.0, LN.SCALE, p(x)))
    got = get_bf16(x, (n, 16))
    want = N.input_channels(h.reshape(1, 1, n), kd.reshape(1, 1, n), 20.0, 3.0, LN.SCALE,
                            base=base.reshape(1, 1, n)).reshape(n, 16)
    close_bf16(got, want, frac=0.999)
    assert not got[:, 4:].any()


def test_adam_step(ws):
    rng = np.random.default_rng(6)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    dw, dm, dv, dg = (DeviceArray.from_numpy(a) for a in
                      (w, np.zeros(n, np.float32), np.zeros(n, np.float32), g))
    wb = DeviceArray(n, np.uint16)
    L = dtdiff.load()
    for step in (1, 2):
        ws.check(L.dtd_nn_adam(ws._ctx, n, p(dw), p(dm), p(dv), p(dg), 1e-2, 0.9, 0.999, 1e-8,
                               step, p(wb)))
    m1, v1 = 0.1 * g, 0.001 * g * g
    w1 = w - 1e-2 * (m1 / 0.1) / (np.sqrt(v1 / 0.001) + 1e-8)
    m2, v2 = 0.9 * m1 + 0.1 * g, 0.999 * v1 + 0.001 * g * g
    w2 = w1 - 1e-2 * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + 1e-8)
    got = get_f32(dw, (n,))
    assert np.allclose(got, w2, rtol=1e-5, atol=1e-6)
    assert np.array_equal(get_bf16(wb, (n,)), N.bf16(got))


def test_network_output_and_gradient_match_the_oracle(ws):
    """The whole U-Net forward, loss and backward on the device against the oracle with bf16 at
    the same places, at 16 x 128 (four levels down to 2 x 16)."""
    rng = np.random.default_rng(8)
    params = LN.init_params(3)
    offs, _ = LN.offsets()
    a, b = offs["out"]
    params[a:b] = (rng.standard_normal(b - a) * 0.05).astype(np.float32)  # a live output layer
    d = dtdiff.Dims(128, 16, 1.0, True)
    h0 = 1000.0 * rng.standard_normal(d.h * d.w)
    kd = np.exp(rng.standard_normal(d.h * d.w))
    base = h0 + 30.0 * rng.standard_normal(d.h * d.w)
    t = h0 + 100.0 * rng.standard_normal(d.h * d.w)
    net = LN.UNet(ws, params)
    hd, kdd, bd, td = (DeviceArray.from_numpy(v) for v in (h0, kd, base, t))
    net.set_input(d, hd, kdd, bd, 20.0)
    net.forward()
    loss, dy = net.loss(td, bd)
    net.backward(dy)
    g_dev = net.flat_grad()
    out_dev = get_f32(net.B["out32"], (1, d.h, d.w, LN.OUTC))

    shp = (1, d.h, d.w)
    x0 = N.input_channels(h0.reshape(shp), kd.reshape(shp), 20.0, 1.0, LN.SCALE,
                          base=base.reshape(shp))
    P = N.unpack(params.astype(np.float64))
    out, A = N.forward(P, x0, bf=True)
    L0, dy0 = N.loss(out, ((t - base) / LN.SCALE).reshape(shp), bf=True)
    g0 = N.flat_grad(N.backward(P, A, dy0, bf=True))
    assert abs(loss - L0) <= 1e-2 * L0
    assert np.linalg.norm(out_dev - out) <= 2e-2 * np.linalg.norm(out)
    for name, ci, co, _ in LN.LAYERS:
        lo, hi = offs[name][0], offs[name][1] + co
        rel = np.linalg.norm(g_dev[lo:hi] - g0[lo:hi]) / np.linalg.norm(g0[lo:hi])
        assert rel <= 5e-2, (name, rel)
    # The guess adds the predicted correction to the base in f64.
    gd = DeviceArray(d.w * d.h)
    ws.check(dtdiff.load().dtd_nn_guess(ws._ctx, d.w * d.h, p(bd), p(net.B["out32"]), LN.OUTC,
                                        LN.SCALE, p(gd)))
    synchronize()
    want = base + LN.SCALE * out_dev[0, ..., 0].reshape(-1).astype(np.float64)
    assert np.allclose(gd.numpy(), want, rtol=0, atol=1e-9 * LN.SCALE)


def test_training_lowers_the_loss(ws):
    """Sixty Adam steps on one small solve pair cut the loss by at least 30 %."""
    rng = np.random.default_rng(9)
    d = dtdiff.Dims(128, 32, 1.0, True)
    h0 = 1000.0 * np.sin(np.linspace(0, 12 * np.pi, d.w))[None, :] * \
        np.cos(np.linspace(0, 3 * np.pi, d.h))[:, None]
    h0 = (h0 + 200.0 * rng.standard_normal(h0.shape)).reshape(-1)
    kd = np.ones(d.h * d.w)
    hd, kdd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(kd)
    td = DeviceArray.from_numpy(h0)
    dtdiff.solve_rel_device(ws, d, td, kdd, 10.0, 1e-10)
    bd = DeviceArray(d.w * d.h)
    dtdiff.lod_device(ws, d, hd, kdd, 10.0, bd)
    net = LN.UNet(ws)
    losses = []
    for k in range(60):
        net.set_input(d, hd, kdd, bd, 10.0)
        net.forward()
        loss, dy = net.loss(td, bd)
        losses.append(loss)
        net.backward(dy)
        net.adam(1e-3)
    assert losses[-1] < 0.7 * losses[0], losses[::10]


def test_solve_from_a_guess_meets_the_same_test(ws):
    """Any guess ends at the same true residual: here h0 itself and a poor one (0.5 h0)."""
    rng = np.random.default_rng(10)
    d = dtdiff.Dims(256, 128, 1.0, True)
    h0 = 1000.0 * rng.standard_normal(d.w * d.h)
    kd = np.exp(0.5 * rng.standard_normal(d.w * d.h))
    kdd = DeviceArray.from_numpy(kd)
    ref = DeviceArray.from_numpy(h0)
    dtdiff.solve_rel_device(ws, d, ref, kdd, 30.0, 1e-10)
    for g in (h0, 0.5 * h0):
        hd, gd = DeviceArray.from_numpy(h0), DeviceArray.from_numpy(g)
        st = dtdiff.solve_rel_guess_device(ws, d, hd, kdd, 30.0, 1e-10, gd)
        assert st.rel_residual <= 1e-10 and st.iterations > 0
        synchronize()
        x, r = hd.numpy(), ref.numpy()
        assert np.linalg.norm(x - r) <= 1e-8 * np.linalg.norm(r)


def test_trained_guess_beats_the_splitting(ws):
    """What the committed network gains over the method without it, on held-out 1024 x 512
    samples (fields from seeds outside the training pools): its guess is several times closer to
    the converged answer than the splitting's, the solve from it takes no more iterations, and
    both solves meet the same true relative residual."""
    rows = LN.evaluate(LN.load_params(), count=2, seed=123)
    for r in rows:
        assert r["err_net_m"] < 0.5 * r["err_lod_m"], r
        assert r["iterations_net"] <= r["iterations_lod"], r
        assert r["rel_net"] <= 1e-10 and r["rel_lod"] <= 1e-10, r


def test_network_arguments_are_checked(ws):
    """Shapes, channel counts, alignment and nulls are refused with a message, before any
    launch."""
    L = dtdiff.load()
    a = DeviceArray(1 << 16, np.uint16)
    b = DeviceArray(1 << 16, np.uint16)

    def err(rc):
        assert rc != 0
        return L.dtd_last_error(ws._ctx).decode()

    # w not a multiple of 16, channels not multiples of 16.
    assert "multiple of 16" in err(L.dtd_nn_conv3x3(ws._ctx, 1, 4, 24, 16, 16, p(a), p(b), None, 0,
                                                    p(a), None, 0))
    assert "multiples of 16" in err(L.dtd_nn_conv3x3(ws._ctx, 1, 4, 32, 24, 16, p(a), p(b), None,
                                                     0, p(a), None, 0))
    assert "null" in err(L.dtd_nn_conv3x3(ws._ctx, 1, 4, 32, 16, 16, p(a), p(b), None, 0, None,
                                          None, 0))
    # A tensor off its 16-byte alignment (2 bytes in).
    off = ctypes.c_void_p(a.data_ptr() + 2)
    assert "aligned" in err(L.dtd_nn_pool(ws._ctx, 1, 4, 32, 16, off, p(b)))
    assert "aligned" in err(L.dtd_nn_relu_bwd(ws._ctx, 64, off, p(b), p(b)))
    # Odd sizes for pooling, nulls, a zero Adam step.
    assert "even" in err(L.dtd_nn_pool(ws._ctx, 1, 5, 32, 16, p(a), p(b)))
    assert "null" in err(L.dtd_nn_up_cat(ws._ctx, 1, 4, 32, 16, 16, p(a), None, p(b)))
    f = DeviceArray(64, np.float32)
    assert "step" in err(L.dtd_nn_adam(ws._ctx, 64, p(f), p(f), p(f), p(f), 1e-3, 0.9, 0.999,
                                       1e-8, 0, None))
    # The guessed solve needs its guess.
    d = dtdiff.Dims(64, 32, 1.0, True)
    h = DeviceArray.from_numpy(np.zeros(d.w * d.h))
    k = DeviceArray.from_numpy(np.ones(d.w * d.h))
    with pytest.raises(dtdiff.DtdError):
        dtdiff.solve_rel_guess_device(ws, d, h, k, 10.0, 1e-10, 0)
    # The guess's output needs at least one channel.
    g = DeviceArray(64, np.float64)
    assert "outc" in err(L.dtd_nn_guess(ws._ctx, 64, p(g), p(f), 0, 1.0, p(g)))
    assert "outc" in err(L.dtd_nn_guess(ws._ctx, 64, p(g), p(f), -3, 1.0, p(g)))


@pytest.mark.parametrize("w,h,f", [(2048, 1024, 8), (256, 64, 8), (96, 48, 4), (32, 16, 1)])
def test_coarse_correction_transfers_match_the_oracle(ws, w, h, f):
    """The learned coarse correction's block means (plain and geometric) and its bilinear
    interpolation back onto the grid, against the oracle."""
    rng = np.random.default_rng(w + h + f)
    a = rng.standard_normal(w * h) * 100.0
    k = np.exp(rng.standard_normal(w * h))
    L = dtdiff.load()
    nc = (w // f) * (h // f)
    ad, kd_, out = DeviceArray.from_numpy(a), DeviceArray.from_numpy(k), DeviceArray(nc)
    # The calls are asynchronous on the context's stream: wait before reading.
    ws.check(L.dtd_nn_block_mean(ws._ctx, w, h, f, 0, p(ad), p(out)))
    assert np.allclose(get_f32(out, (nc,)), N.block_mean(a, w, h, f), rtol=0, atol=1e-12 * 100)
    ws.check(L.dtd_nn_block_mean(ws._ctx, w, h, f, 1, p(kd_), p(out)))
    assert np.allclose(get_f32(out, (nc,)), N.block_mean(k, w, h, f, True), rtol=1e-13, atol=0)
    c = rng.standard_normal(nc).astype(np.float32)
    c2 = np.stack([c, np.zeros_like(c)], axis=1).reshape(-1)  # two channels, the first read
    base = rng.standard_normal(w * h)
    cd, bd, g = DeviceArray.from_numpy(c2), DeviceArray.from_numpy(base), DeviceArray(w * h)
    ws.check(L.dtd_nn_guess_coarse(ws._ctx, w, h, f, p(bd), p(cd), 2, 1000.0, p(g)))
    want = base + 1000.0 * N.interp_centres(c, w, h, f)
    assert np.allclose(get_f32(g, (w * h,)), want, rtol=0, atol=1000.0 * 2e-6 * np.abs(c).max())
    if f > 1:  # a grid that is not whole blocks is refused
        assert L.dtd_nn_block_mean(ws._ctx, w + 1, h, f, 0, p(ad), p(out)) != 0
        assert "multiples of f" in L.dtd_last_error(ws._ctx).decode()


def test_trained_coarse_correction_gains_over_the_splitting(ws):
    """The committed coarse-correction network on held-out fields at 2048x1024 (the network on
    256x128 block means): a closer first guess than the splitting on every sample, no more CG
    iterations, the same true-residual stop."""
    from dtdiff import learned as LN

    rows = LN.evaluate_coarse(LN.load_params(LN.WEIGHTS_COARSE), count=4)
    for r in rows:
        print(r, flush=True)
        assert r["err_coarse_m"] < 0.85 * r["err_lod_m"]
        assert r["iterations_coarse"] <= r["iterations_lod"]
        assert r["rel_coarse"] <= 1e-10 and r["rel_lod"] <= 1e-10
