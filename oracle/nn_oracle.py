"""CPU restatement of the learned first guess's network (numpy). TEST INFRASTRUCTURE ONLY.

Only tests/ load this module, as the checker of the device network (deep-time_amd/csrc/nn.hpp
and dtdiff/learned.py).  The network has no counterpart in the reference (which starts its
conjugate gradients from the splitting, diffusion.rs:110-114); what pins this restatement is
mathematics, in tests/test_oracle_nn.py: the convolution against its definition written as loops,
each backward pass as the exact adjoint of its forward pass (<A x, y> = <x, A^T y>), and the whole
network's gradient against central finite differences in f64.

Tensors are NHWC float arrays (n, h, w, c); a 3x3 convolution wraps in x and reads zeros past
the first and last rows; weights are (cout, 9, cin) with tap = 3 (dy + 1) + dx + 1.  With
`bf16=True` every stored activation and gradient is rounded to bf16 where the device stores bf16,
and sums are formed in f64 (the device's f32 accumulation differs from them by rounding only).
"""
from __future__ import annotations

import numpy as np

TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def bf16(a):
    """Round to the nearest bf16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isfinite(a), out, a).astype(np.float32)


def shift(x, dy, dx):
    """xs[n, y, x] = x[n, y + dy, (x + dx) mod w], zero where y + dy is outside the rows."""
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    rolled = np.roll(x, -dx, axis=2)
    if dy >= 0:
        out[:, :h - dy] = rolled[:, dy:]
    else:
        out[:, -dy:] = rolled[:, :h + dy]
    return out


def conv3x3(x, wt, bias=None):
    """Pre-activation of the 3x3 convolution (f64 sums)."""
    x = x.astype(np.float64)
    wt = wt.astype(np.float64)
    out = np.zeros(x.shape[:3] + (wt.shape[0],))
    for t, (dy, dx) in enumerate(TAPS):
        out += shift(x, dy, dx) @ wt[:, t, :].T
    if bias is not None:
        out += bias.astype(np.float64)
    return out


def conv3x3_bwd_input(dz, wt):
    """The adjoint of conv3x3 in its input: dx = sum over taps of shift(dz, -d) wt[:, tap]."""
    dz = dz.astype(np.float64)
    wt = wt.astype(np.float64)
    out = np.zeros(dz.shape[:3] + (wt.shape[2],))
    for t, (dy, dx) in enumerate(TAPS):
        out += shift(dz, -dy, -dx) @ wt[:, t, :]
    return out


def conv3x3_bwd_weight(x, dz):
    """dw[co, tap, ci] = sum over pixels of dz[p, co] x[p + d(tap), ci]; db[co] = sum dz[., co]."""
    x = x.astype(np.float64)
    dz = dz.astype(np.float64)
    co, ci = dz.shape[3], x.shape[3]
    dzf = dz.reshape(-1, co)
    dw = np.zeros((co, 9, ci))
    for t, (dy, dx) in enumerate(TAPS):
        dw[:, t, :] = dzf.T @ shift(x, dy, dx).reshape(-1, ci)
    return dw, dzf.sum(axis=0)


def pool(x):
    n, h, w, c = x.shape
    x = x.astype(np.float64).reshape(n, h // 2, 2, w // 2, 2, c)
    return 0.25 * ((x[:, :, 0, :, 0] + x[:, :, 0, :, 1]) + (x[:, :, 1, :, 0] + x[:, :, 1, :, 1]))


def pool_bwd(dy):
    """The adjoint of pool: each child gets 0.25 of its parent's gradient."""
    return 0.25 * np.repeat(np.repeat(dy.astype(np.float64), 2, axis=1), 2, axis=2)


def up(u):
    return np.repeat(np.repeat(u.astype(np.float64), 2, axis=1), 2, axis=2)


def up_bwd(d):
    n, h, w, c = d.shape
    d = d.astype(np.float64).reshape(n, h // 2, 2, w // 2, 2, c)
    return (d[:, :, 0, :, 0] + d[:, :, 0, :, 1]) + (d[:, :, 1, :, 0] + d[:, :, 1, :, 1])


# The U-Net: (name, cin, cout, relu), in the order of the flat parameter vector (each layer's
# weights then its bias).
LAYERS = [
    ("enc0a", 16, 32, 1), ("enc0b", 32, 32, 1), ("enc1a", 32, 64, 1), ("enc1b", 64, 64, 1),
    ("enc2a", 64, 128, 1), ("enc2b", 128, 128, 1), ("enc3a", 128, 256, 1), ("enc3b", 256, 256, 1),
    ("dec2a", 384, 128, 1), ("dec2b", 128, 128, 1), ("dec1a", 192, 64, 1), ("dec1b", 64, 64, 1),
    ("dec0a", 96, 32, 1), ("dec0b", 32, 32, 1), ("out", 32, 16, 0),
]
OUTC = 2  # output channels kept in f32: the correction to the base guess, and a spare


def offsets():
    """(weights offset, bias offset) of each layer in the flat vector, and its length."""
    offs, at = {}, 0
    for name, ci, co, _ in LAYERS:
        offs[name] = (at, at + co * 9 * ci)
        at += co * 9 * ci + co
    return offs, at


def unpack(flat):
    offs, _ = offsets()
    p = {}
    for name, ci, co, _ in LAYERS:
        a, b = offs[name]
        p[name] = (flat[a:b].reshape(co, 9, ci), flat[b:b + co])
    return p


def forward(params, x0, bf=True):
    """The network on x0 (n, h, w, 16): returns (out32 (n, h, w, OUTC), activations)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    wq = {k: (r(w), b) for k, (w, b) in params.items()}

    def layer(name, x):
        w, b = wq[name]
        z = conv3x3(x, w, b)
        relu = dict((n, f) for n, _, _, f in LAYERS)[name]
        return r(np.maximum(z, 0.0) if relu else z), z

    A = {"x0": r(x0)}
    A["e0a"], _ = layer("enc0a", A["x0"]); A["e0"], _ = layer("enc0b", A["e0a"])
    A["p1"] = r(pool(A["e0"]))
    A["e1a"], _ = layer("enc1a", A["p1"]); A["e1"], _ = layer("enc1b", A["e1a"])
    A["p2"] = r(pool(A["e1"]))
    A["e2a"], _ = layer("enc2a", A["p2"]); A["e2"], _ = layer("enc2b", A["e2a"])
    A["p3"] = r(pool(A["e2"]))
    A["e3a"], _ = layer("enc3a", A["p3"]); A["e3"], _ = layer("enc3b", A["e3a"])
    A["c2"] = np.concatenate([up(A["e3"]), A["e2"]], axis=3)
    A["d2a"], _ = layer("dec2a", A["c2"]); A["d2"], _ = layer("dec2b", A["d2a"])
    A["c1"] = np.concatenate([up(A["d2"]), A["e1"]], axis=3)
    A["d1a"], _ = layer("dec1a", A["c1"]); A["d1"], _ = layer("dec1b", A["d1a"])
    A["c0"] = np.concatenate([up(A["d1"]), A["e0"]], axis=3)
    A["d0a"], _ = layer("dec0a", A["c0"]); A["d0"], _ = layer("dec0b", A["d0a"])
    A["y"], z = layer("out", A["d0"])
    return z[..., :OUTC], A


def loss(out32, target, bf=True):
    """Mean of (out32[..., 0] - target)^2 (target = (h1 - base) / scale) and its gradient dy
    (n, h, w, 16)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    d = out32[..., 0] - target
    npx = d.size
    dy = np.zeros(out32.shape[:3] + (16,))
    dy[..., 0] = 2.0 * d / npx
    return float(np.mean(d * d)), r(dy)


def backward(params, A, dy, bf=True):
    """Gradient of the loss (through dy, the output's gradient) for every layer: {name: (dw, db)}."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    wq = {k: (r(w), b) for k, (w, b) in params.items()}
    G = {}

    def back(name, x_in, y_out, dy_out, relu, need_input=True):
        dz = r(np.where(y_out > 0, dy_out, 0.0)) if relu else dy_out
        G[name] = conv3x3_bwd_weight(x_in, dz)
        return r(conv3x3_bwd_input(dz, wq[name][0])) if need_input else None

    dd0 = back("out", A["d0"], A["y"], dy, 0)
    dd0a = back("dec0b", A["d0a"], A["d0"], dd0, 1)
    dc0 = back("dec0a", A["c0"], A["d0a"], dd0a, 1)
    dd1 = r(up_bwd(dc0[..., :64]))
    dd1a = back("dec1b", A["d1a"], A["d1"], dd1, 1)
    dc1 = back("dec1a", A["c1"], A["d1a"], dd1a, 1)
    dd2 = r(up_bwd(dc1[..., :128]))
    dd2a = back("dec2b", A["d2a"], A["d2"], dd2, 1)
    dc2 = back("dec2a", A["c2"], A["d2a"], dd2a, 1)
    de3 = r(up_bwd(dc2[..., :256]))
    de3a = back("enc3b", A["e3a"], A["e3"], de3, 1)
    dp3 = back("enc3a", A["p3"], A["e3a"], de3a, 1)
    de2 = r(dc2[..., 256:] + pool_bwd(dp3))
    de2a = back("enc2b", A["e2a"], A["e2"], de2, 1)
    dp2 = back("enc2a", A["p2"], A["e2a"], de2a, 1)
    de1 = r(dc1[..., 128:] + pool_bwd(dp2))
    de1a = back("enc1b", A["e1a"], A["e1"], de1, 1)
    dp1 = back("enc1a", A["p1"], A["e1a"], de1a, 1)
    de0 = r(dc0[..., 64:] + pool_bwd(dp1))
    de0a = back("enc0b", A["e0a"], A["e0"], de0, 1)
    back("enc0a", A["x0"], A["e0a"], de0a, 1, need_input=False)
    return G


def flat_grad(G):
    offs, n = offsets()
    g = np.zeros(n)
    for name, ci, co, _ in LAYERS:
        a, b = offs[name]
        dw, db = G[name]
        g[a:b] = dw.reshape(-1)
        g[b:b + co] = db
    return g


def input_channels(h, kd, dt, dx, scale, base=None):
    """The network's 16-channel input: h / scale, ln kd, ln(dt kd / dx^2) / 4, base / scale (0
    without a base), zeros."""
    k = np.maximum(kd, 1e-30)
    x = np.zeros(h.shape + (16,))
    x[..., 0] = h / scale
    x[..., 1] = np.log(k)
    x[..., 2] = 0.25 * np.log(dt * k / (dx * dx))
    if base is not None:
        x[..., 3] = base / scale
    return x


# ---- the learned coarse correction's transfers ------------------------------------------------

def block_mean(a, w, h, f, log_mean=False):
    """The f x f blocks' means of a (w x h, row-major); log_mean: exp of the mean of
    log(max(a, 1e-30)), the geometric mean.  f64."""
    x = np.asarray(a, dtype=np.float64).reshape(h // f, f, w // f, f)
    if log_mean:
        return np.exp(np.log(np.maximum(x, 1e-30)).mean(axis=(1, 3))).reshape(-1)
    return x.mean(axis=(1, 3)).reshape(-1)


def interp_centres(c, w, h, f):
    """Bilinear interpolation onto the fine w x h cells of the coarse values c (w/f x h/f) placed
    at the blocks' centres: periodic in x, constant past the first and last rows' centres."""
    wc, hc = w // f, h // f
    c = np.asarray(c, dtype=np.float64).reshape(hc, wc)
    X = (np.arange(w) + 0.5) / f - 0.5
    Y = (np.arange(h) + 0.5) / f - 0.5
    x0 = np.floor(X).astype(int)
    tx = X - x0
    i0, i1 = x0 % wc, (x0 + 1) % wc
    y0 = np.floor(Y).astype(int)
    ty = (Y - y0)[:, None]
    j0, j1 = np.clip(y0, 0, hc - 1), np.clip(y0 + 1, 0, hc - 1)
    a = (1 - tx) * c[j0][:, i0] + tx * c[j0][:, i1]
    b = (1 - tx) * c[j1][:, i0] + tx * c[j1][:, i1]
    return ((1 - ty) * a + ty * b).reshape(-1)
