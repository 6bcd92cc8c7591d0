"""The learned first guess: a 4-level U-Net that predicts h1 from (h0, ln kd, dt), trained on the
multigrid's own solves, run on the device by the library's network kernels (include/dtdiff.h,
`dtd_nn_*`).  No reference counterpart: the reference starts its conjugate gradients from the
splitting (crates/deeptime-core/src/surface/diffusion.rs:110-114).  The network refines that
splitting: it sees it as a fourth input and predicts the correction h1 - splitting, which is added
to the splitting in f64.  A guess only changes where the solve starts; `solve_with_guess` ends on
the same true f64 residual test as the plain solve.

Network: 3x3 convolutions (periodic in x, zero-padded past the pole rows, as the grid), 32 / 64 /
128 / 256 channels over four levels, 2x2 mean pooling down, nearest upsampling up with skip
concatenation, ReLU; bf16 weights and activations with f32 accumulation on the matrix cores, f32
master weights and Adam.  Input channels: h0 / SCALE, ln kd, ln(dt kd / dx^2) / 4, splitting /
SCALE; output channel 0 is the correction (h1 - splitting) / SCALE, channel 1 a spare.

    python -m dtdiff.learned train --steps 20000 --out weights.npz   (on a GPU)
"""
from __future__ import annotations

import argparse
import ctypes
import os
import time

import numpy as np

from . import (Dims, Stats, Workspace, load, lod_device, solve_rel_device,
               solve_rel_guess_device)
from .hipmem import DeviceArray, synchronize

SCALE = 1000.0  # metres: the k^-2 fields' standard deviation
LAYERS = [
    ("enc0a", 16, 32, 1), ("enc0b", 32, 32, 1), ("enc1a", 32, 64, 1), ("enc1b", 64, 64, 1),
    ("enc2a", 64, 128, 1), ("enc2b", 128, 128, 1), ("enc3a", 128, 256, 1), ("enc3b", 256, 256, 1),
    ("dec2a", 384, 128, 1), ("dec2b", 128, 128, 1), ("dec1a", 192, 64, 1), ("dec1b", 64, 64, 1),
    ("dec0a", 96, 32, 1), ("dec0b", 32, 32, 1), ("out", 32, 16, 0),
]
OUTC = 2
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_guess.npz")
# The learned coarse correction: the same network on f x f block means (f = COARSE).
COARSE = 8
WEIGHTS_COARSE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                              "unet_coarse.npz")


def offsets():
    offs, at = {}, 0
    for name, ci, co, _ in LAYERS:
        offs[name] = (at, at + co * 9 * ci)
        at += co * 9 * ci + co
    return offs, at


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases; the output layer starts at zero (the guess starts as h0)."""
    rng = np.random.default_rng(seed)
    offs, n = offsets()
    flat = np.zeros(n, dtype=np.float32)
    for name, ci, co, _ in LAYERS:
        a, b = offs[name]
        if name != "out":
            flat[a:b] = rng.standard_normal(b - a) * np.sqrt(2.0 / (9 * ci))
    return flat


class UNet:
    """The network's parameters and buffers on the device (one context, one input shape at a
    time)."""

    def __init__(self, ws: Workspace, params: np.ndarray | None = None):
        self.ws, self.L = ws, load()
        self.offs, self.n = offsets()
        flat = init_params() if params is None else np.asarray(params, dtype=np.float32)
        assert flat.size == self.n
        self.w = DeviceArray.from_numpy(flat)                   # f32 masters
        self.wb = DeviceArray(self.n, np.uint16)                # bf16 copies
        self.g = DeviceArray(self.n, np.float32)                # gradient
        self.m = DeviceArray.from_numpy(np.zeros(self.n, np.float32))
        self.v = DeviceArray.from_numpy(np.zeros(self.n, np.float32))
        self.step_count = 0
        self._chk(self.L.dtd_nn_to_bf16(ws._ctx, self.n, self._p(self.w), self._p(self.wb)))
        self.shape = None
        self.spec = {name: (ci, co, relu) for name, ci, co, relu in LAYERS}

    # -- plumbing --------------------------------------------------------------------------
    def _chk(self, rc):
        self.ws.check(rc)

    @staticmethod
    def _p(a, off: int = 0):
        return ctypes.c_void_p(a.data_ptr() + off * a.dtype.itemsize)

    def params(self) -> np.ndarray:
        synchronize()
        return self.w.numpy()

    def _alloc(self, n: int, h: int, w: int):
        if self.shape == (n, h, w):
            return
        if h % 8 or w % 128:
            raise ValueError("the network needs h a multiple of 8 and w of 128")
        self.shape = (n, h, w)
        px = [n * (h >> l) * (w >> l) for l in range(4)]
        B = {}

        def t(name, level, c):
            B[name] = DeviceArray(px[level] * c, np.uint16)

        t("x0", 0, 16); t("e0a", 0, 32); t("e0", 0, 32); t("p1", 1, 32)
        t("e1a", 1, 64); t("e1", 1, 64); t("p2", 2, 64); t("e2a", 2, 128); t("e2", 2, 128)
        t("p3", 3, 128); t("e3a", 3, 256); t("e3", 3, 256); t("c2", 2, 384); t("d2a", 2, 128)
        t("d2", 2, 128); t("c1", 1, 192); t("d1a", 1, 64); t("d1", 1, 64); t("c0", 0, 96)
        t("d0a", 0, 32); t("d0", 0, 32); t("y", 0, 16)
        # Gradients: two ping-pong buffers per level sized for the widest tensor there, the
        # concat gradients kept for the encoder's skip terms.
        t("g0a", 0, 96); t("g0b", 0, 96); t("g1a", 1, 192); t("g1b", 1, 192)
        t("g2a", 2, 384); t("g2b", 2, 384); t("g3a", 3, 256); t("g3b", 3, 256)
        t("dc0", 0, 96); t("dc1", 1, 192); t("dc2", 2, 384)
        B["out32"] = DeviceArray(px[0] * OUTC, np.float32)
        B["parts"] = DeviceArray(px[0] // 256 + 1, np.float32)
        planes, part = 0, 0
        for name, ci, co, _ in LAYERS:
            lvl = {"0": 0, "1": 1, "2": 2, "3": 3}[name[3]] if name != "out" else 0
            f = ctypes.c_size_t()
            pl = self.L.dtd_nn_bwd_weight_scratch(n, h >> lvl, w >> lvl, ci, co, ctypes.byref(f))
            planes, part = max(planes, pl), max(part, f.value)
        B["planes"] = DeviceArray(planes, np.uint16)
        B["part"] = DeviceArray(part, np.float32)
        B["wflip"] = DeviceArray(max(co * 9 * ci for _, ci, co, _ in LAYERS), np.uint16)
        self.B = B

    def _level_shape(self, lvl):
        n, h, w = self.shape
        return n, h >> lvl, w >> lvl

    # -- forward -----------------------------------------------------------------------------
    def _conv(self, name, lvl, x, y, out32=None):
        ci, co, relu = self.spec[name]
        a, b = self.offs[name]
        n, h, w = self._level_shape(lvl)
        self._chk(self.L.dtd_nn_conv3x3(self.ws._ctx, n, h, w, ci, co, self._p(x),
                                        self._p(self.wb, a), self._p(self.w, b), relu,
                                        self._p(y) if y is not None else None,
                                        self._p(out32) if out32 else None,
                                        OUTC if out32 else 0))

    def _pool(self, lvl, c, x, y):
        n, h, w = self._level_shape(lvl)
        self._chk(self.L.dtd_nn_pool(self.ws._ctx, n, h, w, c, self._p(x), self._p(y)))

    def _up_cat(self, lvl, cu, cs, u, skip, cat):
        n, h, w = self._level_shape(lvl)
        self._chk(self.L.dtd_nn_up_cat(self.ws._ctx, n, h, w, cu, cs, self._p(u),
                                       self._p(skip), self._p(cat)))

    def set_input(self, d: Dims, h_dev, kd_dev, base_dev, dt: float, n: int = 1):
        """The input channels from device arrays h (n images stacked), kd and the base guess."""
        self._alloc(n, d.h, d.w)
        self._chk(self.L.dtd_nn_input(self.ws._ctx, n * d.w * d.h, self._p(h_dev),
                                      self._p(kd_dev), self._p(base_dev), dt, d.dx, SCALE,
                                      self._p(self.B["x0"])))

    def _conv_cat(self, name, lvl, cu, u, skip, y):
        """Layer `name` on [upsampled u, skip] read in place (no concatenation formed)."""
        ci, co, relu = self.spec[name]
        a, b = self.offs[name]
        n, h, w = self._level_shape(lvl)
        self._chk(self.L.dtd_nn_conv3x3_cat(self.ws._ctx, n, h, w, cu, ci - cu, co, self._p(u),
                                            self._p(skip), self._p(self.wb, a),
                                            self._p(self.w, b), relu, self._p(y), None, 0))

    def forward(self, train: bool = True):
        """The network on the input set by set_input.  train keeps every tensor the backward pass
        reads; without it the decoder reads its concatenations in place (w a multiple of 512 at
        level 0, so that the rows of levels 0-2 are multiples of 128)."""
        B = self.B
        n, h, w = self.shape
        if not train and w % 512 == 0:
            return self._forward_infer()
        self._conv("enc0a", 0, B["x0"], B["e0a"]); self._conv("enc0b", 0, B["e0a"], B["e0"])
        self._pool(0, 32, B["e0"], B["p1"])
        self._conv("enc1a", 1, B["p1"], B["e1a"]); self._conv("enc1b", 1, B["e1a"], B["e1"])
        self._pool(1, 64, B["e1"], B["p2"])
        self._conv("enc2a", 2, B["p2"], B["e2a"]); self._conv("enc2b", 2, B["e2a"], B["e2"])
        self._pool(2, 128, B["e2"], B["p3"])
        self._conv("enc3a", 3, B["p3"], B["e3a"]); self._conv("enc3b", 3, B["e3a"], B["e3"])
        self._up_cat(2, 256, 128, B["e3"], B["e2"], B["c2"])
        self._conv("dec2a", 2, B["c2"], B["d2a"]); self._conv("dec2b", 2, B["d2a"], B["d2"])
        self._up_cat(1, 128, 64, B["d2"], B["e1"], B["c1"])
        self._conv("dec1a", 1, B["c1"], B["d1a"]); self._conv("dec1b", 1, B["d1a"], B["d1"])
        self._up_cat(0, 64, 32, B["d1"], B["e0"], B["c0"])
        self._conv("dec0a", 0, B["c0"], B["d0a"]); self._conv("dec0b", 0, B["d0a"], B["d0"])
        self._conv("out", 0, B["d0"], B["y"], out32=B["out32"])

    def _forward_infer(self):
        B = self.B
        self._conv("enc0a", 0, B["x0"], B["e0a"]); self._conv("enc0b", 0, B["e0a"], B["e0"])
        self._pool(0, 32, B["e0"], B["p1"])
        self._conv("enc1a", 1, B["p1"], B["e1a"]); self._conv("enc1b", 1, B["e1a"], B["e1"])
        self._pool(1, 64, B["e1"], B["p2"])
        self._conv("enc2a", 2, B["p2"], B["e2a"]); self._conv("enc2b", 2, B["e2a"], B["e2"])
        self._pool(2, 128, B["e2"], B["p3"])
        self._conv("enc3a", 3, B["p3"], B["e3a"]); self._conv("enc3b", 3, B["e3a"], B["e3"])
        self._conv_cat("dec2a", 2, 256, B["e3"], B["e2"], B["d2a"])
        self._conv("dec2b", 2, B["d2a"], B["d2"])
        self._conv_cat("dec1a", 1, 128, B["d2"], B["e1"], B["d1a"])
        self._conv("dec1b", 1, B["d1a"], B["d1"])
        self._conv_cat("dec0a", 0, 64, B["d1"], B["e0"], B["d0a"])
        self._conv("dec0b", 0, B["d0a"], B["d0"])
        self._conv("out", 0, B["d0"], None, out32=B["out32"])  # the guess reads only out32

    # -- backward ----------------------------------------------------------------------------
    def _back(self, name, lvl, x_in, y_out, dy, dz, dx=None):
        """Layer `name`'s weight gradient into self.g, and its input gradient into dx (if given);
        dz is scratch for the pre-activation gradient (may be dy itself without ReLU)."""
        ci, co, relu = self.spec[name]
        a, b = self.offs[name]
        n, h, w = self._level_shape(lvl)
        ctx, L = self.ws._ctx, self.L
        if relu:
            self._chk(L.dtd_nn_relu_bwd(ctx, n * h * w * co, self._p(y_out), self._p(dy),
                                        self._p(dz)))
        else:
            dz = dy
        B = self.B
        self._chk(L.dtd_nn_conv3x3_bwd_weight(ctx, n, h, w, ci, co, self._p(x_in), self._p(dz),
                                              self._p(B["planes"]), B["planes"].n,
                                              self._p(B["part"]), B["part"].n,
                                              self._p(self.g, a), self._p(self.g, b)))
        if dx is not None:
            self._chk(L.dtd_nn_conv3x3_bwd_input(ctx, n, h, w, ci, co, self._p(dz),
                                                 self._p(self.wb, a), self._p(B["wflip"]),
                                                 self._p(dx)))

    def _up_cat_bwd(self, lvl, cu, cs, dcat, dpool, du, dskip):
        n, h, w = self._level_shape(lvl)
        self._chk(self.L.dtd_nn_up_cat_bwd(self.ws._ctx, n, h, w, cu, cs, self._p(dcat),
                                           self._p(dpool) if dpool else None,
                                           self._p(du) if du else None,
                                           self._p(dskip) if dskip else None))

    def backward(self, dy):
        """Gradient of the loss whose output gradient is dy (16 channels at level 0)."""
        B = self.B
        g0a, g0b, g1a, g1b = B["g0a"], B["g0b"], B["g1a"], B["g1b"]
        g2a, g2b, g3a, g3b = B["g2a"], B["g2b"], B["g3a"], B["g3b"]
        self._back("out", 0, B["d0"], B["y"], dy, None, dx=g0a)
        self._back("dec0b", 0, B["d0a"], B["d0"], g0a, g0b, dx=g0a)
        self._back("dec0a", 0, B["c0"], B["d0a"], g0a, g0b, dx=B["dc0"])
        self._up_cat_bwd(0, 64, 32, B["dc0"], None, g1a, None)
        self._back("dec1b", 1, B["d1a"], B["d1"], g1a, g1b, dx=g1a)
        self._back("dec1a", 1, B["c1"], B["d1a"], g1a, g1b, dx=B["dc1"])
        self._up_cat_bwd(1, 128, 64, B["dc1"], None, g2a, None)
        self._back("dec2b", 2, B["d2a"], B["d2"], g2a, g2b, dx=g2a)
        self._back("dec2a", 2, B["c2"], B["d2a"], g2a, g2b, dx=B["dc2"])
        self._up_cat_bwd(2, 256, 128, B["dc2"], None, g3a, None)
        self._back("enc3b", 3, B["e3a"], B["e3"], g3a, g3b, dx=g3a)
        self._back("enc3a", 3, B["p3"], B["e3a"], g3a, g3b, dx=g3a)     # dp3 in g3a
        self._up_cat_bwd(2, 256, 128, B["dc2"], g3a, None, g2a)           # de2
        self._back("enc2b", 2, B["e2a"], B["e2"], g2a, g2b, dx=g2a)
        self._back("enc2a", 2, B["p2"], B["e2a"], g2a, g2b, dx=g2a)     # dp2
        self._up_cat_bwd(1, 128, 64, B["dc1"], g2a, None, g1a)            # de1
        self._back("enc1b", 1, B["e1a"], B["e1"], g1a, g1b, dx=g1a)
        self._back("enc1a", 1, B["p1"], B["e1a"], g1a, g1b, dx=g1a)     # dp1
        self._up_cat_bwd(0, 64, 32, B["dc0"], g1a, None, g0a)             # de0
        self._back("enc0b", 0, B["e0a"], B["e0"], g0a, g0b, dx=g0a)
        self._back("enc0a", 0, B["x0"], B["e0a"], g0a, g0b)

    # -- training and the guess ----------------------------------------------------------------
    def loss(self, t_dev, base_dev) -> tuple[float, DeviceArray]:
        """Mean squared error of the predicted correction against (t - base) / SCALE (device, f64,
        metres) and the output gradient."""
        n, h, w = self.shape
        out = ctypes.c_double()
        dy = self.B["g0b"]  # 16-channel view of a level-0 gradient buffer
        self._chk(self.L.dtd_nn_loss(self.ws._ctx, n * h * w, self._p(self.B["out32"]), OUTC,
                                     self._p(t_dev), self._p(base_dev), SCALE,
                                     self._p(self.B["parts"]), self._p(dy), ctypes.byref(out)))
        return out.value, dy

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step_count += 1
        self._chk(self.L.dtd_nn_adam(self.ws._ctx, self.n, self._p(self.w), self._p(self.m),
                                     self._p(self.v), self._p(self.g), lr, beta1, beta2, eps,
                                     self.step_count, self._p(self.wb)))

    def guess(self, d: Dims, h_dev, kd_dev, dt: float, g_dev) -> None:
        """g = the network's h1 for (h, kd, dt) on the device (one image): the splitting into g,
        then g += SCALE x the predicted correction."""
        lod_device(self.ws, d, h_dev, kd_dev, dt, g_dev)
        self.set_input(d, h_dev, kd_dev, g_dev, dt)
        self.forward(train=False)
        self._chk(self.L.dtd_nn_guess(self.ws._ctx, d.w * d.h, self._p(g_dev),
                                      self._p(self.B["out32"]), OUTC, SCALE, self._p(g_dev)))

    def flat_grad(self) -> np.ndarray:
        synchronize()
        self._chk(self.L.dtd_sync(self.ws._ctx))
        return self.g.numpy()


class CoarseGuess:
    """The learned coarse correction of the first guess: the network at 1/f of the grid's
    resolution.  Its inputs are the f x f block means of h0, kd (geometric) and the splitting (the
    diffusion number's channel still in fine cells, so the network sees the same physics at any
    grid size); it predicts the block means of (h1 - splitting) / SCALE, which are interpolated
    bilinearly between the blocks' centres and added to the splitting in f64.  At 8192 x 4096 the
    network runs on 1024 x 512 pixels (1/64 of the full-resolution guess's work).  Needs w a
    multiple of 128 f and h of 8 f."""

    def __init__(self, ws: Workspace, params: np.ndarray | None = None, f: int = COARSE):
        self.ws, self.f, self.L = ws, f, load()
        self.net = UNet(ws, params)
        self.shape = None

    def _alloc(self, d: Dims):
        if self.shape == (d.w, d.h):
            return
        if d.w % (128 * self.f) or d.h % (8 * self.f):
            raise ValueError(f"the coarse correction needs w a multiple of {128 * self.f} and h "
                             f"of {8 * self.f}")
        self.shape = (d.w, d.h)
        nc = (d.w // self.f) * (d.h // self.f)
        self.dc = Dims(d.w // self.f, d.h // self.f, d.dx, d.wrap_x)
        self.hc, self.kc, self.bc, self.tc = (DeviceArray(nc) for _ in range(4))

    def _mean(self, d: Dims, src, dst, log_mean: int = 0):
        self.ws.check(self.L.dtd_nn_block_mean(self.ws._ctx, d.w, d.h, self.f, log_mean,
                                               UNet._p(src), UNet._p(dst)))

    def set_input(self, d: Dims, h_dev, kd_dev, base_dev, dt: float):
        self._alloc(d)
        self._mean(d, h_dev, self.hc)
        self._mean(d, kd_dev, self.kc, 1)
        self._mean(d, base_dev, self.bc)
        self.net.set_input(Dims(self.dc.w, self.dc.h, d.dx, d.wrap_x), self.hc, self.kc,
                           self.bc, dt)

    def loss(self, d: Dims, t_dev):
        """Against the block means of the solve t (the splitting's already in self.bc)."""
        self._mean(d, t_dev, self.tc)
        return self.net.loss(self.tc, self.bc)

    def guess(self, d: Dims, h_dev, kd_dev, dt: float, g_dev) -> None:
        """g = the splitting + the interpolated coarse correction, on the device."""
        lod_device(self.ws, d, h_dev, kd_dev, dt, g_dev)
        self.set_input(d, h_dev, kd_dev, g_dev, dt)
        self.net.forward(train=False)
        self.ws.check(self.L.dtd_nn_guess_coarse(self.ws._ctx, d.w, d.h, self.f,
                                                 UNet._p(g_dev), UNet._p(self.net.B["out32"]),
                                                 OUTC, SCALE, UNet._p(g_dev)))


def _bias_mask() -> np.ndarray:
    offs, n = offsets()
    m = np.zeros(n, dtype=bool)
    for name, ci, co, _ in LAYERS:
        m[offs[name][1]:offs[name][1] + co] = True
    return m


def save(path: str, params: np.ndarray, **meta) -> None:
    """The weights as the network computes with them (bf16, stored as their upper 16 bits) and
    the biases in f32."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    p = np.ascontiguousarray(params, dtype=np.float32)
    u = p.view(np.uint32).astype(np.uint64)
    bits = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)
    mask = _bias_mask()
    np.savez_compressed(path, weights_bf16=bits, biases=p[mask],
                        **{k: np.asarray(v) for k, v in meta.items()})


def load_params(path: str = WEIGHTS) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        flat = (z["weights_bf16"].astype(np.uint32) << 16).view(np.float32).copy()
        flat[_bias_mask()] = z["biases"]
    return flat


def solve_with_guess(ws: Workspace, net: UNet, d: Dims, h_dev, kd_dev, dt: float,
                     rel_tol: float, g_dev=None) -> tuple[Stats, float]:
    """The solve started from the network's guess; returns its stats and the guess's device time
    (ms, network included)."""
    g = g_dev if g_dev is not None else DeviceArray(d.w * d.h)
    ws.check(load().dtd_sync(ws._ctx))
    t0 = time.perf_counter()
    net.guess(d, h_dev, kd_dev, dt, g)
    ws.check(load().dtd_sync(ws._ctx))
    ms = 1e3 * (time.perf_counter() - t0)
    st = solve_rel_guess_device(ws, d, h_dev, kd_dev, dt, rel_tol, g)
    return st, ms


# -- training data: the multigrid's own solves -------------------------------------------------

class Samples:
    """Training pairs (h0, kd, dt) -> h1 at w x h: h0 k^-2 noise and kd constant or log-normal
    (sigma 8 cells) from pools of seeded fields, each sample a random pairing, x shift, row flip,
    sign and dt kd / dx^2 ratio in [5, 200]; h1 from the multigrid solve to rel. residual 1e-9."""

    def __init__(self, ws: Workspace, w: int = 1024, h: int = 512, pool: int = 32, seed: int = 7):
        from . import fields as F

        self.ws, self.w, self.h = ws, w, h
        self.d = Dims(w, h, 1.0, True)
        self.noise = [F.spectral_noise(w, h, seed * 1000 + i) for i in range(pool)]
        self.kd = [np.ones(w * h)] + [F.lognormal_k(w, h, seed * 1000 + 500 + i)
                                      for i in range(pool - 1)]
        self.rng = np.random.default_rng(seed)
        self.h_dev, self.k_dev, self.t_dev, self.b_dev = (DeviceArray(w * h) for _ in range(4))

    def draw(self):
        self.ws.check(load().dtd_sync(self.ws._ctx))  # the last step's kernels are done with
        r = self.rng
        hh = self.noise[r.integers(len(self.noise))].reshape(self.h, self.w)
        kk = self.kd[0 if r.random() < 0.25 else r.integers(1, len(self.kd))]
        kk = kk.reshape(self.h, self.w)
        sx = int(r.integers(self.w))
        hh = np.roll(hh, sx, axis=1)
        kk = np.roll(kk, int(r.integers(self.w)), axis=1)
        if r.random() < 0.5:
            hh = hh[::-1]
        if r.random() < 0.5:
            hh = -hh
        ratio = float(np.exp(r.uniform(np.log(5.0), np.log(200.0))))
        self.h_dev.upload(hh.reshape(-1))
        self.k_dev.upload(kk.reshape(-1))
        self.t_dev.copy_from(self.h_dev)
        st = solve_rel_device(self.ws, self.d, self.t_dev, self.k_dev, ratio, 1e-9)
        lod_device(self.ws, self.d, self.h_dev, self.k_dev, ratio, self.b_dev)
        return ratio, st


def train(steps: int, out: str, lr: float = 1e-3, w: int = 1024, h: int = 512, pool: int = 32,
          log_every: int = 100, params: np.ndarray | None = None, seed: int = 7) -> np.ndarray:
    ws = Workspace()
    S = Samples(ws, w, h, pool, seed)
    net = UNet(ws, params)
    t0 = time.perf_counter()
    run = 0.0
    for k in range(1, steps + 1):
        ratio, _ = S.draw()
        net.set_input(S.d, S.h_dev, S.k_dev, S.b_dev, ratio)
        net.forward()
        loss, dy = net.loss(S.t_dev, S.b_dev)
        net.backward(dy)
        # Cosine decay from lr to lr / 20.
        net.adam(lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * k / steps))))
        run = loss if k == 1 else 0.98 * run + 0.02 * loss
        if k % log_every == 0 or k == 1:
            print(f"step {k:6d}  loss {loss:.3e}  (avg {run:.3e})  "
                  f"{time.perf_counter() - t0:7.1f} s", flush=True)
    params = net.params()
    save(out, params, steps=steps, lr=lr, w=w, h=h, final_loss=run)
    return params


def train_coarse(steps: int, out: str, lr: float = 1e-3, w: int = 2048, h: int = 1024,
                 pool: int = 32, log_every: int = 100, params: np.ndarray | None = None,
                 seed: int = 11) -> np.ndarray:
    """The coarse correction's network on the multigrid's own solves at w x h (the network at
    w / 8 x h / 8), loss on the block means of the splitting's error."""
    ws = Workspace()
    S = Samples(ws, w, h, pool, seed)
    cg = CoarseGuess(ws, params)
    t0 = time.perf_counter()
    run = 0.0
    for k in range(1, steps + 1):
        ratio, _ = S.draw()
        cg.set_input(S.d, S.h_dev, S.k_dev, S.b_dev, ratio)
        cg.net.forward()
        loss, dy = cg.loss(S.d, S.t_dev)
        cg.net.backward(dy)
        cg.net.adam(lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * k / steps))))
        run = loss if k == 1 else 0.98 * run + 0.02 * loss
        if k % log_every == 0 or k == 1:
            print(f"step {k:6d}  loss {loss:.3e}  (avg {run:.3e})  "
                  f"{time.perf_counter() - t0:7.1f} s", flush=True)
    params = cg.net.params()
    save(out, params, steps=steps, lr=lr, w=w, h=h, f=COARSE, final_loss=run)
    return params


def evaluate_coarse(params: np.ndarray, count: int = 8, w: int = 2048, h: int = 1024,
                    seed: int = 99, rel_tol: float = 1e-10) -> list[dict]:
    """Held-out samples: iterations and guess errors from the splitting and from the splitting
    with the learned coarse correction, each solve to the same true relative residual."""
    ws = Workspace()
    S = Samples(ws, w, h, pool=max(2, count // 2), seed=seed)
    cg = CoarseGuess(ws, params)
    rows = []
    g = DeviceArray(w * h)
    for k in range(count):
        ratio, _ = S.draw()
        x = DeviceArray.from_numpy(S.h_dev.numpy())
        plain = solve_rel_device(ws, S.d, x, S.k_dev, ratio, rel_tol)
        x.upload(S.h_dev.numpy())
        ws.check(load().dtd_sync(ws._ctx))
        t0 = time.perf_counter()
        cg.guess(S.d, x, S.k_dev, ratio, g)
        ws.check(load().dtd_sync(ws._ctx))
        ms = 1e3 * (time.perf_counter() - t0)
        st = solve_rel_guess_device(ws, S.d, x, S.k_dev, ratio, rel_tol, g)
        synchronize()
        t, gg, b = S.t_dev.numpy(), g.numpy(), S.b_dev.numpy()
        rows.append({"ratio": ratio, "iterations_lod": plain.iterations,
                     "iterations_coarse": st.iterations, "rel_lod": plain.rel_residual,
                     "rel_coarse": st.rel_residual,
                     "err_lod_m": float(np.sqrt(np.mean((b - t) ** 2))),
                     "err_coarse_m": float(np.sqrt(np.mean((gg - t) ** 2))), "guess_ms": ms})
    ws.close()
    return rows


def evaluate(params: np.ndarray, count: int = 8, w: int = 1024, h: int = 512, seed: int = 99,
             rel_tol: float = 1e-10) -> list[dict]:
    """Held-out samples (fields from seeds outside the training pools): the CG iterations from the
    splitting and from the network's guess, each solve to the same true relative residual."""
    ws = Workspace()
    S = Samples(ws, w, h, pool=max(2, count // 2), seed=seed)
    net = UNet(ws, params)
    rows = []
    g = DeviceArray(w * h)
    for k in range(count):
        ratio, _ = S.draw()
        x = DeviceArray.from_numpy(S.h_dev.numpy())
        plain = solve_rel_device(ws, S.d, x, S.k_dev, ratio, rel_tol)
        x.upload(S.h_dev.numpy())
        st, ms = solve_with_guess(ws, net, S.d, x, S.k_dev, ratio, rel_tol, g)
        synchronize()
        t, gg, b = S.t_dev.numpy(), g.numpy(), S.b_dev.numpy()
        rows.append({"ratio": ratio, "iterations_lod": plain.iterations,
                     "iterations_net": st.iterations, "rel_lod": plain.rel_residual,
                     "rel_net": st.rel_residual,
                     "err_lod_m": float(np.sqrt(np.mean((b - t) ** 2))),
                     "err_net_m": float(np.sqrt(np.mean((gg - t) ** 2))), "guess_ms": ms})
    ws.close()
    return rows


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--steps", type=int, default=20000)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--w", type=int, default=1024)
    t.add_argument("--h", type=int, default=512)
    t.add_argument("--pool", type=int, default=32)
    t.add_argument("--init", default=None, help="start from these weights (fine-tuning)")
    t.add_argument("--seed", type=int, default=7, help="the training fields' pool seed")
    t.add_argument("--coarse", action="store_true",
                   help="the coarse correction's network (default size 2048 x 1024)")
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=None)
    e.add_argument("--count", type=int, default=8)
    e.add_argument("--coarse", action="store_true")
    a = ap.parse_args()
    if a.cmd == "train":
        init = load_params(a.init) if a.init else None
        if a.coarse:
            w, h = (a.w, a.h) if (a.w, a.h) != (1024, 512) else (2048, 1024)
            train_coarse(a.steps, a.out if a.out != WEIGHTS else WEIGHTS_COARSE, a.lr, w, h,
                         a.pool, params=init, seed=a.seed)
        else:
            train(a.steps, a.out, a.lr, a.w, a.h, a.pool, params=init, seed=a.seed)
    elif a.coarse:
        for r in evaluate_coarse(load_params(a.weights or WEIGHTS_COARSE), a.count):
            print(" ".join(f"{k} {v:.4g}" if isinstance(v, float) else f"{k} {v}"
                           for k, v in r.items()), flush=True)
    else:
        a.weights = a.weights or WEIGHTS
        for r in evaluate(load_params(a.weights), a.count):
            print(" ".join(f"{k} {v:.4g}" if isinstance(v, float) else f"{k} {v}"
                           for k, v in r.items()), flush=True)


if __name__ == "__main__":
    main()
