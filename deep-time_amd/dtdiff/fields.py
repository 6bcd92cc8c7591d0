"""Seeded synthetic inputs for the solve's workloads.

Random bits come from SplitMix64 (crates/deeptime-core/src/sphere/mod.rs:43-63), evaluated
vectorised: the k-th output of a generator seeded with s depends only on s + k * 0x9E3779B97F4A7C15,
so a whole field is drawn at once and equals the sequential generator's draws bit for bit.
Normals use the reference's Box-Muller form (`SplitMix64::normal`, sphere/mod.rs:71-76).

The field recipes (k^-2 spectral noise, a Gaussian-smoothed log-normal diffusivity) are this
package's own; the reference holds no generator for them.

    h0:  isotropic power spectrum proportional to |k|^-2, scaled to mean 0 and std 1000 m;
    K:   ln K ~ N(0, 1) white, blurred by a Gaussian of sigma = 8 cells (periodic in x, mirrored
         in y), renormalised to mean 0 and std 1, exponentiated and scaled to mean 1;
    dt:  chosen so that dt * mean(K) / dx^2 equals the workload's ratio, with dx = 1 m.
"""
from __future__ import annotations

import numpy as np

from . import Dims

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix_u64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Outputs start+1 .. start+n of SplitMix64(seed) (`next_u64`, sphere/mod.rs:52-58)."""
    k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def splitmix_f64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """`next_f64` (sphere/mod.rs:61-63): the top 53 bits scaled into [0, 1)."""
    return (splitmix_u64(seed, n, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 2.0**53)


def splitmix_normal(seed: int, n: int) -> np.ndarray:
    """`normal` (sphere/mod.rs:71-76) n times: u = max(next, MIN_POSITIVE), v = next,
    sqrt(-2 ln u) cos(2 pi v)."""
    f = splitmix_f64(seed, 2 * n)
    u = np.maximum(f[0::2], np.finfo(np.float64).tiny)
    v = f[1::2]
    return np.sqrt(-2.0 * np.log(u)) * np.cos(2.0 * np.pi * v)


def spectral_noise(w: int, h: int, seed: int, std: float = 1000.0) -> np.ndarray:
    """Isotropic k^-2 power-spectrum noise on an h x w grid (row-major), mean 0, given std."""
    white = splitmix_normal(seed, w * h).reshape(h, w)
    f = np.fft.rfft2(white)
    ky = np.fft.fftfreq(h)[:, None]
    kx = np.fft.rfftfreq(w)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = np.inf  # no mean
    f *= 1.0 / k      # amplitude |k|^-1: power |k|^-2
    del k
    z = np.fft.irfft2(f, s=(h, w))
    del f
    z -= z.mean()
    z *= std / z.std()
    return z.reshape(-1)


def lognormal_k(w: int, h: int, seed: int, sigma: float = 8.0) -> np.ndarray:
    """exp of a Gaussian-smoothed N(0, 1) field renormalised to std 1, scaled to mean 1."""
    from scipy.ndimage import gaussian_filter1d

    g = splitmix_normal(seed ^ 0x6B64, w * h).reshape(h, w)
    g = gaussian_filter1d(g, sigma, axis=1, mode="wrap")
    g = gaussian_filter1d(g, sigma, axis=0, mode="mirror")
    g -= g.mean()
    g /= g.std()
    k = np.exp(g)
    k /= k.mean()
    return k.reshape(-1)


def case_constant(seed: int = 1, w: int = 512, h: int = 256):
    """Constant K = 1, dt K / dx^2 = 10, k^-2 h0 (512 x 256 by default; the large case is the same
    at 16384 x 8192, seed 3)."""
    d = Dims(w, h, 1.0, True)
    return d, spectral_noise(w, h, seed), np.ones(w * h), 10.0


def case_lognormal(seed: int = 2, w: int = 8192, h: int = 4096, ratio: float = 50.0):
    """Log-normal K (sigma 8 cells), dt mean(K) / dx^2 = 50, k^-2 h0 (8192 x 4096 by default)."""
    d = Dims(w, h, 1.0, True)
    kd = lognormal_k(w, h, seed)
    return d, spectral_noise(w, h, seed), kd, ratio / float(kd.mean())


def case_belt(seed: int = 5, w: int = 8192, h: int = 4096, ratio: float = 50.0):
    """A mostly empty grid: log-normal K (sigma 8 cells) in the belt of rows [0.4 h, 0.6 h), K = +0
    on every other row (a range with nothing to diffuse around it: the rows the Jacobi step's
    `Live` skips), dt mean(K in the belt) / dx^2 = 50, k^-2 h0 offset by 2000 m (so that every
    elevation is normal)."""
    d = Dims(w, h, 1.0, True)
    kd = lognormal_k(w, h, seed).reshape(h, w)
    a, b = (4 * h) // 10, (6 * h) // 10
    kd[:a] = 0.0
    kd[b:] = 0.0
    mean = float(kd[a:b].mean()) if b > a else 1.0
    return d, spectral_noise(w, h, seed) + 2000.0, kd.reshape(-1), ratio / mean


def case_regional(seed: int = 4, count: int = 512):
    """Independent regional grids: w uniform in {64, 96, ..., 1024}, h in {64, 96, ..., 512}, each
    periodic in x with its own log-normal K (sigma 8 cells) and k^-2 h0, dt mean(K) / dx^2 uniform
    in [5, 200].  Returns a list of (Dims, h0, kd, dt)."""
    u = splitmix_u64(seed, 3 * count)
    out = []
    for k in range(count):
        w = 64 + 32 * int(u[3 * k] % np.uint64(31))
        h = 64 + 32 * int(u[3 * k + 1] % np.uint64(15))
        ratio = 5.0 + 195.0 * float(u[3 * k + 2] >> np.uint64(11)) * 2.0 ** -53
        out.append(case_lognormal(seed=1000 * seed + k, w=w, h=h, ratio=ratio))
    return out


# The benchmark's workloads by name: generator and its arguments.
WORKLOADS = {
    "constant-512x256": (case_constant, dict(seed=1, w=512, h=256)),
    "lognormal-8192x4096": (case_lognormal, dict(seed=2, w=8192, h=4096)),
    "constant-16384x8192": (case_constant, dict(seed=3, w=16384, h=8192)),
    "regional-512": (case_regional, dict(seed=4, count=512)),
    "belt-8192x4096": (case_belt, dict(seed=5, w=8192, h=4096)),
}
