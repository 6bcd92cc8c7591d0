"""The oracle's Jacobi solve (`step_in`, diffusion.rs:262-283) pinned on the CPU.

The reference pins its fused Jacobi solve on the bits of the floats (diffusion/parity.rs): against
the textbook loop it replaced (seven regions an iteration, `dot` as row sums added in row order),
and with the rows that have nothing to diffuse skipped (`Live`) against the solve of every row.
The same checks here: the oracle's fused solve against a textbook loop written from parity.rs:
216-270 (numpy elementwise arithmetic, the sums as sequential Python sums from -0.0, the operator
from the oracle's `apply`, itself pinned against the definition), and its skipping against its
own every-row solve, on the reference's test grids (`rough`, `ranges`, `blobs`, the pokes that
turn the skipping off, the non-finite fallback).
"""
import functools

import numpy as np
import pytest

import oracle as O

DX, DT = 4800.0, 2.0e6


def seq_sum(vals):
    """`.sum()` of an f64 iterator: left to right from -0.0."""
    return functools.reduce(lambda a, b: a + b, (float(v) for v in vals), -0.0)


def dot_rows(w, a, b):
    """parity.rs:205-212: per row the sequential sum of a*b, then the rows in order."""
    prods = a * b
    return seq_sum(seq_sum(prods[j * w:(j + 1) * w]) for j in range(len(a) // w))


def textbook_step(w, h, wrap, hgt, kd, dt):
    """`reference::step`, parity.rs:24-28 and 216-270, with CG_TOL_M = 1e-6."""
    b = np.array(hgt, dtype=np.float64)
    x = O.lod(w, h, DX, wrap, b, kd, dt)
    e, s, g = O.operator(w, h, DX, wrap, kd, dt)
    n = w * h
    ax = O.apply(w, h, wrap, e, s, g, x)
    r = b - ax
    tol2 = 1e-6 * 1e-6 * n
    its = 0
    if dot_rows(w, r, r) > tol2:
        z = r / g
        p = z.copy()
        rz = dot_rows(w, r, z)
        while its < 400:
            its += 1
            ap = O.apply(w, h, wrap, e, s, g, p)
            pap = dot_rows(w, p, ap)
            if pap <= 0.0:
                break
            alpha = rz / pap
            x = x + alpha * p
            r = r - alpha * ap
            if dot_rows(w, r, r) <= tol2:
                break
            z = r / g
            rz_new = dot_rows(w, r, z)
            beta = rz_new / rz
            rz = rz_new
            p = z + beta * p
    ax = O.apply(w, h, wrap, e, s, g, x)
    lost = seq_sum(b - ax)
    weight = seq_sum(g - 1.0)
    if weight > 0.0 and lost != 0.0:
        x = x + (lost / weight) * (g - 1.0)
    return x, its


def same_bits(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rough(w, h, seed):
    """parity.rs:374-398: +-2000 m white noise, kd cycling 0, 1e-20, 5, 300, 1e5."""
    rng = O.SplitMix64(seed)
    hgt = np.array([rng.next_f64() * 4000.0 - 2000.0 for _ in range(w * h)])
    pat = [0.0, 1.0e-20, 5.0, 300.0, 1.0e5]
    kd = np.array([pat[(c + seed) % 5] for c in range(w * h)])
    return hgt, kd


def ranges(w, h, seed, top_pct):
    """parity.rs:631-651: fractal terrain 6 km up, the spreading's diffusivity above the strength
    height `top_pct` per cent of the cells pass."""
    hgt = O.fractal(w, h, seed) + 6000.0
    n = w * h
    hs = np.sort(hgt)[min(int(n * (1.0 - top_pct / 100.0)), n - 1)]
    kd = np.array([O.spread_diffusivity(z, hs, 30000.0) for z in hgt])
    return hgt, kd


def blobs(w, h, on):
    """parity.rs:653-669: 1-3 km white noise, kd 1000 where `on`, +0 elsewhere."""
    rng = O.SplitMix64(w * 7919 + h)
    hgt = np.array([1000.0 + rng.next_f64() * 2000.0 for _ in range(w * h)])
    kd = np.array([1000.0 if on(c % w, c // w) else 0.0 for c in range(w * h)])
    return hgt, kd


def steps(fn, w, h, wrap, hgt, kd, k):
    x = np.array(hgt, dtype=np.float64)
    its = []
    for _ in range(k):
        x, i = fn(w, h, wrap, x, kd)
        its.append(i)
    return x, its


def fused(w, h, wrap, x, kd):
    return O.step_in(w, h, DX, wrap, x, kd, DT)


def every_row(w, h, wrap, x, kd):
    return O.jacobi_step(w, h, DX, wrap, x, kd, DT)


def textbook(w, h, wrap, x, kd):
    return textbook_step(w, h, wrap, x, kd, DT)


@pytest.mark.parametrize("w,h,seed,wrap", [(40, 24, 1, True), (33, 17, 11, False),
                                           (7, 9, 11, True), (2, 5, 11, True), (1, 9, 11, True),
                                           (9, 1, 11, True)])
def test_fused_jacobi_solve_is_the_textbook_solve(w, h, seed, wrap):
    """parity.rs:542-608 (rough, stiff grids, awkward shapes): the same bits and iterations."""
    hgt, kd = rough(w, h, seed)
    got, its = steps(fused, w, h, wrap, hgt, kd, 2)
    want, its_w = steps(textbook, w, h, wrap, hgt, kd, 2)
    assert same_bits(got, want) and its == its_w
    if w * h >= 100:
        assert max(its) > 5


def test_fused_jacobi_solve_with_nothing_to_solve_or_too_much():
    """parity.rs:610-629: kd = 0 (no iteration), constant kd, a flat field, the iteration cap."""
    hgt, kd = rough(32, 16, 5)
    z = np.zeros_like(kd)
    got, its = fused(32, 16, True, hgt, z)
    want, its_w = textbook(32, 16, True, hgt, z)
    assert its == 0 and its_w == 0 and same_bits(got, want)
    c = np.full_like(kd, 50.0)
    assert same_bits(steps(fused, 32, 16, True, hgt, c, 2)[0],
                     steps(textbook, 32, 16, True, hgt, c, 2)[0])
    flat = np.full_like(hgt, 100.0)
    assert same_bits(fused(32, 16, True, flat, c)[0], textbook(32, 16, True, flat, c)[0])
    # A long dt with 1e5 against 1e-20 keeps the tolerance out of reach.
    hgt, kd = rough(64, 32, 8)
    x, its = O.step_in(64, 32, DX, True, hgt, kd, 2.0e8)
    xt, its_t = textbook_step(64, 32, True, hgt, kd, 2.0e8)
    assert its == its_t and its >= 50 and same_bits(x, xt)


@pytest.mark.parametrize("w,h,seed,wrap,top", [(160, 96, 1, True, 2.0), (160, 96, 2, True, 0.5),
                                               (96, 61, 4, False, 2.0), (64, 41, 5, True, 1.0)])
def test_skipping_rows_gives_the_every_row_solve_on_ranges(w, h, seed, wrap, top):
    """parity.rs:675-699: the spreading's diffusivity on fractal ranges skips rows, and the
    result is the every-row solve's, bit for bit, over one and three steps."""
    hgt, kd = ranges(w, h, seed, top)
    live = O.live_of(w, h, DX, hgt, kd, DT)
    assert live is not None and not live.all()
    got, its = steps(fused, w, h, wrap, hgt, kd, 3)
    want, its_w = steps(every_row, w, h, wrap, hgt, kd, 3)
    assert same_bits(got, want) and its == its_w and its[0] > 0


BLOBS = [
    (64, 48, True, lambda i, j: i == 10 and j == 20),
    (64, 48, True, lambda i, j: j == 0 and i % 3 == 0),
    (64, 47, True, lambda i, j: j == 46 and i > 50),
    (64, 48, False, lambda i, j: (j == 7 or j == 8) and i < 5),
    (64, 48, True, lambda i, j: (j == 15 and i == 0) or (j == 40 and i == 63)),
    (1, 40, True, lambda i, j: j == 17),
    (2, 40, True, lambda i, j: j == 9 and i == 1),
    (3, 40, False, lambda i, j: j == 31),
    (33, 25, True, lambda i, j: j >= 16 and i == 32),
    (40, 17, True, lambda i, j: j == 16 and i == 3),
]


@pytest.mark.parametrize("k", range(len(BLOBS)))
def test_skipping_rows_gives_the_every_row_solve_on_blobs(k):
    """parity.rs:700-718: blobs at the edges of strips and of the grid."""
    w, h, wrap, on = BLOBS[k]
    hgt, kd = blobs(w, h, on)
    assert O.live_of(w, h, DX, hgt, kd, DT) is not None
    got, its = steps(fused, w, h, wrap, hgt, kd, 2)
    want, its_w = steps(every_row, w, h, wrap, hgt, kd, 2)
    assert same_bits(got, want) and its == its_w


def test_zero_subnormal_elevation_or_negative_zero_kd_turns_skipping_off():
    """parity.rs:721-750."""
    on = BLOBS[0][3]
    pokes = [0.0, -0.0, 1.0e-310, np.frombuffer(np.uint64(0x0010000000000001).tobytes(),
                                                dtype=np.float64)[0]]
    for poke in pokes:
        for row in (3, 16):
            hgt, kd = blobs(64, 48, on)
            hgt[row * 64 + 5] = poke
            assert O.live_of(64, 48, DX, hgt, kd, DT) is None
            assert same_bits(fused(64, 48, True, hgt, kd)[0], every_row(64, 48, True, hgt, kd)[0])
    hgt, kd = blobs(64, 48, on)
    kd[2 * 64 + 7] = -0.0
    assert same_bits(fused(64, 48, True, hgt, kd)[0], every_row(64, 48, True, hgt, kd)[0])
    hgt, kd = blobs(64, 48, on)
    kd[:] = 0.0
    live = O.live_of(64, 48, DX, hgt, kd, DT)
    assert live is not None and not live.any()
    x, its = fused(64, 48, True, hgt, kd)
    assert its == 0 and same_bits(x, hgt)


def test_a_non_finite_result_is_solved_again_on_every_row():
    """parity.rs:752-776: a NaN in a visited row; the step is solved on every row, and the NaN
    reaches a skipped row as in the every-row solve."""
    hgt, kd = blobs(64, 48, BLOBS[0][3])
    hgt[20 * 64 + 11] = np.nan
    assert O.live_of(64, 48, DX, hgt, kd, DT) is not None
    got, its = fused(64, 48, True, hgt, kd)
    want, its_w = every_row(64, 48, True, hgt, kd)
    assert same_bits(got, want) and its == its_w
    assert np.isnan(got[3 * 64])
