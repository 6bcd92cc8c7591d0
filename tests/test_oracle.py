"""Pins of the CPU oracle (oracle/dtoracle.c) before it is trusted as the checker.

Each pin is chosen so that a plausible mistake in the restatement (a dropped term, a wrong
index or sign, a transposed face) fails it:
  * published SplitMix64 vectors (tests/golden/splitmix64.json);
  * the reference's known-answer ring test (diffusion.rs:1084-1106);
  * mathematics: the closed-form eigen-solution for constant K, dense solves of the
    definitional operator on small grids, A 1 = 1 and symmetry, the Galerkin identity of the
    aggregation levels, the locally one-dimensional splitting as two dense 1-D solves, damped
    Jacobi and a two-level V-cycle in dense matrix form;
  * the reference's own tests, restated: multigrid/tests.rs:181-301 and surface/tests.rs:110-140.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))

# multigrid/tests.rs:166-179
SIZES = [(1, 1), (1, 7), (7, 1), (2, 2), (2, 5), (3, 5), (5, 3), (17, 9), (33, 33), (64, 40),
         (129, 67), (200, 96)]


def rms(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)))


def dense(w, h, dx, wrap, kd, dt):
    """The definitional operator (multigrid/tests.rs:17-43, restated in the oracle) as a matrix."""
    n = w * h
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        A[:, c] = O.apply_reference(w, h, dx, wrap, kd, dt, e)
    return A


# ---- fixtures --------------------------------------------------------------------------------

def test_splitmix64_published_vectors():
    g = json.load(open(os.path.join(HERE, "golden", "splitmix64.json")))
    for v in g["vectors"]:
        s = O.SplitMix64(v["seed"])
        assert [s.next_u64() for _ in v["outputs"]] == [int(x) for x in v["outputs"]]


def test_next_f64_is_the_top_53_bits():
    a, b = O.SplitMix64(99), O.SplitMix64(99)
    for _ in range(100):
        u = a.next_u64()
        assert b.next_f64() == (u >> 11) * 2.0**-53


def test_vectorised_splitmix_matches_the_sequential_one():
    from dtdiff import fields as F
    s = O.SplitMix64(424242)
    seq = np.array([s.next_f64() for _ in range(3000)])
    assert np.array_equal(seq, F.splitmix_f64(424242, 3000))


def test_submarine_kd_limits():
    # Deep water below the wave base decays as exp(-(z - wb)/wb); on land it is `land`.
    d0, wb, tr, land = 100.0, 60.0, 2.0, 3.0
    assert abs(O.submarine_kd(-1e4, d0, wb, tr, land) - land) < 1e-12
    z = 500.0
    f1 = (np.tanh(z / tr) + 1) / 2
    want = d0 * f1 * np.exp(-(z - wb) / wb) + land * (1 - f1)
    assert abs(O.submarine_kd(z, d0, wb, tr, land) - want) <= 1e-15 * want
    assert O.submarine_kd(0.0, d0, wb, tr, land) == pytest.approx(0.5 * d0 + 0.5 * land)


# ---- the operator ----------------------------------------------------------------------------

@pytest.mark.parametrize("w,h,wrap", [(1, 1, True), (1, 5, True), (2, 3, True), (2, 3, False),
                                      (3, 4, True), (5, 3, False), (7, 6, True)])
def test_operator_is_symmetric_and_rows_sum_to_one(w, h, wrap):
    b = O.terrain(w, h, 3)
    kd = O.shelf_kd(b, 3)
    e, s, g = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    n = w * h
    A = np.zeros((n, n))
    for c in range(n):
        v = np.zeros(n)
        v[c] = 1.0
        A[:, c] = O.apply(w, h, wrap, e, s, g, v)
    assert np.allclose(A, A.T, rtol=0, atol=1e-12 * np.abs(A).max())
    assert np.allclose(A.sum(axis=1), 1.0, rtol=1e-12)
    Ad = dense(w, h, 2400.0, wrap, kd, 2e6)
    assert np.allclose(A, Ad, rtol=0, atol=1e-12 * np.abs(A).max())
    # Positive definite with every eigenvalue >= 1.
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() >= 1 - 1e-9


def test_a_two_cell_ring_has_two_faces_and_one_cell_none():
    kd = np.array([2.0, 4.0])
    e, s, g = O.operator(2, 1, 1.0, True, kd, 1.0)
    # Both faces join the same pair: the coupling is counted twice on the diagonal.
    assert e[0] == 3.0 and e[1] == 3.0 and g[0] == 7.0 and g[1] == 7.0
    e, s, g = O.operator(1, 3, 1.0, True, np.ones(3), 1.0)
    assert (e == 0).all() and list(g) == [2.0, 3.0, 2.0]


@pytest.mark.parametrize("w,h", [(16, 8), (24, 10), (3, 5)])
def test_closed_form_solution_for_constant_k(w, h):
    """Periodic x, no-flux y: cos(2 pi k i / w) cos(pi l (j + 1/2) / h) is an eigenvector of the
    five-point operator with eigenvalue 1 + c (4 sin^2(pi k / w) + 4 sin^2(pi l / (2 h)))."""
    c = 10.0
    kd = np.ones(w * h)
    for k, l in [(1, 0), (2, 3), (0, 1), (w // 2, h - 1)]:
        i = np.arange(w)[None, :]
        j = np.arange(h)[:, None]
        b = (np.cos(2 * np.pi * k * i / w) * np.cos(np.pi * l * (j + 0.5) / h)).reshape(-1) * 1e3
        lam = 1 + c * (4 * np.sin(np.pi * k / w) ** 2 + 4 * np.sin(np.pi * l / (2 * h)) ** 2)
        x, _ = O.solve(w, h, 1.0, True, b, kd, c, 1e-13 * rms(b, 0) + 1e-300)
        assert np.allclose(x, b / lam, rtol=0, atol=1e-9 * np.abs(b).max())


@pytest.mark.parametrize("w,h", [(w, h) for (w, h) in SIZES if w * h <= 600])
@pytest.mark.parametrize("wrap", [True, False])
def test_solve_matches_a_dense_cholesky_solve(w, h, wrap):
    b = O.terrain(w, h, 5 + w)
    kd = O.shelf_kd(b, 5)
    A = dense(w, h, 2400.0, wrap, kd, 2e6)
    L = np.linalg.cholesky(0.5 * (A + A.T))
    exact = np.linalg.solve(L.T, np.linalg.solve(L, b))
    x, its = O.solve(w, h, 2400.0, wrap, b, kd, 2e6, 1e-11)
    assert its < 400
    assert rms(x, exact) <= 1e-9
    xr = O.reference_cg(w, h, 2400.0, wrap, b, kd, 2e6)
    assert rms(xr, exact) <= 1e-9


# ---- the first guess -------------------------------------------------------------------------

def test_cyclic_solves_a_ring():
    """diffusion.rs:1084-1106."""
    n = 5
    k = [0.3, 1.2, 0.7, 2.0, 0.1]
    lo = np.array([-k[(i + n - 1) % n] for i in range(n)])
    up = np.array([-k[i] for i in range(n)])
    di = np.array([1.0 + k[(i + n - 1) % n] + k[i] for i in range(n)])
    b0 = np.array([1.0, -2.0, 3.0, 0.5, 4.0])
    x = O.cyclic(lo, di, up, up[n - 1], lo[0], b0)
    for i in range(n):
        ax = di[i] * x[i] + lo[i] * x[(i + n - 1) % n] + up[i] * x[(i + 1) % n]
        assert abs(ax - b0[i]) < 1e-10


def test_thomas_matches_a_dense_solve():
    rng = np.random.default_rng(0)
    n = 40
    lo, up = -rng.random(n), -rng.random(n)
    di = 1.0 + rng.random(n) + np.abs(lo) + np.abs(up)
    b = rng.standard_normal(n)
    M = np.diag(di) + np.diag(lo[1:], -1) + np.diag(up[:-1], 1)
    assert np.allclose(O.thomas(lo, di, up, b), np.linalg.solve(M, b), rtol=1e-12, atol=1e-12)


def _faces_1d(k, dt, s2, ring):
    n = len(k)
    f = np.zeros(n)
    for i in range(n):
        if i + 1 < n:
            f[i] = 0.5 * (k[i] + k[i + 1]) * dt / s2
        elif ring and n > 1:
            f[i] = 0.5 * (k[n - 1] + k[0]) * dt / s2
    return f


def _tri(f, ring):
    """I + dt L along one line with faces f (f[i] between i and i+1, the last one wrapping)."""
    n = len(f)
    M = np.eye(n)
    for i in range(n):
        if i + 1 < n or (ring and n > 1):
            a, b = i, (i + 1) % n
            M[a, a] += f[i]
            M[b, b] += f[i]
            M[a, b] -= f[i]
            M[b, a] -= f[i]
    return M


@pytest.mark.parametrize("w,h,wrap", [(1, 1, True), (1, 6, True), (2, 5, True), (2, 5, False),
                                      (3, 4, True), (9, 7, True), (9, 7, False), (16, 3, True)])
def test_lod_is_two_dense_one_dimensional_solves(w, h, wrap):
    b = O.terrain(w, h, 7)
    kd = O.shelf_kd(b, 7)
    dt, dx = 2e6, 2400.0
    s2 = dx * dx
    x = b.reshape(h, w).copy()
    for j in range(h):
        x[j] = np.linalg.solve(_tri(_faces_1d(kd.reshape(h, w)[j], dt, s2, wrap), wrap), x[j])
    for i in range(w):
        col = kd.reshape(h, w)[:, i]
        x[:, i] = np.linalg.solve(_tri(_faces_1d(col, dt, s2, False), False), x[:, i])
    got = O.lod(w, h, dx, wrap, b, kd, dt)
    assert np.allclose(got, x.reshape(-1), rtol=0, atol=1e-10 * np.abs(b).max())


# ---- the cycle -------------------------------------------------------------------------------

def _level_matrix(L):
    """A level's f32 operator (as f64) from its east/south/diag arrays."""
    w, h, wrap = L["w"], L["h"], L["wrap"]
    n = w * h
    A = np.diag(L["diag"].astype(np.float64))
    e, s = L["east"].astype(np.float64), L["south"].astype(np.float64)
    for j in range(h):
        for i in range(w):
            c = j * w + i
            if i + 1 < w or (wrap and w > 1):
                m = j * w + (i + 1) % w
                A[c, m] -= e[c]
                A[m, c] -= e[c]
            if j + 1 < h:
                A[c, c + w] -= s[c]
                A[c + w, c] -= s[c]
    return A


def _prolong(wf, hf, wc, hc):
    P = np.zeros((wf * hf, wc * hc))
    for j in range(hf):
        for i in range(wf):
            P[j * wf + i, (j // 2) * wc + i // 2] = 1.0
    return P


@pytest.mark.parametrize("w,h,wrap", [(17, 9, True), (10, 13, False), (4, 40, True),
                                      (3, 30, True)])
@pytest.mark.parametrize("scale", [1.0, 0.5])
def test_coarse_levels_are_the_scaled_galerkin_operator(w, h, wrap, scale):
    b = O.terrain(w, h, 2)
    kd = O.shelf_kd(b, 2)
    e, s, g = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    hy = O.Hierarchy(w, h, wrap, e, s, g, 0.8, 2, scale)
    assert hy.levels() >= 2
    for l in range(hy.levels() - 1):
        F, C = hy.level(l), hy.level(l + 1)
        assert (C["w"], C["h"]) == ((F["w"] + 1) // 2, (F["h"] + 1) // 2)
        Af, Ac = _level_matrix(F), _level_matrix(C)
        P = _prolong(F["w"], F["h"], C["w"], C["h"])
        gal = P.T @ Af @ P
        mass = np.diag(P.T @ np.diag(F["mass"].astype(np.float64)) @ P).copy()
        want = scale * (gal - np.diag(mass)) + np.diag(mass)
        assert np.allclose(Ac, want, rtol=0, atol=3e-6 * np.abs(want).max())
        assert np.allclose(C["mass"], mass, rtol=0)
        assert np.allclose(C["wd"], 0.8 / C["diag"], rtol=1e-7)
    last = hy.level(hy.levels() - 1)
    assert last["w"] * last["h"] <= 64


def _jacobi(A, r, om, zs, sweeps):
    D = np.diag(A)
    z = zs + 0.0
    for _ in range(sweeps):
        z = z + om * (r - A @ z) / D
    return z


def test_single_level_cycle_is_sixteen_damped_jacobi_sweeps():
    w, h = 8, 7
    b = O.terrain(w, h, 4)
    kd = O.shelf_kd(b, 4)
    e, s, g = O.operator(w, h, 2400.0, True, kd, 2e6)
    hy = O.Hierarchy(w, h, True, e, s, g)
    assert hy.levels() == 1
    A = _level_matrix(hy.level(0))
    r = np.random.default_rng(1).standard_normal(w * h)
    z, rz = hy.precondition(r.astype(np.float32), r)
    r32 = r.astype(np.float32).astype(np.float64)
    want = _jacobi(A, r32, 0.8, 0.8 * r32 / np.diag(A), 15)
    assert np.allclose(z, want, rtol=0, atol=1e-5 * np.abs(want).max())
    assert abs(rz - float(r @ z.astype(np.float64))) <= 1e-12 * abs(rz)


def test_two_level_cycle_in_matrix_form():
    w, h = 10, 12  # 120 cells: one coarse level of 5 x 6 = 30 cells
    b = O.terrain(w, h, 6)
    kd = O.shelf_kd(b, 6)
    e, s, g = O.operator(w, h, 2400.0, True, kd, 2e6)
    hy = O.Hierarchy(w, h, True, e, s, g)
    assert hy.levels() == 2
    A, Ac = _level_matrix(hy.level(0)), _level_matrix(hy.level(1))
    P = _prolong(w, h, 5, 6)
    r = np.random.default_rng(2).standard_normal(w * h)
    r32 = r.astype(np.float32).astype(np.float64)
    om = 0.8
    z = om * r32 / np.diag(A)
    z = _jacobi(A, r32, om, z, 1)
    rc = P.T @ (r32 - A @ z)
    zc = _jacobi(Ac, rc, om, om * rc / np.diag(Ac), 15)
    z = _jacobi(A, r32, om, z + P @ zc, 2)
    got, _ = hy.precondition(r.astype(np.float32), r)
    assert np.allclose(got, z, rtol=0, atol=2e-5 * np.abs(z).max())


# ---- the reference's own tests, restated -----------------------------------------------------

def test_the_multigrid_solve_meets_its_tolerance_and_keeps_the_mass():
    """multigrid/tests.rs:181-217."""
    solved = 0
    for (w, h) in SIZES:
        for wrap in (True, False):
            for seed in (1, 2):
                b = O.terrain(w, h, seed * 7 + w)
                kd = O.shelf_kd(b, seed)
                for dt, tol in [(2.0e6, 1.0e-6), (2.0e6, 1.0e-3), (2.0e4, 1.0e-6)]:
                    exact = O.reference_cg(w, h, 2400.0, wrap, b, kd, dt)
                    bx = O.lod(w, h, 2400.0, wrap, b, kd, dt)
                    x, its = O.cg_on(w, h, 2400.0, wrap, bx, b, kd, dt, tol)
                    assert its < 400
                    lim = O.bound(w, h, 2400.0, wrap, kd, dt, tol)
                    assert rms(x, exact) <= lim, (w, h, wrap, seed, dt, tol)
                    scale = max(np.abs(b).sum(), 1.0)
                    assert abs(b.sum() - x.sum()) <= 1e-12 * scale
                    x0, _ = O.cg_on(w, h, 2400.0, wrap, b, b, kd, dt, tol)
                    assert rms(x0, exact) <= lim
                    solved += 1
    assert solved >= 100


def test_the_multigrid_and_jacobi_solves_agree_within_their_tolerance():
    """multigrid/tests.rs:219-238."""
    for (w, h) in [(64, 40), (129, 67), (5, 3)]:
        for wrap in (True, False):
            b = O.terrain(w, h, 11 + w)
            kd = O.shelf_kd(b, 3)
            xm, _ = O.solve(w, h, 2400.0, wrap, b, kd, 2e6, 1e-6)
            xj, _ = O.jacobi_step(w, h, 2400.0, wrap, b, kd, 2e6)
            lim = 2.0 * O.bound(w, h, 2400.0, wrap, kd, 2e6, 1e-6)
            assert rms(xm, xj) <= lim


@pytest.mark.parametrize("w,h,wrap", [(64, 40, True), (37, 23, False), (5, 3, True),
                                      (2, 9, True), (1, 1, False)])
def test_the_cycle_is_a_symmetric_positive_definite_operator(w, h, wrap):
    """multigrid/tests.rs:242-289."""
    z = O.terrain(w, h, 5)
    kd = O.shelf_kd(z, 5)
    e, s, g = O.operator(w, h, 2400.0, wrap, kd, 2e6)
    for (om, nu, sc) in [(0.8, 2, 0.5), (0.7, 2, 1.0)]:
        hy = O.Hierarchy(w, h, wrap, e, s, g, om, nu, sc)
        rng = O.SplitMix64(9)
        r1 = np.array([rng.next_f64() - 0.5 for _ in range(w * h)])
        r2 = np.array([rng.next_f64() - 0.5 for _ in range(w * h)])
        z1, rz1 = hy.precondition(r1.astype(np.float32), r1)
        z2, rz2 = hy.precondition(r2.astype(np.float32), r2)
        z1, z2 = z1.astype(np.float64), z2.astype(np.float64)
        assert rz1 > 0 and rz2 > 0
        assert abs(rz1 - z1 @ r1) <= 1e-9 * max(abs(rz1), 1e-30)
        assert abs(z1 @ r2 - r1 @ z2) <= 1e-4 * np.sqrt(rz1 * rz2)


def test_the_cycle_takes_far_fewer_iterations_than_jacobi_on_a_shelf():
    """multigrid/tests.rs:291-301."""
    w, h = 256, 128
    b = O.terrain(w, h, 21)
    kd = O.shelf_kd(b, 21)
    _, mg = O.solve(w, h, 2400.0, True, b, kd, 2e6, 1e-6)
    _, jac = O.jacobi_step(w, h, 2400.0, True, b, kd, 2e6)
    assert 3 * mg <= jac


@pytest.mark.parametrize("solver", ["multigrid", "jacobi"])
def test_diffusion_conserves_mass_and_makes_no_new_extremes(solver):
    """surface/tests.rs:110-140 (stiff diffusivities 0, 1e-20, 5, 300, 1e5)."""
    rng = O.SplitMix64(7)
    for (w, h, wrap) in [(64, 32, True), (33, 17, False), (2, 5, True), (1, 9, True)]:
        n = w * h
        z = np.array([rng.next_f64() * 4000.0 - 2000.0 for _ in range(n)])
        kd = np.array([[0.0, 1.0e-20, 5.0, 300.0, 1.0e5][c % 5] for c in range(n)])
        before = z.sum()
        for _ in range(10):
            if solver == "multigrid":
                z, _ = O.solve(w, h, 4800.0, wrap, z, kd, 2e6, 1e-6)
            else:
                z, _ = O.jacobi_step(w, h, 4800.0, wrap, z, kd, 2e6)
        assert abs(before - z.sum()) <= 1e-9 * n * 2000.0
        assert np.isfinite(z).all() and (z > -2000.1).all() and (z < 2000.1).all()


def test_results_do_not_depend_on_the_thread_count():
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import oracle as O, numpy as np;"
            "b = O.terrain(300, 129, 3); kd = O.shelf_kd(b, 3);"
            "x, its = O.solve(300, 129, 2400.0, True, b, kd, 2e6, 1e-7);"
            "sys.stdout.write(x.tobytes().hex()[:4000] + str(its) + repr(float(x.sum())))"
            ) % os.path.dirname(O.__file__)
    outs = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.add(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                                text=True, check=True).stdout)
    assert len(outs) == 1


def test_the_config_fields():
    from dtdiff import fields as F
    d, h0, kd, dt = F.case_constant(w=128, h=64)
    assert abs(h0.mean()) < 1e-9 and abs(h0.std() - 1000.0) < 1e-9
    assert (kd == 1).all() and dt == 10.0
    # k^-2 power: the radially averaged spectrum falls by ~4x per doubling of |k|.
    f = np.abs(np.fft.fft2(h0.reshape(64, 128))) ** 2
    ky, kx = np.meshgrid(np.fft.fftfreq(64), np.fft.fftfreq(128), indexing="ij")
    k = np.sqrt(kx ** 2 + ky ** 2)
    lo = f[(k > 0.05) & (k < 0.07)].mean()
    hi = f[(k > 0.1) & (k < 0.14)].mean()
    assert 2.5 < lo / hi < 6.5
    d, h0, kd, dt = F.case_lognormal(w=256, h=128)
    assert abs(kd.mean() - 1.0) < 1e-12 and dt == pytest.approx(50.0)
    lk = np.log(kd)
    assert abs(lk.std() - 1.0) < 1e-9
