"""Synthetic cube grids for the cube-grid diffusion solve: the inputs `CubeGrid` hands the solve.

The solve (crates/deeptime-core/src/surface/cube/diffusion.rs) reads three things of the grid:
each cell's eight neighbour slots in D8 order (`CubeTopology::slots`, E, NE, N, NW, W, SW, S, SE
in the cell's face frame, none for the diagonal across a cube corner), the Laplacian's couplings
(`CubeGrid::lap`) and the cells' areas with their mean.  This module builds them for an n x n
cube on a sphere of radius R, with the metric rules of crates/deeptime-core/src/sphere/cube.rs:
- storage: face by face, row by row, cell (face, i, j) at face * n^2 + j * n + i, i east and j
  south in the face's frame (cube.rs:6-9);
- areas: R^2 times the solid angle of the cell's corner quadrilateral as two geodesic triangles
  (Van Oosterom & Strackee 1983, cube.rs:180-187, 248-252);
- couplings: P1 finite elements on the cell centres, each quadrilateral of four centres cut along
  both diagonals and the cuts averaged, the triangle of three centres at a cube corner taken whole
  (cotangent weights, cube.rs:279-307); every pair takes the value its lower-indexed cell computed
  (cube.rs:310-325).
The map from a face's local coordinates to the sphere is this module's own, the equiangular
gnomonic map p = normalize(n_f + tan(pi a / 4) e_f + tan(pi b / 4) s_f); the reference uses
Moser's equal-area map, which the solve does not see.  Neighbours across a face's edge are found
geometrically: a step across an edge lands on the next face's cell at the same place along the
edge (the map's coordinate along an edge is the same from both faces).  The tests check the
topology (mutual neighbours, seven neighbours at the 24 corner cells, eight elsewhere) and the
metric (areas summing to 4 pi R^2, symmetric couplings).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NO_CELL = 0xFFFFFFFF  # sphere/cube.rs:65
# D8 offsets (surface/grid.rs:21-31): E, NE, N, NW, W, SW, S, SE; north is j - 1.
DI = np.array([1, 1, 0, -1, -1, -1, 0, 1])
DJ = np.array([0, -1, -1, -1, 0, 1, 1, 1])

# Each face's centre normal, east (i) and south (j) axes, right-handed as n = e x s.
_FACES = [
    ((1, 0, 0), (0, 1, 0), (0, 0, -1)),
    ((0, 1, 0), (-1, 0, 0), (0, 0, -1)),
    ((-1, 0, 0), (0, -1, 0), (0, 0, -1)),
    ((0, -1, 0), (1, 0, 0), (0, 0, -1)),
    ((0, 0, 1), (0, 1, 0), (1, 0, 0)),
    ((0, 0, -1), (0, 1, 0), (-1, 0, 0)),
]


@dataclass
class CubeGrid:
    n: int
    radius: float
    slots: np.ndarray  # (6 n^2, 8) uint32, NO_CELL where there is no neighbour
    lap: np.ndarray  # (6 n^2, 8) f64 couplings (0 where there is no neighbour)
    area: np.ndarray  # (6 n^2,) m^2
    mean_area: float
    unit: np.ndarray  # (6 n^2, 3) cell centres on the unit sphere

    @property
    def cells(self) -> int:
        return 6 * self.n * self.n


def _map(face, a, b):
    nf, ef, sf = (np.array(v, dtype=np.float64) for v in _FACES[face])
    ta, tb = np.tan(np.pi * a / 4.0), np.tan(np.pi * b / 4.0)
    p = nf[None, :] + ta[:, None] * ef[None, :] + tb[:, None] * sf[None, :]
    return p / np.linalg.norm(p, axis=1, keepdims=True)


def _solid_angle(a, b, c):
    num = np.abs(np.einsum("ij,ij->i", a, np.cross(b, c)))
    den = 1.0 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + \
        np.einsum("ij,ij->i", c, a)
    return 2.0 * np.arctan2(num, den)


def _cot_at(pi, pj, pk):
    e1, e2 = pi - pk, pj - pk
    return np.einsum("ij,ij->i", e1, e2) / np.linalg.norm(np.cross(e1, e2), axis=1)


def cube_grid(n: int, radius: float = 6.371e6) -> CubeGrid:
    """The n x n cube grid (n >= 2) on a sphere of `radius` metres."""
    if n < 2:
        raise ValueError("a cube face needs at least 2 x 2 cells")
    nn, cells = n * n, 6 * n * n
    i = np.tile(np.arange(n), n)
    j = np.repeat(np.arange(n), n)
    unit = np.empty((cells, 3))
    for f in range(6):
        unit[f * nn:(f + 1) * nn] = _map(f, (2 * i + 1 - n) / n, (2 * j + 1 - n) / n)
    normals = np.array([v[0] for v in _FACES], dtype=np.float64)
    slots = np.full((cells, 8), NO_CELL, dtype=np.uint32)
    for f in range(6):
        base = f * nn
        for k in range(8):
            ii, jj = i + DI[k], j + DJ[k]
            in_i, in_j = (ii >= 0) & (ii < n), (jj >= 0) & (jj < n)
            inside = in_i & in_j
            slots[base + np.flatnonzero(inside), k] = base + (jj * n + ii)[inside]
            # One side out: the neighbour is the next face's cell at the same place along the
            # shared edge (the map's coordinate along an edge is the same from both faces), in
            # its first row or column.
            edge = in_i ^ in_j
            if not edge.any():
                continue
            e_i = np.where(ii[edge] < 0, -1.0, 1.0)
            e_j = np.where(jj[edge] < 0, -1.0, 1.0)
            a = np.where(in_i[edge], (2 * ii[edge] + 1 - n) / n, e_i)
            b = np.where(in_j[edge], (2 * jj[edge] + 1 - n) / n, e_j)
            beyond = _map(f, np.where(in_i[edge], a, a * (1 + 1.0 / n)),
                          np.where(in_j[edge], b, b * (1 + 1.0 / n)))
            g = np.argmax(beyond @ normals.T, axis=1)
            p = _map(f, a, b)
            m = np.empty(p.shape[0], dtype=np.int64)
            for gf in np.unique(g):
                sel = g == gf
                ng, eg, sg = (np.array(v, dtype=np.float64) for v in _FACES[gf])
                pn = p[sel] @ ng
                ag = np.arctan((p[sel] @ eg) / pn) * 4.0 / np.pi
                bg = np.arctan((p[sel] @ sg) / pn) * 4.0 / np.pi
                on_a = np.abs(np.abs(ag) - 1.0) < np.abs(np.abs(bg) - 1.0)
                gi = np.where(on_a, np.where(ag < 0, 0, n - 1),
                              np.rint(((ag + 1.0) * n - 1.0) / 2.0))
                gj = np.where(on_a, np.rint(((bg + 1.0) * n - 1.0) / 2.0),
                              np.where(bg < 0, 0, n - 1))
                m[sel] = gf * nn + gj.astype(np.int64) * n + gi.astype(np.int64)
            slots[base + np.flatnonzero(edge), k] = m.astype(np.uint32)
    # Corners of each face, (n + 1)^2 points; areas from the two triangles of each cell.
    q = np.arange(n + 1)
    qa, qb = np.tile(q, n + 1), np.repeat(q, n + 1)
    area = np.empty(cells)
    r2 = radius * radius
    for f in range(6):
        corner = _map(f, (2 * qa - n) / n, (2 * qb - n) / n).reshape(n + 1, n + 1, 3)
        q00 = corner[j, i]
        q10 = corner[j, i + 1]
        q01 = corner[j + 1, i]
        q11 = corner[j + 1, i + 1]
        area[f * nn:(f + 1) * nn] = r2 * (_solid_angle(q00, q10, q11) +
                                          _solid_angle(q00, q11, q01))
    # Couplings computed from each cell, then every pair given its lower-indexed cell's value.
    has = slots != NO_CELL
    nb = np.where(has, slots, 0).astype(np.int64)
    raw = np.zeros((cells, 8))
    pc = unit
    err = np.seterr(invalid="ignore", divide="ignore")  # the quad terms of corner cells, unused
    for k in (1, 3, 5, 7):
        ka, kb = k - 1, (k + 1) % 8
        pa, pb, pm = unit[nb[:, ka]], unit[nb[:, kb]], unit[nb[:, k]]
        corner = ~has[:, k]
        tri_a, tri_b = 0.5 * _cot_at(pc, pa, pb), 0.5 * _cot_at(pc, pb, pa)
        qa_ = 0.25 * (_cot_at(pc, pa, pm) + _cot_at(pc, pa, pb))
        qb_ = 0.25 * (_cot_at(pc, pb, pm) + _cot_at(pc, pb, pa))
        qm_ = 0.25 * (_cot_at(pc, pm, pa) + _cot_at(pc, pm, pb))
        raw[:, ka] += np.where(corner, tri_a, qa_)
        raw[:, kb] += np.where(corner, tri_b, qb_)
        raw[:, k] = np.where(corner, 0.0, qm_)
    np.seterr(**err)
    lap = raw.copy()
    c_idx = np.arange(cells)
    for k in range(8):
        m = nb[:, k]
        take = has[:, k] & (m < c_idx)
        if take.any():
            back = np.argmax(slots[m[take]] == c_idx[take, None].astype(np.uint32), axis=1)
            lap[take, k] = raw[m[take], back]
    lap[~has] = 0.0
    total = float(sum(area[f * nn:(f + 1) * nn].reshape(n, n).sum(axis=1).sum()
                      for f in range(6)))
    return CubeGrid(n, radius, slots, lap, area, total / cells, unit)


def cone(g: CubeGrid, centre, r0: float, hgt: float) -> np.ndarray:
    """A cone of height hgt and angular radius r0 around the unit vector `centre`."""
    c = np.asarray(centre, dtype=np.float64)
    c = c / np.linalg.norm(c)
    ang = np.arccos(np.clip(g.unit @ c, -1.0, 1.0))
    return np.maximum(0.0, hgt * (1.0 - ang / r0))
