"""Brute-force checker of the ray-traced wall sampling.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

The reference traces a wall-normal segment through face-connected
tetrahedra and accepts a sample only when the whole segment from the wall
to it lies in the fluid (turbulence/sst_gpu_v2/source/
puretet_profile5_eq5_hxt5b/modules/wall/portable/tet_ray_trace.hpp:1-5,
TetRayTrace::trace :319-335).  Both properties are checked here without
any traversal:

* ``locate_all``: every cell whose barycentric coordinates at a point are
  all >= -tol (exhaustive search over the mesh);
* ``segment_cover``: the ray parameter interval of every cell (barycentric
  coordinates are affine along a line, so each tetrahedron holds one
  interval), and whether their union covers [0, t] without a gap, i.e.
  whether the segment stays inside the meshed fluid; the first gap is the
  distance at which it leaves.

Meshes for the pins: ``kuhn_box`` (cubes split into six tetrahedra around
the main diagonal, optionally with jittered interior vertices) and
``u_channel`` (the same box with a notch cut out of the top, so that a
segment can leave the fluid and re-enter it).
"""
from __future__ import annotations

import numpy as np

# six tetrahedra of the unit cube around the diagonal 0-7 (vertex index =
# x + 2y + 4z)
_KUHN = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]


def kuhn_box(nx, ny, nz, h=1.0, jitter=0.0, seed=0, keep=None):
    """points (np,3), tets (nt,4) of an nx x ny x nz box of cubes of side h;
    keep(i, j, k) -> bool drops cubes; interior vertices move by up to
    jitter*h in each coordinate.  Tetrahedra are positively oriented."""
    idx = -np.ones((nx + 1, ny + 1, nz + 1), dtype=np.int64)
    pts = []
    cubes = [(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx) if keep is None or keep(i, j, k)]
    used = set()
    for i, j, k in cubes:
        for c in range(8):
            used.add((i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)))
    for v in sorted(used, key=lambda t: (t[2], t[1], t[0])):
        idx[v] = len(pts)
        pts.append([v[0] * h, v[1] * h, v[2] * h])
    pts = np.array(pts, dtype=float)
    tets = []
    for i, j, k in cubes:
        corner = [idx[i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)] for c in range(8)]
        for t in _KUHN:
            tets.append([corner[a] for a in t])
    tets = np.array(tets, dtype=np.int64)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        boundary = _boundary_vertices(tets, len(pts))
        move = ~boundary
        pts[move] += jitter * h * rng.uniform(-1, 1, (int(move.sum()), 3))
    # orient positively
    d = _dets(pts, tets)
    flip = d < 0
    tets[flip, 0], tets[flip, 1] = tets[flip, 1].copy(), tets[flip, 0].copy()
    return pts, tets


def u_channel(n=3, h=1.0 / 3):
    """A 3n x 2n x 1 box of cubes with the top middle n x n block removed:
    two arms of width n joined by a base."""
    return kuhn_box(3 * n, 2 * n, 1, h, keep=lambda i, j, k: not (n <= i < 2 * n and j >= n))


def _dets(pts, tets):
    a = pts[tets[:, 1]] - pts[tets[:, 0]]
    b = pts[tets[:, 2]] - pts[tets[:, 0]]
    c = pts[tets[:, 3]] - pts[tets[:, 0]]
    return np.einsum("ij,ij->i", a, np.cross(b, c))


def boundary_faces(tets):
    """(cell, local face j opposite vertex j) of every face that belongs to
    one tetrahedron only."""
    keys = {}
    for c, t in enumerate(tets):
        for j in range(4):
            k = tuple(sorted(int(t[a]) for a in range(4) if a != j))
            keys.setdefault(k, []).append((c, j))
    return [v[0] for v in keys.values() if len(v) == 1]


def _boundary_vertices(tets, n):
    out = np.zeros(n, dtype=bool)
    for c, j in boundary_faces(tets):
        for a in range(4):
            if a != j:
                out[tets[c, a]] = True
    return out


def face_frame(pts, tets, c, j):
    """Centroid-free description of boundary face j of cell c: its three
    vertices (3,3) and its inward unit normal (towards vertex j)."""
    v = pts[[tets[c, a] for a in range(4) if a != j]]
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n /= np.linalg.norm(n)
    if np.dot(pts[tets[c, j]] - v[0], n) < 0:
        n = -n
    return v, n


def bary_all(pts, tets, p):
    """(nt,4) barycentric coordinates of point p in every cell."""
    o = pts[tets[:, 0]]
    M = np.stack([pts[tets[:, 1]] - o, pts[tets[:, 2]] - o, pts[tets[:, 3]] - o], axis=2)  # columns
    lam = np.linalg.solve(M, (p - o)[:, :, None])[:, :, 0]
    return np.concatenate([1 - lam.sum(axis=1, keepdims=True), lam], axis=1)


def locate_all(pts, tets, p, tol=1e-9):
    return np.flatnonzero(bary_all(pts, tets, p).min(axis=1) >= -tol)


def cell_intervals(pts, tets, origin, d, tol=1e-9):
    """(lo, hi) per cell: ray parameters (unit direction d) at which every
    barycentric coordinate is >= -tol."""
    d = np.asarray(d, float) / np.linalg.norm(d)
    l0 = bary_all(pts, tets, origin)
    l1 = bary_all(pts, tets, origin + d)
    s = l1 - l0
    lo = np.full(len(tets), -np.inf)
    hi = np.full(len(tets), np.inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        at = (-tol - l0) / s
    pos, neg = s > 1e-300, s < -1e-300
    lo = np.max(np.where(pos, at, -np.inf), axis=1)
    hi = np.min(np.where(neg, at, np.inf), axis=1)
    dead = np.any((np.abs(s) <= 1e-300) & (l0 < -tol), axis=1)
    lo[dead], hi[dead] = np.inf, -np.inf
    return lo, hi


def segment_cover(pts, tets, origin, d, t_end, tol=1e-9, gap=1e-9):
    """(inside, reach): whether the cell intervals cover [0, t_end] with no
    gap wider than `gap` (in length units), and the distance up to which
    they do."""
    lo, hi = cell_intervals(pts, tets, origin, d, tol)
    ok = (hi > lo + gap) & (hi > 0) & (lo < t_end)
    order = np.argsort(lo[ok])
    reach = 0.0
    for a, b in zip(lo[ok][order], hi[ok][order]):
        if a > reach + gap:
            break
        reach = max(reach, b)
        if reach >= t_end:
            return True, t_end
    return reach >= t_end, min(reach, t_end)
