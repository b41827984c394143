"""Structured tetrahedral mesh generators that write OpenFOAM polyMesh directories.

The H8 solver path reads an OpenFOAM ``constant/polyMesh`` directory of
triangular faces (reference reader ``gpu/serial_cuda/foam_mesh_g2.hpp:55``),
so every generated case is written in that format and read back through the
same reader as any external mesh.  Velocity DOF numbering follows face order
(vertices first, then faces; ``gpu/serial_cuda/g3_host.hpp:29-55``), which
makes the face order written here part of the case definition.

Generators:

* ``kuhn_cube(n)``: unit cube, ``n**3`` hexes, six Kuhn tetrahedra per hex
  sharing the main diagonal.
* ``ogrid_section(core, ring)``: O-grid disc of ``core**2 + 4*core*ring``
  quads (224 for ``core=8, ring=5``; 576 for ``core=12, ring=9``).
* ``swept_tets(...)``: sweep a quad section along a path, split every
  hex into two triangular prisms and every prism into three tetrahedra with
  the smallest-global-vertex diagonal rule, which keeps neighbouring faces
  conforming.  On a Kuhn-numbered cube this reproduces the Kuhn split.
"""
from __future__ import annotations

import math
import os
from pathlib import Path
from typing import Callable, List, Sequence, Tuple

import numpy as np

__all__ = [
    "kuhn_cube", "ogrid_section", "swept_tets", "straight_pipe", "helical_pipe",
    "tets_to_polymesh", "write_polymesh", "write_cavity_case", "write_pipe_case",
    "PolyMesh",
]


# ---------------------------------------------------------------------------
# volume meshes
# ---------------------------------------------------------------------------

def kuhn_cube(n: int, length: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Unit cube with ``n**3`` hexes, each split into six Kuhn tetrahedra.

    Vertex ``(i, j, k)`` has index ``i + (n+1)*(j + (n+1)*k)``.  The six
    tetrahedra of a hex are the monotone lattice paths from its corner
    ``v000`` to ``v111``.
    """
    m = n + 1
    g = np.arange(m, dtype=np.float64) * (length / n)
    k, j, i = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    ar = np.arange(n)
    kk, jj, ii = np.meshgrid(ar, ar, ar, indexing="ij")   # hex order: i fastest
    base = ii.ravel() + m * (jj.ravel() + m * kk.ravel())
    step = np.array([1, m, m * m])
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        a = base
        b = a + step[perm[0]]
        c = b + step[perm[1]]
        d = c + step[perm[2]]
        tets.append(np.stack([a, b, c, d], axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4)
    return pts, tets.astype(np.int64)


def ogrid_section(core: int = 8, ring: int = 5, radius: float = 0.5,
                  core_frac: float = 0.4) -> Tuple[np.ndarray, np.ndarray]:
    """O-grid disc: a ``core x core`` square of quads and ``ring`` layers.

    Returns 2-D points ``(np, 2)`` and quads ``(nq, 4)`` (counter-clockwise).
    """
    a = core_frac * radius
    pts: List[Tuple[float, float]] = []
    for j in range(core + 1):
        for i in range(core + 1):
            pts.append((a * (2.0 * i / core - 1.0), a * (2.0 * j / core - 1.0)))

    def core_index(i: int, j: int) -> int:
        return i + (core + 1) * j

    nper = 4 * core
    perim = []
    for m in range(nper):
        side, s = divmod(m, core)
        if side == 0:
            ij = (s, 0)
        elif side == 1:
            ij = (core, s)
        elif side == 2:
            ij = (core - s, core)
        else:
            ij = (0, core - s)
        perim.append(core_index(*ij))
    base = len(pts)
    for layer in range(1, ring + 1):
        t = layer / ring
        for m in range(nper):
            sx, sy = pts[perim[m]]
            th = math.atan2(sy, sx)
            cx, cy = radius * math.cos(th), radius * math.sin(th)
            pts.append(((1.0 - t) * sx + t * cx, (1.0 - t) * sy + t * cy))

    def ring_index(layer: int, m: int) -> int:
        m %= nper
        if layer == 0:
            return perim[m]
        return base + (layer - 1) * nper + m

    quads = []
    for j in range(core):
        for i in range(core):
            quads.append((core_index(i, j), core_index(i + 1, j),
                          core_index(i + 1, j + 1), core_index(i, j + 1)))
    for layer in range(ring):
        for m in range(nper):
            quads.append((ring_index(layer, m), ring_index(layer + 1, m),
                          ring_index(layer + 1, m + 1), ring_index(layer, m + 1)))
    return np.asarray(pts, dtype=np.float64), np.asarray(quads, dtype=np.int64)


def _prism_tets(a: int, b: int, c: int, nsec: int) -> List[Tuple[int, int, int, int]]:
    """Three tetrahedra of the prism (a,b,c)x(a+nsec,b+nsec,c+nsec).

    Every quadrilateral side face is split along the diagonal through its
    smallest global vertex index; the bottom vertices always have the smaller
    indices, so the diagonal starts at min(bottom pair).
    """
    tri = [a, b, c]
    r = tri.index(min(tri))
    a, b, c = tri[r:] + tri[:r]
    A, B, C = a + nsec, b + nsec, c + nsec
    out = [(a, B, C, A)]
    if b < c:
        out += [(a, b, c, C), (a, b, C, B)]
    else:
        out += [(a, b, c, B), (a, c, C, B)]
    return out


def swept_tets(pts2d: np.ndarray, quads: np.ndarray, nslab: int,
               place: Callable[[np.ndarray, np.ndarray], np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    """Sweep a quad section through ``nslab`` slabs; 6 tets per hex.

    ``place(xy, s)`` maps section coordinates ``xy (np,2)`` at path
    parameters ``s (nslab+1,)`` to 3-D points ``(nslab+1, np, 3)``.
    Vertex index = ``section_vertex + slab_plane * nsec``.
    """
    nsec = pts2d.shape[0]
    s = np.arange(nslab + 1, dtype=np.float64) / nslab
    xyz = place(pts2d, s).reshape(-1, 3)
    per_quad = []
    for q in quads:
        q = [int(v) for v in q]
        mn = q.index(min(q))
        q = q[mn:] + q[:mn]
        # diagonal through the smallest vertex q[0]
        per_quad.append(_prism_tets(q[0], q[1], q[2], nsec) + _prism_tets(q[0], q[2], q[3], nsec))
    per_quad = np.asarray(per_quad, dtype=np.int64)  # (nq, 6, 4)
    tets = []
    for k in range(nslab):
        tets.append(per_quad + k * nsec)
    tets = np.concatenate(tets, axis=0).reshape(-1, 4)
    return xyz, tets


def straight_pipe(core: int = 8, ring: int = 5, nslab: int = 30, diameter: float = 1.0,
                  length_d: float = 10.0) -> Tuple[np.ndarray, np.ndarray]:
    """Straight circular pipe along +z from z=0 to z=length_d*diameter."""
    pts2d, quads = ogrid_section(core, ring, 0.5 * diameter)
    L = length_d * diameter

    def place(xy, s):
        out = np.empty((s.size, xy.shape[0], 3))
        out[:, :, 0] = xy[None, :, 0]
        out[:, :, 1] = xy[None, :, 1]
        out[:, :, 2] = s[:, None] * L
        return out

    return swept_tets(pts2d, quads, nslab, place)


def helical_pipe(core: int = 12, ring: int = 9, nslab: int = 60, diameter: float = 1.0,
                 coil_radius_d: float = 5.0, pitch_d: float = 2.0, turns: float = 2.0) -> Tuple[np.ndarray, np.ndarray]:
    """Helical pipe: section swept along a helix with a Frenet-like frame."""
    pts2d, quads = ogrid_section(core, ring, 0.5 * diameter)
    Rc = coil_radius_d * diameter
    c = pitch_d * diameter / (2.0 * math.pi)
    T = 2.0 * math.pi * turns

    def place(xy, s):
        t = s * T
        cen = np.stack([Rc * np.cos(t), Rc * np.sin(t), c * t], axis=1)
        tan = np.stack([-Rc * np.sin(t), Rc * np.cos(t), np.full_like(t, c)], axis=1)
        tan /= np.linalg.norm(tan, axis=1, keepdims=True)
        nor = np.stack([-np.cos(t), -np.sin(t), np.zeros_like(t)], axis=1)
        bin_ = np.cross(tan, nor)
        out = (cen[:, None, :] + xy[None, :, 0:1] * nor[:, None, :]
               + xy[None, :, 1:2] * bin_[:, None, :])
        return out

    return swept_tets(pts2d, quads, nslab, place)


# ---------------------------------------------------------------------------
# polyMesh construction and output
# ---------------------------------------------------------------------------

class PolyMesh:
    """Triangle-face polyMesh arrays in OpenFOAM order."""

    def __init__(self, points, faces, owner, neighbour, patches):
        self.points = points          # (np,3) float64
        self.faces = faces            # (nf,3) int64
        self.owner = owner            # (nf,)
        self.neighbour = neighbour    # (ni,)
        self.patches = patches        # list of (name, start, count)

    @property
    def n_cells(self) -> int:
        return int(max(self.owner.max(), self.neighbour.max() if self.neighbour.size else 0)) + 1


def tets_to_polymesh(points: np.ndarray, tets: np.ndarray,
                     classify: Callable[[np.ndarray, np.ndarray], np.ndarray],
                     patch_names: Sequence[str]) -> PolyMesh:
    """Build OpenFOAM face lists from tetrahedra.

    Internal faces are sorted by (owner, neighbour) with owner < neighbour,
    boundary faces are grouped by patch in ``patch_names`` order and sorted
    by owner.  Face normals point out of the owner cell.  ``classify`` maps
    boundary face centroids ``(nb,3)`` and unit normals ``(nb,3)`` to patch
    indices.
    """
    nc = tets.shape[0]
    loc = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    allf = tets[:, loc].reshape(-1, 3)
    cell = np.repeat(np.arange(nc), 4)
    key = np.sort(allf, axis=1)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    # pair detection
    first = np.ones(ks.shape[0], dtype=bool)
    first[1:] = ~same
    gid = np.cumsum(first) - 1
    nface = gid[-1] + 1
    counts = np.bincount(gid, minlength=nface)
    if counts.max() > 2:
        raise ValueError("non-manifold face in tetrahedral mesh")
    fcell = cell[order]
    fverts = allf[order]
    start = np.flatnonzero(first)
    c0 = fcell[start]
    internal = counts == 2
    c1 = np.where(internal, fcell[np.minimum(start + 1, ks.shape[0] - 1)], -1)
    own = np.where(internal, np.minimum(c0, c1), c0)
    nei = np.where(internal, np.maximum(c0, c1), -1)
    verts = fverts[start]
    centroids = points[tets].mean(axis=1)
    fc = points[verts].mean(axis=1)
    nrm = np.cross(points[verts[:, 1]] - points[verts[:, 0]], points[verts[:, 2]] - points[verts[:, 0]])
    flip = np.einsum("ij,ij->i", nrm, fc - centroids[own]) < 0.0
    verts[flip] = verts[flip][:, [0, 2, 1]]
    nrm[flip] = -nrm[flip]
    # internal faces (owner, neighbour) order
    ii = np.flatnonzero(internal)
    ii = ii[np.lexsort((nei[ii], own[ii]))]
    bi = np.flatnonzero(~internal)
    unit = nrm[bi] / np.linalg.norm(nrm[bi], axis=1, keepdims=True)
    patch = np.asarray(classify(fc[bi], unit), dtype=np.int64)
    if patch.min() < 0 or patch.max() >= len(patch_names):
        raise ValueError("boundary face not classified")
    bi = bi[np.lexsort((own[bi], patch))]
    patch_sorted = np.sort(patch, kind="stable")
    faces = np.concatenate([verts[ii], verts[bi]], axis=0)
    owner = np.concatenate([own[ii], own[bi]])
    neighbour = nei[ii]
    patches = []
    start_face = ii.size
    for p, name in enumerate(patch_names):
        cnt = int(np.count_nonzero(patch_sorted == p))
        patches.append((name, start_face, cnt))
        start_face += cnt
    return PolyMesh(points, faces, owner, neighbour, patches)


_HEADER = """/*--------------------------------*- C++ -*----------------------------------*\\
| polyMesh written by nodals_amd.meshgen                                      |
\\*---------------------------------------------------------------------------*/
FoamFile
{{
    format      ascii;
    class       {cls};
    location    "constant/polyMesh";
    object      {obj};
}}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

"""


def _write_list(path: Path, cls: str, obj: str, lines: Sequence[str]) -> None:
    with open(path, "w", encoding="ascii") as f:
        f.write(_HEADER.format(cls=cls, obj=obj))
        f.write(f"\n{len(lines)}\n(\n")
        f.write("\n".join(lines))
        f.write("\n)\n\n// ************************************************************************* //\n")


def write_polymesh(directory: os.PathLike, mesh: PolyMesh) -> Path:
    """Write ``points faces owner neighbour boundary`` under ``directory``."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    p = mesh.points
    _write_list(d / "points", "vectorField", "points",
                ["(%.17g %.17g %.17g)" % (x, y, z) for x, y, z in p])
    _write_list(d / "faces", "faceList", "faces",
                ["3(%d %d %d)" % (a, b, c) for a, b, c in mesh.faces])
    _write_list(d / "owner", "labelList", "owner", [str(int(v)) for v in mesh.owner])
    _write_list(d / "neighbour", "labelList", "neighbour", [str(int(v)) for v in mesh.neighbour])
    blines = []
    for name, start, cnt in mesh.patches:
        blines.append(f"    {name}\n    {{\n        type            patch;\n"
                      f"        nFaces          {cnt};\n        startFace       {start};\n    }}")
    with open(d / "boundary", "w", encoding="ascii") as f:
        f.write(_HEADER.format(cls="polyBoundaryMesh", obj="boundary"))
        f.write(f"\n{len(blines)}\n(\n" + "\n".join(blines) + "\n)\n")
    return d


def write_pipe_case(directory: os.PathLike, core: int = 8, ring: int = 5, nslab: int = 30,
                    diameter: float = 1.0, length_d: float = 10.0) -> Path:
    """Straight pipe polyMesh with the H8 patch names.

    ``patch_0_0`` wall, ``patch_2_0`` inlet (z=0), ``patch_1_0`` outlet
    (z=L), the defaults of ``gpu/serial_cuda/h8_main_base.cu:90``.
    """
    pts, tets = straight_pipe(core, ring, nslab, diameter, length_d)
    L = length_d * diameter
    tol = 1e-9 * max(1.0, L)

    def classify(fc, n):
        out = np.zeros(fc.shape[0], dtype=np.int64)
        out[np.abs(fc[:, 2]) < tol] = 1
        out[np.abs(fc[:, 2] - L) < tol] = 2
        return out

    mesh = tets_to_polymesh(pts, tets, classify, ["patch_0_0", "patch_2_0", "patch_1_0"])
    return write_polymesh(directory, mesh)


def write_helix_case(directory: os.PathLike, core: int = 12, ring: int = 9, nslab: int = 60,
                     diameter: float = 1.0, coil_radius_d: float = 5.0, pitch_d: float = 2.0,
                     turns: float = 2.0) -> Path:
    """Helical pipe polyMesh with the H8 patch names (run it with --duct 1).

    ``patch_0_0`` wall, ``patch_2_0`` inlet (the section at the start of the
    helix), ``patch_1_0`` outlet (the section at its end).  The defaults are
    the spiral configuration's section (576 quads: a 12x12 core and 4 x
    (12 x 9) ring) swept through 60 slabs, 6 tets per hex: 207,360 tets.
    """
    pts, tets = helical_pipe(core, ring, nslab, diameter, coil_radius_d, pitch_d, turns)
    Rc = coil_radius_d * diameter
    c = pitch_d * diameter / (2.0 * math.pi)
    T = 2.0 * math.pi * turns
    tol = 1e-8 * diameter

    def plane(t):
        cen = np.array([Rc * math.cos(t), Rc * math.sin(t), c * t])
        tan = np.array([-Rc * math.sin(t), Rc * math.cos(t), c])
        return cen, tan / np.linalg.norm(tan)

    c0, t0 = plane(0.0)
    c1, t1 = plane(T)

    def classify(fc, n):
        out = np.zeros(fc.shape[0], dtype=np.int64)
        out[np.abs((fc - c0) @ t0) < tol] = 1
        out[np.abs((fc - c1) @ t1) < tol] = 2
        return out

    mesh = tets_to_polymesh(pts, tets, classify, ["patch_0_0", "patch_2_0", "patch_1_0"])
    return write_polymesh(directory, mesh)


def write_helix_msh2(path: os.PathLike, core: int = 12, ring: int = 9, nslab: int = 60, diameter: float = 1.0,
                     coil_radius_d: float = 5.0, pitch_d: float = 2.0, turns: float = 2.0) -> Path:
    """The helix of write_helix_case as a Gmsh ASCII 2.2 file with the SST
    application's physical tags (inlet 1, outlet 2, wall 3, volume 4)."""
    from nodals_amd import hxt1

    tmp = Path(path).with_suffix(".polymesh_tmp")
    write_helix_case(tmp, core, ring, nslab, diameter, coil_radius_d, pitch_d, turns)
    pts, tets = helical_pipe(core, ring, nslab, diameter, coil_radius_d, pitch_d, turns)
    faces = np.asarray(_read_foam_faces(tmp / "faces"))
    patches = _read_foam_boundary(tmp / "boundary")
    groups = {}
    for name, start, count in patches:
        groups[name] = faces[start:start + count]
    hxt1.write_msh2(path, pts, tets, groups["patch_2_0"], groups["patch_1_0"], groups["patch_0_0"])
    import shutil
    shutil.rmtree(tmp)
    return Path(path)


def _read_foam_faces(path: Path):
    out = []
    for line in Path(path).read_text().splitlines():
        line = line.strip()
        if line.startswith("3(") and line.endswith(")"):
            out.append([int(v) for v in line[2:-1].split()])
    return out


def _read_foam_boundary(path: Path):
    text = Path(path).read_text()
    out = []
    import re
    for m in re.finditer(r"(\w+)\s*\{[^}]*?nFaces\s+(\d+);[^}]*?startFace\s+(\d+);", text):
        out.append((m.group(1), int(m.group(3)), int(m.group(2))))
    return out


def write_cavity_case(directory: os.PathLike, n: int = 16) -> Path:
    """Lid-driven cavity polyMesh: ``lid`` at z=1 and ``walls`` elsewhere."""
    pts, tets = kuhn_cube(n)

    def classify(fc, nrm):
        out = np.zeros(fc.shape[0], dtype=np.int64)
        out[np.abs(fc[:, 2] - 1.0) < 1e-12] = 1
        return out

    mesh = tets_to_polymesh(pts, tets, classify, ["walls", "lid"])
    return write_polymesh(directory, mesh)


if __name__ == "__main__":  # pragma: no cover - convenience CLI
    import argparse

    ap = argparse.ArgumentParser(description="write a structured polyMesh")
    ap.add_argument("kind", choices=["pipe", "cavity"])
    ap.add_argument("out")
    ap.add_argument("--core", type=int, default=8)
    ap.add_argument("--ring", type=int, default=5)
    ap.add_argument("--slabs", type=int, default=30)
    ap.add_argument("--n", type=int, default=16)
    a = ap.parse_args()
    if a.kind == "pipe":
        write_pipe_case(a.out, a.core, a.ring, a.slabs)
    else:
        write_cavity_case(a.out, a.n)
