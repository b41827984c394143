"""Gmsh ASCII 2.2 tetrahedral meshes to the HXT1 binary mesh of the SST
application, and HXT1 reading.

Input contract (turbulence/sst_gpu_v2/tools/msh2_to_hxt1.py:11-45): ASCII
MSH 2.2; physical tags inlet = 1, outlet = 2, wall = 3 on triangles and 4 on
tetrahedra; line and point elements are ignored, any other element type is
an error.  The conversion (:47-93) orients every tetrahedron positively
(vertices 2 and 3 swapped for a negative Jacobian), rejects non-manifold
faces, requires the tagged triangles to be exactly the exterior faces and
every node to be used, and writes ``HXT1_mesh.bin`` and ``mesh_audit.json``.

HXT1 layout (source/hybrid_hxt1/hxt1_mesh.hpp:20-131, produced by
tools/prepare_hxt1.py): magic ``HXT1BIN1``, uint32 version 1, uint32
reserved, thirteen uint64 counts, then the arrays in count order.  For a
tetrahedral mesh: points [nv][3] f64; tets [nt][4] i32; the velocity DOFs
[nt][8] i32 (the four vertices, then nv + the rank of each face among all
distinct faces sorted by their vertex triple, local face lf opposite vertex
lf with vertex triples (1,2,3), (0,3,2), (0,1,3), (0,2,1)); pressure ids
0..nt-1; geometry [nt][12] f64 (volume, longest edge squared, det J and the
row-major inverse of J = [p1-p0 | p2-p0 | p3-p0]); the inlet and outlet
boundary records (cell, local face, bubble DOF) [][3] i32.
"""
from __future__ import annotations

import collections
import hashlib
import json
import math
import struct
from pathlib import Path

import numpy as np

MAGIC = b"HXT1BIN1"
TET_FACES = ((1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1))
TAG_INLET, TAG_OUTLET, TAG_WALL, TAG_VOLUME = 1, 2, 3, 4


class MeshError(RuntimeError):
    pass


def read_msh2(path: Path):
    """(points (n,3), tets (nt,4) local indices, {tag: triangles (m,3)},
    counts of (element type, physical tag))."""
    lines = Path(path).read_text(encoding="utf8").splitlines()
    pos = 0

    def block(mark):
        nonlocal pos
        while pos < len(lines):
            pos += 1
            if lines[pos - 1].strip() == mark:
                return
        raise MeshError("Missing Gmsh block " + mark)

    def nextline():
        nonlocal pos
        if pos >= len(lines):
            raise MeshError("Truncated Gmsh file")
        pos += 1
        return lines[pos - 1]

    block("$MeshFormat")
    fmt = nextline().split()
    if len(fmt) < 2 or fmt[0] != "2.2" or fmt[1] != "0":
        raise MeshError("Expected ASCII MSH 2.2 (gmsh -format msh2 -bin 0)")
    block("$Nodes")
    n = int(nextline())
    ids = np.empty(n, dtype=np.int64)
    pts = np.empty((n, 3))
    for k in range(n):
        rec = nextline().split()
        if len(rec) != 4:
            raise MeshError("Bad Gmsh node record")
        ids[k] = int(rec[0])
        pts[k] = [float(x) for x in rec[1:]]
    if not np.isfinite(pts).all() or len(set(ids.tolist())) != n:
        raise MeshError("Non-finite node coordinates or repeated node ids")
    local = {int(g): i for i, g in enumerate(ids)}
    block("$Elements")
    ne = int(nextline())
    tets, tris = [], {TAG_INLET: [], TAG_OUTLET: [], TAG_WALL: []}
    counts = collections.Counter()
    for _ in range(ne):
        rec = nextline().split()
        if len(rec) < 4:
            raise MeshError("Bad Gmsh element record")
        etype, ntags = int(rec[1]), int(rec[2])
        if ntags < 1:
            raise MeshError("Element without a physical tag")
        tag = int(rec[3])
        nodes = rec[3 + ntags:]
        counts[(etype, tag)] += 1
        if etype == 4:
            if tag != TAG_VOLUME or len(nodes) != 4:
                raise MeshError(f"Tetrahedron with physical tag {tag} or {len(nodes)} nodes")
            tets.append([local[int(v)] for v in nodes])
        elif etype == 2:
            if tag not in tris or len(nodes) != 3:
                raise MeshError(f"Triangle with physical tag {tag} or {len(nodes)} nodes")
            tris[tag].append([local[int(v)] for v in nodes])
        elif etype not in (1, 15):
            raise MeshError(f"Element type {etype}: a pure tetrahedral mesh is required")
    tets = np.asarray(tets, dtype=np.int32).reshape(-1, 4)
    if len(tets) == 0 or any(len(v) == 0 for v in tris.values()):
        raise MeshError("Missing volume, inlet, outlet or wall elements")
    return pts, tets, {k: np.asarray(v, dtype=np.int32).reshape(-1, 3) for k, v in tris.items()}, counts


def _jacobians(pts, tets):
    p = pts[tets]
    return np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)


def hxt1_arrays(pts, tets, inlet, outlet):
    """HXT1 header counts and arrays for an oriented tetrahedral mesh."""
    nv, nt = len(pts), len(tets)
    keys = {}
    for c in range(nt):
        for lf, loc in enumerate(TET_FACES):
            keys.setdefault(tuple(sorted(int(tets[c, j]) for j in loc)), []).append((c, lf))
    order = sorted(keys)
    gid = {k: nv + i for i, k in enumerate(order)}
    tvel = np.empty((nt, 8), dtype=np.int32)
    tvel[:, :4] = tets
    for c in range(nt):
        for lf, loc in enumerate(TET_FACES):
            tvel[c, 4 + lf] = gid[tuple(sorted(int(tets[c, j]) for j in loc))]

    def records(tris):
        out = []
        for t in tris:
            k = tuple(sorted(int(v) for v in t))
            own = keys.get(k, [])
            if len(own) != 1:
                raise MeshError(f"boundary face {k} has {len(own)} cells")
            c, lf = own[0]
            out.append((c, lf, gid[k]))
        return np.asarray(out, dtype=np.int32).reshape(-1, 3)

    J = _jacobians(pts, tets)
    det = np.linalg.det(J)
    inv = np.linalg.inv(J)
    p = pts[tets]
    h2 = np.max([((p[:, a] - p[:, b]) ** 2).sum(1) for a in range(4) for b in range(a + 1, 4)], axis=0)
    geom = np.concatenate([(det / 6.0)[:, None], h2[:, None], det[:, None], inv.reshape(nt, 9)], axis=1)
    rin, rout = records(inlet), records(outlet)
    header = [nv, 0, nt, 0, len(order), nv + len(order), nt, 0, 0, 0, len(rin), 0, len(rout)]
    return header, dict(points=pts, tets=tets, tvel=tvel, tpres=np.arange(nt, dtype=np.int32), geom=geom,
                        inlet=rin, outlet=rout)


def write_hxt1(path: Path, header, arrays) -> None:
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<II", 1, 0))
        f.write(struct.pack("<13Q", *[int(x) for x in header]))
        f.write(np.ascontiguousarray(arrays["points"], dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(arrays["tets"], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(arrays["tvel"], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(arrays["tpres"], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(arrays["geom"], dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(arrays["inlet"], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(arrays["outlet"], dtype="<i4").tobytes())


def read_hxt1(path: Path) -> dict:
    """Tetrahedral HXT1 file as a dict of numpy arrays and the counts."""
    data = Path(path).read_bytes()
    if data[:8] != MAGIC:
        raise MeshError("bad HXT1 magic")
    version, _ = struct.unpack_from("<II", data, 8)
    if version != 1:
        raise MeshError("unsupported HXT1 version")
    h = struct.unpack_from("<13Q", data, 16)
    nv, nhex, nt, _, ntf, nvel, npres, ninter, nwh, ninh, nint, nouth, noutt = h
    if nhex or ninter or nwh or ninh or nouth:
        raise MeshError("hexahedral HXT1 content is not supported")
    off = 16 + 13 * 8

    def take(dtype, count, shape):
        nonlocal off
        a = np.frombuffer(data, dtype=dtype, count=count, offset=off).reshape(shape)
        off += a.nbytes
        return a

    out = dict(counts=dict(zip(["nv", "nhex", "ntet", "nhexface", "ntetface", "nvel", "npressure", "ninterface",
                                "nwallhex", "ninhex", "nintet", "nouthex", "nouttet"], h)))
    out["points"] = take("<f8", nv * 3, (nv, 3))
    out["tets"] = take("<i4", nt * 4, (nt, 4))
    out["tvel"] = take("<i4", nt * 8, (nt, 8))
    out["tpres"] = take("<i4", nt, (nt,))
    out["geom"] = take("<f8", nt * 12, (nt, 12))
    out["inlet"] = take("<i4", nint * 3, (nint, 3))
    out["outlet"] = take("<i4", noutt * 3, (noutt, 3))
    if off != len(data):
        raise MeshError("trailing bytes in HXT1 file")
    return out


def convert(msh: Path, outdir: Path) -> dict:
    """msh2 -> outdir/HXT1_mesh.bin + outdir/mesh_audit.json; returns the audit."""
    pts, tets, tris, counts = read_msh2(msh)
    det = np.linalg.det(_jacobians(pts, tets))
    if not np.isfinite(det).all() or (np.abs(det) < 1e-18).any():
        raise MeshError("Degenerate tetrahedra present")
    neg = det < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    use = collections.Counter(tuple(sorted(int(v) for v in tets[c, list(loc)]))
                              for c in range(len(tets)) for loc in TET_FACES)
    if any(n > 2 for n in use.values()):
        raise MeshError("Non-manifold face (more than two cells)")
    exterior = {k for k, n in use.items() if n == 1}
    tagged = collections.Counter(tuple(sorted(int(v) for v in t)) for a in tris.values() for t in a)
    if any(n != 1 for n in tagged.values()):
        raise MeshError("Boundary triangle tagged more than once")
    if exterior != set(tagged):
        raise MeshError(f"Tagged triangles are not the exterior faces: missing={len(exterior - set(tagged))}, "
                        f"extra={len(set(tagged) - exterior)}")
    used = set(tets.ravel().tolist())
    for a in tris.values():
        used.update(a.ravel().tolist())
    if used != set(range(len(pts))):
        raise MeshError("Unused nodes in the mesh")

    def area(t):
        return float(0.5 * np.linalg.norm(np.cross(pts[t[:, 1]] - pts[t[:, 0]], pts[t[:, 2]] - pts[t[:, 0]]),
                                          axis=1).sum())

    ai, ao = area(tris[TAG_INLET]), area(tris[TAG_OUTLET])
    if not (math.isfinite(ai) and math.isfinite(ao) and ai > 0 and ao > 0):
        raise MeshError("Invalid inlet/outlet area")
    outdir = Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    print(f"NODALS_MSH_PREFLIGHT=PASS tets={len(tets)} points={len(pts)} inletFaces={len(tris[TAG_INLET])} "
          f"outletFaces={len(tris[TAG_OUTLET])} wallFaces={len(tris[TAG_WALL])} "
          f"minTetVolume={float(np.abs(det).min() / 6.0)}", flush=True)
    header, arrays = hxt1_arrays(pts, tets, tris[TAG_INLET], tris[TAG_OUTLET])
    target = outdir / "HXT1_mesh.bin"
    write_hxt1(target, header, arrays)
    audit = {"tetrahedra": len(tets), "vertices": len(pts), "inlet_faces": len(tris[TAG_INLET]),
             "outlet_faces": len(tris[TAG_OUTLET]), "wall_faces": len(tris[TAG_WALL]), "inlet_area": ai,
             "outlet_area": ao, "min_tet_volume": float(np.abs(det).min() / 6.0),
             "tag_counts": {str(k): v for k, v in counts.items()}, "hxt1_path": str(target.resolve())}
    (outdir / "mesh_audit.json").write_text(json.dumps(audit, indent=2) + "\n")
    print("NODALS_HXT1_CONVERSION=PASS path=" + str(target), flush=True)
    return audit


def sha256(path: Path) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for b in iter(lambda: f.read(1 << 20), b""):
            h.update(b)
    return h.hexdigest()


def write_msh2(path: Path, points, tets, inlet, outlet, wall) -> None:
    """A tetrahedral mesh as Gmsh ASCII 2.2 with the physical tags above."""
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(points))]
    lines += [f"{i + 1} {float(x)!r} {float(y)!r} {float(z)!r}" for i, (x, y, z) in enumerate(np.asarray(points, dtype=float))]
    lines += ["$EndNodes", "$Elements"]
    recs = []
    for tag, tris in ((TAG_INLET, inlet), (TAG_OUTLET, outlet), (TAG_WALL, wall)):
        recs += [f"2 2 {tag} {tag} " + " ".join(str(int(v) + 1) for v in t) for t in tris]
    recs += [f"4 2 {TAG_VOLUME} {TAG_VOLUME} " + " ".join(str(int(v) + 1) for v in t) for t in tets]
    lines.append(str(len(recs)))
    lines += [f"{i + 1} {r}" for i, r in enumerate(recs)]
    lines.append("$EndElements")
    Path(path).write_text("\n".join(lines) + "\n")
