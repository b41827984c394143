"""Duct mode (--duct 1) setup on a helical pipe (no GPU).

Pins: the inlet faces of the generated helix are its first section; the
inlet Dirichlet values carry the uniform inflow bulk * (-n) on every inlet
vertex off the wall and the bubbles make every inlet face's average exact,
so the boundary flux through B equals -bulk * A_inlet; D is the hydraulic
diameter 2 sqrt(A/pi); the vertex wall-distance field equals a brute-force
point-to-triangle minimum over every wall face; the patch pressure drop is
the difference of the area-weighted inlet and outlet cell pressures.
"""
import numpy as np
import pytest

from oracle import turb_oracle as T

BASE = ["--duct", "1", "--run-mode", "converge", "--verbose", "0"]


@pytest.fixture(scope="module")
def helix(mesh_dir):
    from nodals_amd._lib import H8

    h = H8(["--mesh", mesh_dir["helix"]] + BASE, device=False)
    yield h, T.mesh_from_product(h)
    h.close()


def brute_distance(m, wall_patch):
    wall = np.flatnonzero(m.face_patch == wall_patch)
    a, b, c = (m.points[m.faces[wall]][:, k] for k in range(3))

    def seg(p, x, y):
        t = np.clip(((p - x) * (y - x)).sum(-1) / ((y - x) ** 2).sum(-1), 0, 1)
        return np.linalg.norm(p - (x + (y - x) * t[:, None]), axis=-1)

    n = np.cross(b - a, c - a)
    nn = n / np.linalg.norm(n, axis=1, keepdims=True)
    out = np.empty(m.points.shape[0])
    for i, p in enumerate(m.points):
        h = ((p - a) * nn).sum(-1)
        q = p - h[:, None] * nn
        # barycentric inside test of the projection
        v0, v1, v2 = b - a, c - a, q - a
        d00, d01, d11 = (v0 * v0).sum(-1), (v0 * v1).sum(-1), (v1 * v1).sum(-1)
        d20, d21 = (v2 * v0).sum(-1), (v2 * v1).sum(-1)
        den = d00 * d11 - d01 * d01
        v = (d11 * d20 - d01 * d21) / den
        w = (d00 * d21 - d01 * d20) / den
        inside = (v >= 0) & (w >= 0) & (v + w <= 1)
        dface = np.where(inside, np.abs(h), np.inf)
        out[i] = np.minimum.reduce([dface, seg(p, a, b), seg(p, b, c), seg(p, c, a)]).min()
    return out


def test_helix_patches_and_inflow(helix):
    h, m = helix
    pg = h.get("pipe_geometry", 12)
    inlet, wall = int(pg[9]), int(pg[10])
    nin = int((m.face_patch == inlet).sum())
    nout = int((m.face_patch == int(pg[11])).sum())
    assert nin == nout == 2 * 48  # one section: 16 core + 32 ring quads, two triangles each
    f = np.flatnonzero(m.face_patch == inlet)
    tri = m.points[m.faces[f]]
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1).sum()
    D, bulk = pg[3], pg[4]
    assert abs(D - 2.0 * np.sqrt(area / np.pi)) < 1e-12
    u_in = pg[6:9]
    assert abs(np.linalg.norm(u_in) - bulk) < 1e-12
    # the inlet velocity points into the duct, along the helix tangent at its start
    assert u_in[1] > 0.9 and abs(u_in[0]) < 1e-12
    fd = h.get("fixed_div", h.info["cells"])
    assert abs(fd.sum() + bulk * area) < 1e-12 * area


def test_inlet_values_uniform(helix):
    h, m = helix
    pg = h.get("pipe_geometry", 12)
    inlet, wall = int(pg[9]), int(pg[10])
    slot = h.get("fixed_slot", m.ns)
    fv = h.get("fixed_value", h.info["nfixed"] * 3).reshape(-1, 3)
    on_wall = np.zeros(m.nvert, bool)
    on_wall[m.faces[m.face_patch == wall].ravel()] = True
    inlet_v = np.unique(m.faces[m.face_patch == inlet].ravel())
    for v in inlet_v:
        want = np.zeros(3) if on_wall[v] else pg[6:9]
        assert np.allclose(fv[slot[v]], want, atol=0, rtol=0)
    # each inlet face: vertex mean + 9/20 bubble = exact face average (the inlet velocity)
    for f in np.flatnonzero(m.face_patch == inlet):
        vm = fv[slot[m.faces[f]]].mean(axis=0)
        bub = fv[slot[m.nvert + f]]
        assert np.allclose(vm + 9.0 / 20.0 * bub, pg[6:9], rtol=0, atol=1e-14)


def test_wall_distance_brute_force(helix):
    h, m = helix
    pg = h.get("pipe_geometry", 12)
    wd = h.get("wall_distance", m.nvert)
    ref = brute_distance(m, int(pg[10]))
    assert np.abs(wd - ref).max() < 1e-12
    on_wall = np.zeros(m.nvert, bool)
    on_wall[m.faces[m.face_patch == int(pg[10])].ravel()] = True
    assert np.all(wd[on_wall] == 0.0) and np.all(wd[~on_wall] > 0.0)


def test_straight_pipe_duct_mode_matches_pipe_geometry(mesh_dir):
    from nodals_amd._lib import H8

    hp = H8(["--mesh", mesh_dir["small"], "--run-mode", "converge", "--verbose", "0"], device=False)
    hd = H8(["--mesh", mesh_dir["small"]] + BASE, device=False)
    gp, gd = hp.get("pipe_geometry", 12), hd.get("pipe_geometry", 12)
    # same inflow direction; D from the polygonal section's area, within 3% of the circle's
    assert np.allclose(gp[6:9], gd[6:9], atol=1e-14)
    assert abs(gd[3] - gp[3]) < 0.03 * gp[3]
    hp.close()
    hd.close()
