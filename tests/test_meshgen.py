"""Structured mesh generators and the polyMesh writer/reader round trip."""
import numpy as np
import pytest

from nodals_amd import meshgen


def _vol(p, t):
    return np.einsum("ij,ij->i", np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]]), p[t[:, 3]] - p[t[:, 0]]) / 6


def test_kuhn_cube_counts_and_volume():
    p, t = meshgen.kuhn_cube(4)
    assert p.shape == (125, 3) and t.shape == (6 * 64, 4)
    v = _vol(p, t)
    assert np.allclose(np.abs(v), 1 / (6 * 64)) and abs(np.abs(v).sum() - 1.0) < 1e-14
    # every Kuhn tet contains the hex main diagonal v000-v111
    assert np.all(t[:, 3] - t[:, 0] == 1 + 5 + 25)


@pytest.mark.parametrize("core,ring,quads", [(8, 5, 224), (12, 9, 576)])
def test_ogrid_quad_counts(core, ring, quads):
    pts, q = meshgen.ogrid_section(core, ring)
    assert q.shape[0] == quads
    assert pts.shape[0] == (core + 1) ** 2 + ring * 4 * core


def test_config_pipe_counts():
    p, t = meshgen.straight_pipe(8, 5, 30)
    assert t.shape[0] == 40320 and p.shape[0] == 7471
    v = np.abs(_vol(p, t))
    assert v.min() > 0
    # polygonal O-grid volume below the circle, within 1%
    assert 0.99 * np.pi * 0.25 * 10 < v.sum() < np.pi * 0.25 * 10


def test_helix_counts():
    p, t = meshgen.helical_pipe(12, 9, 60)
    assert t.shape[0] == 207360
    assert np.abs(_vol(p, t)).min() > 0


def test_polymesh_conformity_and_patches(tmp_path):
    d = meshgen.write_pipe_case(tmp_path / "pipe", core=4, ring=2, nslab=3)
    from nodals_amd._lib import H8

    h = H8(["--mesh", str(d), "--max-outer", "10"], device=False)
    assert h.info["cells"] == 48 * 3 * 6
    mesh = meshgen.tets_to_polymesh(*meshgen.straight_pipe(4, 2, 3),
                                    lambda fc, n: np.zeros(fc.shape[0], dtype=int), ["all"])
    # closed surface: boundary faces = 2 * quads (ends) + wall faces
    nb = mesh.faces.shape[0] - mesh.neighbour.size
    assert nb == 2 * 48 * 2 + 16 * 3 * 2
    assert np.all(mesh.owner[: mesh.neighbour.size] < mesh.neighbour)
