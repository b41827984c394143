"""Lid-driven cavity mode of the H8 path (config: unit cube of Kuhn tets,
lid u = (1, 0, 0), no-slip elsewhere, pressure pinned in cell 0)."""
import numpy as np
import pytest
import scipy.sparse as sp


@pytest.fixture(scope="module")
def cavity(tmp_path_factory):
    from nodals_amd import meshgen

    return str(meshgen.write_cavity_case(tmp_path_factory.mktemp("cav") / "c4", n=4))


def _h8(mesh, **kw):
    from nodals_amd._lib import H8

    return H8(["--mesh", mesh, "--lid", "lid", "--wall", "walls", "--re", "100", "--max-outer", "10",
               "--amg-smoother", "jacobi"], **kw)


def test_cavity_boundary_values_and_closed_domain(cavity):
    h = _h8(cavity, device=False)
    nc = h.info["cells"]
    pipe = h.get("pipe", 5)
    assert pipe[0] == pytest.approx(0.01)  # nu = U L / Re
    assert int(h.get("pin_cell", 1)[0]) == 0
    fv = h.get("fixed_value", h.info["nfixed"] * 3).reshape(-1, 3)
    assert np.all(fv[:, 1:] == 0.0) and fv[:, 0].max() == 1.0
    # no inflow or outflow: the fixed-velocity divergence sums to zero, so the
    # pressure system is consistent
    fd = h.get("fixed_div", nc)
    assert abs(fd.sum()) < 1e-13
    # every vertex on the box surface is fixed: 4^3 cube has 5^3 - 3^3 = 98
    # boundary vertices and 6 * 2 * 16 = 192 boundary faces
    assert h.info["nfixed"] == 98 + 192
    h.close()


def test_cavity_pinned_operator_is_spd(cavity):
    from oracle import h8_oracle as O

    h = _h8(cavity, device=False)
    case = O.case_from_product(h)
    A, _ = O.fine_refresh(case, case.init_rau)
    M = A.toarray()
    assert np.allclose(M, M.T, atol=1e-14 * abs(M).max())
    assert np.linalg.eigvalsh(M).min() > 0
    assert np.all(M[0, 1:] == 0) and np.all(M[1:, 0] == 0)
    # without the pin the operator has the constants in its kernel
    case.pin = -1
    A1, _ = O.fine_refresh(case, case.init_rau)
    assert np.abs(A1 @ np.ones(case.nc)).max() < 1e-12 * abs(A1).max()
    h.close()
