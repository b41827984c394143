"""Host setup of the product vs the reference's own host builders.

The reference's g2/g3/g4/g5_sa host headers are compiled from the reference
tree into oracle/_ref/libnodals_ref.so; the product's setup must reproduce
them bit for bit: DOF numbering, momentum CSR pattern and diffusion values,
greedy colouring, cell plans and row slots, B coefficients, Dirichlet values
and right-hand sides, the setup rAU and the whole SA hierarchy.
"""
import os

import numpy as np
import pytest

from oracle import refbind

pytestmark = pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built (needs the reference tree)")

REF_MESH_40K = "/root/reference/meshes/pipe/40k/constant/polyMesh"


def _product(mesh, extra=()):
    from nodals_amd._lib import H8

    return H8(["--mesh", mesh, "--max-outer", "10", "--amg-smoother", "cheb2", *extra], device=False)


def _compare_all(mesh, dg_inlet=False):
    R = refbind.RefCase(mesh, dg_inlet=dg_inlet)
    R.build_sa()
    P = _product(mesh, ["--dg-inlet", "1"] if dg_inlet else [])
    I = P.info
    nv, nc, nnz = I["nv"], I["cells"], I["mom_nnz"]
    assert R.nv == nv and R.nc == nc
    exact = [("mom_row", "mom_row", np.int64, nv + 1), ("mom_col", "mom_col", np.int32, nnz),
             ("color", "color", np.int32, nv), ("cell_ref", "cell_ref", np.int32, nc * 8),
             ("cell_slot", "cell_slot", np.uint8, nc * 64), ("mom_val", "mom_diff", np.float64, nnz),
             ("bcoef", "bcoef", np.float64, nc * 24), ("static_rhs", "static_rhs", np.float64, 3 * nv),
             ("fixed_div", "fixed_div", np.float64, nc), ("init_rau", "init_rau", np.float64, nv),
             ("fixed_value", "fixed_value", np.float64, 3 * I["nfixed"]), ("cell_det", "cell_det", np.float64, nc),
             ("cell_invJ", "cell_invJ", np.float64, 9 * nc)]
    for rname, pname, dt, n in exact:
        a = R.get(rname, dt)
        b = P.get(pname, n, dt)
        assert np.array_equal(a, b), f"{rname} differs from the reference"
    g2 = R.get("g2free", np.int32)
    assert np.array_equal(g2, P.get("g2free", g2.size, np.int32))
    dims = R.get("sa_dims", np.int32)
    assert I["amg_levels"] == dims[0] and I["terminal"] == dims[2]
    assert np.array_equal(R.get("sa_agg0", np.int32), P.get("agg0", nc, np.int32))
    for l in range(int(dims[1])):
        pr = R.get(f"sa_P{l}_row", np.int64)
        assert np.array_equal(pr, P.get(f"P{l}_row", pr.size, np.int64))
        assert np.array_equal(R.get(f"sa_P{l}_col", np.int32), P.get(f"P{l}_col", int(pr[-1]), np.int32))
        assert np.array_equal(R.get(f"sa_P{l}_val", np.float64), P.get(f"P{l}_val", int(pr[-1])))
    for l in range(int(dims[0])):
        ar = R.get(f"sa_A{l}_row", np.int64)
        assert np.array_equal(ar, P.get(f"A{l}_row", ar.size, np.int64))
        assert np.array_equal(R.get(f"sa_A{l}_col", np.int32), P.get(f"A{l}_col", int(ar[-1]), np.int32))
        assert np.array_equal(R.get(f"sa_A{l}_val", np.float64), P.get(f"A{l}_val", int(ar[-1])))
    t = R.get("sa_terminal_inv", np.float64)
    assert np.array_equal(t, P.get("terminal_inv", t.size))
    lam = R.get("sa_lambda", np.float64)
    assert lam[0] == P.get("fine_lambda", 1)[0]
    return R, P


def test_small_pipe_bitwise(mesh_dir):
    _compare_all(mesh_dir["small"])


def test_medium_pipe_bitwise(mesh_dir):
    _compare_all(mesh_dir["medium"])


@pytest.mark.parametrize("which", ["small", "medium"])
def test_dg_inlet_setup_bitwise(mesh_dir, which):
    """build_g4_setup with the DG weak inlet (g4_host.hpp:289-313): free
    inlet DOFs, the effective B without the inlet face trace, the prescribed
    inflow flux in the fixed divergence and the setup rAU of the plug start;
    the inlet really changes the numbering and the B table."""
    R, P = _compare_all(mesh_dir[which], dg_inlet=True)
    R0 = refbind.RefCase(mesh_dir[which])
    assert R.nv > R0.nv
    plug = P.get("plug", 3 * R.nv).reshape(R.nv, 3)
    pipe = P.get("pipe_geometry", 12)
    g2 = R.get("g2free", np.int32)
    npts = P.get("mesh_sizes", 4, np.int64)[0]
    free_vert = g2[:npts][g2[:npts] >= 0]
    assert np.array_equal(plug[free_vert], np.tile(pipe[6:9], (free_vert.size, 1)))
    assert np.all(np.delete(plug, free_vert, axis=0) == 0)


@pytest.mark.skipif(not os.path.isdir(REF_MESH_40K), reason="reference 40k pipe mesh not present")
@pytest.mark.parametrize("dg_inlet", [False, True])
def test_reference_40k_pipe_bitwise(dg_inlet):
    """The reference's own shipped mesh (40,620 unstructured tets), with and
    without the DG inlet."""
    _compare_all(REF_MESH_40K, dg_inlet)


def test_unsmoothed_hierarchy_when_no_spectrum(mesh_dir):
    """Jacobi smoother + spectrum policy auto runs no power iteration; the
    reference then degenerates to plain aggregation (lambda=0, P entries 1)."""
    from nodals_amd._lib import H8

    P = H8(["--mesh", mesh_dir["small"], "--max-outer", "10", "--amg-smoother", "jacobi"], device=False)
    pr = P.get("P0_row", P.info["cells"] + 1, np.int64)
    pv = P.get("P0_val", int(pr[-1]))
    assert np.all(np.diff(pr) == 1) and np.all(pv == 1.0)


def test_pressure_drop_fit_matches_reference(mesh_dir):
    R = refbind.RefCase(mesh_dir["small"])
    rng = np.random.default_rng(3)
    p = rng.standard_normal(R.nc)
    from oracle import h8_oracle

    P = _product(mesh_dir["small"])
    case = h8_oracle.case_from_product(P, with_hierarchy=False)
    assert abs(h8_oracle.pressure_drop_fit(case, p) - R.pressure_drop(p)) <= 1e-12 * abs(R.pressure_drop(p))
