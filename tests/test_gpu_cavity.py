"""Lid-driven cavity (Re = 100, lid u = (1, 0, 0), pressure pinned in cell 0)
on the GPU: ten SIMPLE iterations against the oracle (1e-9 relative),
converged fields against the oracle's converged SIMPLE (1e-6), and the
centreline velocity of a converged 8^3 cavity within the range of the
published 3-D Re = 100 cubic-cavity profile (a sanity band: minimum of u on
the vertical centreline near -0.2)."""
import numpy as np
import pytest

from oracle import h8_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def cav(tmp_path_factory):
    from nodals_amd import meshgen

    d = tmp_path_factory.mktemp("cavity")
    return {"c4": str(meshgen.write_cavity_case(d / "c4", n=4)), "c8": str(meshgen.write_cavity_case(d / "c8", n=8))}


ARGS = ["--lid", "lid", "--wall", "walls", "--re", "100", "--amg-smoother", "jacobi", "--pressure-solver", "fgmres",
        "--verbose", "0"]


def test_cavity_fixed10_matches_oracle(cav):
    from nodals_amd._lib import H8

    print("\n[gpu] cavity 4^3 fixed10 vs oracle", flush=True)
    h = H8(["--mesh", cav["c4"], "--run-mode", "fixed10", "--max-outer", "10", "--p-rtol", "0.5", "--p-max", "20"]
           + ARGS)
    r = h.run()
    assert r["pass_"] == 1
    case = O.case_from_product(h)
    assert case.pin == 0
    ref = O.simple(case, p_solver="fgmres", p_rtol=0.5, p_max=20, smoother=O.Smoother(kind="jacobi"),
                   run_mode="fixed10", max_outer=10)
    for d in range(3):
        assert rel(h.get(f"u{d}", case.nv), ref["U"][d]) < 1e-9
    assert rel(h.get("p", case.nc), ref["p"]) < 1e-9
    h.close()


def test_cavity_converged_matches_oracle_and_physics(cav):
    from nodals_amd._lib import H8

    print("\n[gpu] cavity 8^3 converge vs oracle", flush=True)
    h = H8(["--mesh", cav["c8"], "--run-mode", "converge", "--max-outer", "5000", "--simple-tol", "1e-7",
            "--p-rtol", "1e-2", "--p-max", "200"] + ARGS)
    r = h.run()
    assert r["converged"] == 1, r
    case = O.case_from_product(h)
    ref = O.simple(case, p_solver="fgmres", p_rtol=1e-2, p_max=200, smoother=O.Smoother(kind="jacobi"),
                   run_mode="converge", max_outer=5000, simple_tol=1e-7)
    assert ref["converged"]
    U = [h.get(f"u{d}", case.nv) for d in range(3)]
    for d in range(3):
        assert rel(U[d], ref["U"][d]) < 1e-6
    # u on the vertical centreline x = y = 0.5 (vertex DOFs of the points
    # on that line)
    pts = _points(cav["c8"])
    g2free = h.get("g2free", h.info["nv"] + h.info["nfixed"], np.int32)
    on = np.flatnonzero((np.abs(pts[:, 0] - 0.5) < 1e-9) & (np.abs(pts[:, 1] - 0.5) < 1e-9))
    free = g2free[on] >= 0
    umin = float(U[0][g2free[on][free]].min())
    print(f"[gpu] cavity 8^3 centreline min u = {umin:.4f}", flush=True)
    assert -0.35 < umin < -0.08
    h.close()


def _points(mesh):
    """Points of an ASCII polyMesh (count, then one "(x y z)" per line)."""
    from pathlib import Path

    lines = (Path(mesh) / "points").read_text().splitlines()
    k = next(i for i, l in enumerate(lines) if l.strip().isdigit())
    n = int(lines[k])
    return np.array([[float(v) for v in lines[k + 2 + i].strip("()").split()] for i in range(n)])
