"""Duct mode on a helical pipe on the GPU: ten SIMPLE iterations against
the oracle (uniform inflow, generic geometry), a converged laminar run, and
ten coupled SST iterations with the vertex wall-distance field against the
oracle's SIMPLE loop with the same hooks."""
import numpy as np
import pytest

from oracle import h8_oracle as H
from oracle import refbind
from oracle import turb_oracle as T

pytestmark = pytest.mark.gpu

# The helix's cells are long along the duct (aspect ratio about 10 here),
# which the aggregation AMG handles slowly: PCG gets up to 200 iterations.
FIX10 = ["--duct", "1", "--max-outer", "10", "--alpha-u", "0.5", "--alpha-p", "0.5", "--p-rtol", "0.5", "--p-max",
         "200", "--amg-smoother", "jacobi", "--pressure-solver", "pcg", "--verbose", "0"]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def test_duct_fixed10_matches_oracle(mesh_dir):
    from nodals_amd._lib import H8

    print("\n[gpu] helix duct fixed10 vs oracle", flush=True)
    h = H8(["--mesh", mesh_dir["helix"]] + FIX10)
    case = H.case_from_product(h)
    r = h.run()
    assert r["outer"] == 10 and r["pass_"] == 1
    ref = H.simple(case, p_solver="pcg", p_max=200, smoother=H.Smoother(kind="jacobi"), max_outer=10)
    errs = [rel(h.get(f"u{d}", case.nv), ref["U"][d]) for d in range(3)] + [rel(h.get("p", case.nc), ref["p"])]
    print(f"[gpu]   rel errors u0 u1 u2 p: {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-9
    h.close()


def test_duct_converges(mesh_dir):
    from nodals_amd._lib import H8

    print("[gpu] helix duct laminar Re=20 to 1e-6", flush=True)
    h = H8(["--mesh", mesh_dir["helix"], "--duct", "1", "--run-mode", "converge", "--max-outer", "6000",
            "--simple-tol", "1e-6", "--amg-smoother", "jacobi", "--pressure-solver", "fgmres", "--p-rtol", "0.5",
            "--p-max", "200", "--amg-precision", "fp32", "--verbose", "0"])
    r = h.run()
    print(f"[gpu]   outer={r['outer']} converged={r['converged']} dp={r['pressure_drop']:.6f}", flush=True)
    assert r["converged"] == 1 and r["pressure_drop"] > 0 and np.isfinite(r["pressure_drop"])
    h.close()


@pytest.mark.skipif(not refbind.sst_available(), reason="reference SST cores not built (oracle/_ref)")
def test_duct_coupled_sst_fixed10_matches_oracle(mesh_dir):
    from nodals_amd._lib import H8

    print("[gpu] helix duct coupled SST fixed10 vs oracle (wall-distance field)", flush=True)
    h = H8(["--mesh", mesh_dir["helix"]] + FIX10)
    h.turb_init(["--sst-wall-active-points", "1"])  # the one-point fit: a scalar root, no rounding-dependent branch
    h.turb_couple(True)
    case = H.case_from_product(h)
    m = T.mesh_from_product(h)
    st = T.Step(m, case, h.get("pipe_geometry", 12), T.Params(active_points=1), refbind.RefSst(),
                wall_distance=h.get("wall_distance", m.nvert))
    r = h.run()
    assert r["outer"] == 10 and r["turb_ms"] > 0
    ref = H.simple(case, p_solver="pcg", p_max=200, smoother=H.Smoother(kind="jacobi"), max_outer=10,
                   extra_blocks=st.eddy_blocks, callback=lambda it, U, p: st.step(U))
    ns = m.ns
    errs = [rel(h.get(f"u{d}", case.nv), ref["U"][d]) for d in range(3)] + [rel(h.get("p", case.nc), ref["p"]),
                                                                          rel(h.get("k", ns), st.k),
                                                                          rel(h.get("omega", ns), st.w)]
    print(f"[gpu]   rel errors u0 u1 u2 p k omega: {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-8
    h.close()
