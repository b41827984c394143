"""The reference-named solve entry points on the GPU, against the calls they
wrap (the calls themselves are tested against the oracle elsewhere):

* nodals_h8_momentum_solve "m3-fgs" = the wall path's block sweeps
  (nodals_h8_wall_momentum) bit for bit, "fgs1" = nodals_h8_momentum_work,
  with residuals that drop;
* nodals_h8_pcg_amg reaches its tolerance on the live fine Schur operator
  (checked with nodals_h8_fine_spmv) and equals nodals_h8_pressure_solve on
  the continuity right-hand side;
* nodals_h8_scalar_solve for k then omega = one nodals_h8_turb_step bit for
  bit, with residuals that drop."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WALL = ["--duct", "1", "--wall-model", "spalding", "--re", "5000", "--alpha-u", "0.7", "--alpha-p", "0.3",
        "--amg-hierarchy", "cf", "--cf-interp", "exti", "--amg-smoother", "mcgs", "--pressure-solver", "pcg",
        "--p-rtol", "1e-3", "--p-max", "600", "--max-outer", "10", "--verbose", "0"]


def _wall_pair(mesh):
    from nodals_amd._lib import H8

    hs = []
    for _ in range(2):
        h = H8(["--mesh", mesh, *WALL])
        h.turb_init([])
        h.wall_init([])
        hs.append(h)
    return hs


def test_momentum_solve_m3_is_the_wall_block_sweep(mesh_dir):
    print("\n[gpu] momentum_solve m3-fgs vs wall_momentum, small helix", flush=True)
    a, b = _wall_pair(mesh_dir["helix"])
    nv = a.info["nv"]
    for h in (a, b):
        h.wall_assemble()
    r = a.momentum_solve("m3-fgs", 20, 0.7)
    b.wall_momentum()
    for d in range(3):
        assert np.array_equal(a.get(f"u{d}", nv).view(np.int64), b.get(f"u{d}", nv).view(np.int64))
    print(f"[gpu]   sweeps {r['its']} residual {r['r0']:.3e} -> {r['rel']:.3e}", flush=True)
    assert r["its"] == 20 and r["ok"] and r["rel"] < r["r0"]
    a.close()
    b.close()


def test_momentum_solve_fgs1_and_pcg_amg(mesh_dir):
    from nodals_amd._lib import H8

    print("\n[gpu] momentum_solve fgs1 and pcg_amg, small pipe", flush=True)
    args = ["--mesh", mesh_dir["small"], "--max-outer", "10", "--verbose", "0"]
    a, b = H8(args), H8(args)
    nv, nc = a.info["nv"], a.info["cells"]
    for h in (a, b):
        h.assemble(0.5)
    r = a.momentum_solve("fgs1", 1, 1.0)
    b.momentum_work("fgs1", 1.0, 0)
    for d in range(3):
        assert np.array_equal(a.get(f"u{d}", nv).view(np.int64), b.get(f"u{d}", nv).view(np.int64))
    assert r["its"] == 1 and r["rel"] < r["r0"]
    # PCG-AMG on the refreshed fine operator
    a.continuity()
    a.pressure_refresh()
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(nc)
    x, res = a.pcg_amg(rhs, 1e-10, 0.0, 400)
    err = np.linalg.norm(a.fine_spmv(x) - rhs) / np.linalg.norm(rhs)
    print(f"[gpu]   pcg_amg its {res['its']} rel {res['rel']:.2e} true {err:.2e}", flush=True)
    assert res["ok"] and err < 1e-9
    # on the continuity right-hand side it is the pressure solve
    b.continuity()
    b.pressure_refresh()
    cont = b.get("cont", nc)
    xb, _ = b.pcg_amg(cont, 1e-6, 0.0, 200)
    b.continuity()
    pr = b.pressure_solve("pcg", 1e-6, 0.0, 200)
    dp = b.get("dp", nc)
    assert np.array_equal(xb.view(np.int64), dp.view(np.int64)), pr
    a.close()
    b.close()


def test_scalar_solve_is_the_turbulence_step(mesh_dir):
    from nodals_amd._lib import H8

    print("\n[gpu] scalar_solve k, omega vs turb_step, small pipe", flush=True)
    args = ["--mesh", mesh_dir["small"], "--max-outer", "10", "--verbose", "0", "--re", "200"]
    a, b = H8(args), H8(args)
    rng = np.random.default_rng(5)
    nv = a.info["nv"]
    u = [0.1 * rng.standard_normal(nv) for _ in range(3)]
    for h in (a, b):
        h.turb_init([])
        for d in range(3):
            h.set(f"u{d}", u[d])
    rk = a.scalar_solve(0)
    rw = a.scalar_solve(1)
    b.turb_step()
    ns = a.turb["ns"]
    for name in ("k", "omega"):
        assert np.array_equal(a.get(name, ns).view(np.int64), b.get(name, ns).view(np.int64)), name
    print(f"[gpu]   k {rk['r0']:.3e} -> {rk['rel']:.3e}; omega {rw['r0']:.3e} -> {rw['rel']:.3e}", flush=True)
    assert rk["rel"] < rk["r0"] and rw["rel"] < rw["r0"]
    a.close()
    b.close()
