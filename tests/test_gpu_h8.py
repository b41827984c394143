"""Parity of the HIP H8 kernels with the CPU FP64 oracle, through the C ABI.

Setup data for the oracle come from the product host (pinned bit-exactly
against the reference builders in test_setup_vs_reference.py).  Tolerances:
FP64 device arithmetic differs from the oracle only by summation order and
FMA contraction, so single-kernel results must agree to ~1e-12 relative;
whole SIMPLE trajectories are compared to 1e-9 (fixed 10 iterations) and
converged fields to 1e-6 relative L2 (the tolerance the case definition
states).
"""
import numpy as np
import pytest

from oracle import h8_oracle as O

pytestmark = pytest.mark.gpu

BASE = ["--max-outer", "10", "--alpha-u", "0.5", "--alpha-p", "0.5", "--p-rtol", "0.5", "--p-max", "20"]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def pipe(mesh_dir):
    from nodals_amd._lib import H8

    print(f"\n[gpu] H8 create mesh={mesh_dir['small']} precision=fp64", flush=True)
    h = H8(["--mesh", mesh_dir["small"], *BASE, "--amg-smoother", "jacobi", "--pressure-solver", "pcg"])
    case = O.case_from_product(h)
    yield h, case
    h.close()


def _random_state(case, seed=7):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((3, case.nv)) * 0.3
    p = rng.standard_normal(case.nc)
    return U, p


def _set_state(h, U, p):
    for d in range(3):
        h.set(f"u{d}", U[d])
    h.set("p", p)


def test_assembly_relaxation_rhs(pipe):
    h, case = pipe
    U, p = _random_state(case)
    _set_state(h, U, p)
    h.assemble(0.5)
    _, T = O.tensors()
    aval, conv, _ = O.assemble_central(case, U, T)
    delta, diag, rau = O.finalize_relax(case, aval, 0.5)
    b = O.momentum_rhs(case, conv, O.bt_apply(case, p), delta, U)
    assert rel(h.get("mom_val", case.nnz), aval) < 1e-12
    assert rel(h.get("rau", case.nv), rau) < 1e-12
    assert rel(h.get("delta", case.nv), delta) < 1e-12
    assert rel(h.get("diag", case.nv), diag) < 1e-12
    for d in range(3):
        assert rel(h.get(f"b{d}", case.nv), b[d]) < 1e-11


def test_gauss_seidel_sweep(pipe):
    h, case = pipe
    U, p = _random_state(case, 11)
    _set_state(h, U, p)
    h.assemble(0.7)
    _, T = O.tensors()
    aval, conv, _ = O.assemble_central(case, U, T)
    delta, diag, rau = O.finalize_relax(case, aval, 0.7)
    b = O.momentum_rhs(case, conv, O.bt_apply(case, p), delta, U)
    A = O.mom_matrix(case, aval)
    for mode, omega in (("fgs1", 1.0), ("sgs1", 0.8)):
        _set_state(h, U, p)
        h.momentum_work(mode, omega, 1)
        ref = O.gs_sweep(case, A, diag, b, U, omega)
        if mode == "sgs1":
            ref = O.gs_sweep(case, A, diag, b, ref, omega, reverse=True)
        for d in range(3):
            assert rel(h.get(f"u{d}", case.nv), ref[d]) < 1e-11, mode
    _set_state(h, U, p)
    assert np.allclose(h.momentum_audit(), O.momentum_audit(A, b, U), rtol=1e-10)


def test_continuity(pipe):
    h, case = pipe
    U, p = _random_state(case, 5)
    _set_state(h, U, p)
    n = h.continuity()
    r = O.continuity(case, U)
    assert abs(n - np.linalg.norm(r)) <= 1e-12 * np.linalg.norm(r)
    assert rel(h.get("cont", case.nc), -r) < 1e-12


def test_fine_schur_pattern_and_refresh(pipe):
    h, case = pipe
    U, p = _random_state(case, 3)
    _set_state(h, U, p)
    h.assemble(0.5)
    h.pressure_refresh()
    topo = O.fine_topology(case)
    assert np.array_equal(h.get("fine_row", case.nc + 1), topo["row"])
    assert np.array_equal(h.get("fine_col", int(topo["row"][-1])), topo["col"])
    rau = h.get("rau", case.nv)
    Af, fdiag = O.fine_refresh(case, rau)
    assert rel(h.get("fine_val", Af.nnz), Af.data) < 1e-12
    assert rel(h.get("fine_diag", case.nc), fdiag) < 1e-12
    x = np.random.default_rng(1).standard_normal(case.nc)
    assert rel(h.fine_spmv(x), Af @ x) < 1e-12
    assert rel(h.fine_spmv(x), O.schur_matrix_free(case, rau, x)) < 1e-12


def test_vcycle_matches_oracle(pipe):
    h, case = pipe
    U, p = _random_state(case, 9)
    _set_state(h, U, p)
    h.assemble(0.5)
    h.pressure_refresh()
    Af, fdiag = O.fine_refresh(case, h.get("rau", case.nv))
    vc = O.VCycle(case.hier, O.Smoother(kind="jacobi", jacobi_omega=0.7))
    vc.set_fine(Af, fdiag)
    r = np.random.default_rng(2).standard_normal(case.nc)
    assert rel(h.precond(r), vc.apply(r)) < 1e-12


_EAGER_SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
from nodals_amd._lib import H8
h = H8(["--mesh", sys.argv[2], "--max-outer", "10", "--amg-smoother", "jacobi", "--verbose", "0"])
rng = np.random.default_rng(7)
U = rng.standard_normal((3, h.info["nv"])) * 0.3
for d in range(3):
    h.set(f"u{d}", U[d])
h.assemble(0.5)
h.pressure_refresh()
r = np.random.default_rng(2).standard_normal(h.info["cells"])
z = h.precond(r)
h.momentum_work("fgs1", 1.0, 1)
np.save(sys.argv[3], np.concatenate([z, h.get("u0", h.info["nv"])]))
"""


def test_graph_replays_match_eager(mesh_dir, tmp_path):
    """V-cycle and GS-sweep graphs replayed several times, with eager work and
    synchronisations in between, equal the eager (graphs disabled) launch
    sequence bit for bit on every replay."""
    import os
    import subprocess
    import sys
    from pathlib import Path

    from nodals_amd._lib import H8

    root = str(Path(__file__).resolve().parents[1])
    script = tmp_path / "eager.py"
    script.write_text(_EAGER_SCRIPT)
    env = dict(os.environ, NODALS_HIP_GRAPHS="0")
    print("\n[gpu] eager (graphs disabled) reference in a subprocess", flush=True)
    subprocess.run([sys.executable, str(script), root, mesh_dir["small"], str(tmp_path / "eager.npy")], env=env,
                   check=True, timeout=240)
    eager = np.load(tmp_path / "eager.npy")
    h = H8(["--mesh", mesh_dir["small"], "--max-outer", "10", "--amg-smoother", "jacobi", "--verbose", "0"])
    nc, nv = h.info["cells"], h.info["nv"]
    rng = np.random.default_rng(7)
    U = rng.standard_normal((3, nv)) * 0.3
    r = np.random.default_rng(2).standard_normal(nc)
    for rep in range(4):
        for d in range(3):
            h.set(f"u{d}", U[d])
        h.assemble(0.5)
        h.pressure_refresh()
        h.continuity()  # eager work + host synchronisation between replays
        z = h.precond(r)
        assert np.array_equal(z, eager[:nc]), rep
        h.momentum_work("fgs1", 1.0, 1)
        assert np.array_equal(h.get("u0", nv), eager[nc:]), rep
    h.close()


@pytest.mark.parametrize("solver", ["pcg", "richardson", "fgmres"])
def test_pressure_solve(pipe, solver):
    h, case = pipe
    U, p = _random_state(case, 13)
    _set_state(h, U, p)
    h.assemble(0.5)
    h.continuity()
    h.pressure_refresh()
    # stationary Richardson contracts slowly on the unsmoothed hierarchy
    rtol = 5e-2 if solver == "richardson" else 1e-8
    res = h.pressure_solve(solver, rtol, 1e-14, 400)
    assert res["ok"]
    Af, fdiag = O.fine_refresh(case, h.get("rau", case.nv))
    rhs = -O.continuity(case, U)
    x = h.get("dp", case.nc)
    assert np.linalg.norm(rhs - Af @ x) <= 1.01 * rtol * np.linalg.norm(rhs)
    vc = O.VCycle(case.hier, O.Smoother(kind="jacobi", jacobi_omega=0.7))
    vc.set_fine(Af, fdiag)
    fn = {"pcg": O.pcg, "richardson": O.richardson, "fgmres": O.fgmres}[solver]
    xo, ro = fn(Af, vc.apply, rhs, rtol, 1e-14, 400)
    assert abs(res["its"] - ro["its"]) <= 1
    assert rel(x, xo) < 1e-6


def test_simple_fixed10_matches_oracle(mesh_dir):
    from nodals_amd._lib import H8

    print("\n[gpu] SIMPLE fixed10 small pipe fp64 pcg+jacobi", flush=True)
    h = H8(["--mesh", mesh_dir["small"], *BASE, "--amg-smoother", "jacobi", "--pressure-solver", "pcg",
            "--verbose", "0"])
    case = O.case_from_product(h)
    r = h.run()
    assert r["pass_"] == 1 and r["outer"] == 10
    ref = O.simple(case, p_solver="pcg", smoother=O.Smoother(kind="jacobi"), max_outer=10)
    for d in range(3):
        assert rel(h.get(f"u{d}", case.nv), ref["U"][d]) < 1e-9
    assert rel(h.get("p", case.nc), ref["p"]) < 1e-9
    h.close()


@pytest.mark.parametrize("work,sweeps", [("fgs1", 4), ("sgs1", 2)])
def test_simple_fixed10_momentum_sweeps_match_oracle(mesh_dir, work, sweeps):
    """Several momentum work units per SIMPLE iteration at omega 0.7 (the SST
    path's fixed M3 sweep count and relaxation) against the oracle's loop
    with the same count."""
    from nodals_amd._lib import H8

    print(f"\n[gpu] SIMPLE fixed10 small pipe momentum {work} x{sweeps} omega 0.7", flush=True)
    h = H8(["--mesh", mesh_dir["small"], *BASE, "--amg-smoother", "jacobi", "--pressure-solver", "pcg",
            "--momentum-work", work, "--mom-sweeps", str(sweeps), "--mom-omega", "0.7", "--verbose", "0"])
    case = O.case_from_product(h)
    r = h.run()
    assert r["pass_"] == 1 and r["outer"] == 10
    ref = O.simple(case, p_solver="pcg", smoother=O.Smoother(kind="jacobi"), max_outer=10, momentum_work=work,
                   mom_sweeps=sweeps, mom_omega=0.7)
    one = O.simple(case, p_solver="pcg", smoother=O.Smoother(kind="jacobi"), max_outer=10, momentum_work=work,
                   mom_omega=0.7)
    for d in range(3):
        got = h.get(f"u{d}", case.nv)
        assert rel(got, ref["U"][d]) < 1e-9
        assert rel(got, one["U"][d]) > 1e-6  # the extra sweeps change the trajectory
    assert rel(h.get("p", case.nc), ref["p"]) < 1e-9
    h.close()


def test_deterministic_bitwise(mesh_dir):
    from nodals_amd._lib import H8

    out = []
    for k in range(2):
        h = H8(["--mesh", mesh_dir["small"], *BASE, "--amg-smoother", "jacobi", "--pressure-solver", "fgmres",
                "--verbose", "0"])
        h.run()
        out.append((h.get("u2", h.info["nv"]), h.get("p", h.info["cells"])))
        h.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_mixed_precision_amg_reaches_same_fields(mesh_dir, precision):
    """FP32 / BF16-finest AMG inside the FP64 FGMRES: converged SIMPLE fields
    equal the FP64-AMG fields within 1e-6 relative L2."""
    from nodals_amd._lib import H8

    args = ["--mesh", mesh_dir["small"], "--run-mode", "converge", "--max-outer", "3000", "--simple-tol", "1e-9",
            "--amg-smoother", "jacobi", "--pressure-solver", "fgmres", "--p-rtol", "1e-2", "--p-max", "200", "--verbose", "0"]
    fields = {}
    for prec in ("fp64", precision):
        print(f"\n[gpu] SIMPLE converge small pipe amg={prec}", flush=True)
        h = H8(args + ["--amg-precision", prec])
        r = h.run()
        assert r["converged"] == 1, (prec, r)
        fields[prec] = [h.get(f"u{d}", h.info["nv"]) for d in range(3)] + [h.get("p", h.info["cells"])]
        h.close()
    for a, b in zip(fields[precision], fields["fp64"]):
        assert rel(a, b) < 1e-6


def test_converged_fields_match_oracle(mesh_dir):
    """Converged FP64 SIMPLE fields vs the oracle's own SIMPLE iteration."""
    from nodals_amd._lib import H8

    print("\n[gpu] SIMPLE converge small pipe fp64 vs oracle", flush=True)
    h = H8(["--mesh", mesh_dir["small"], "--run-mode", "converge", "--max-outer", "3000", "--simple-tol", "1e-8",
            "--amg-smoother", "jacobi", "--pressure-solver", "fgmres", "--p-rtol", "1e-2", "--p-max", "200", "--verbose", "0"])
    r = h.run()
    assert r["converged"] == 1
    case = O.case_from_product(h)
    ref = O.simple(case, p_solver="fgmres", p_rtol=1e-2, p_max=200, smoother=O.Smoother(kind="jacobi"), run_mode="converge",
                   max_outer=3000, simple_tol=1e-8)
    assert ref["converged"]
    for d in range(3):
        assert rel(h.get(f"u{d}", case.nv), ref["U"][d]) < 1e-6
    p = h.get("p", case.nc)
    assert rel(p, ref["p"]) < 1e-6
    assert abs(O.pressure_drop_fit(case, p) - O.pressure_drop_fit(case, ref["p"])) <= 1e-3 * abs(
        O.pressure_drop_fit(case, ref["p"]))
    h.close()


@pytest.mark.parametrize("interp", ["exti", "direct"])
def test_cf_vcycle_matches_oracle(mesh_dir, interp):
    """V-cycle on the CF/PMIS hierarchy (the translator's default) equals the
    oracle V-cycle on the same hierarchy."""
    from nodals_amd._lib import H8

    print(f"\n[gpu] H8 create small pipe amg-hierarchy=cf interp={interp}", flush=True)
    h = H8(["--mesh", mesh_dir["small"], *BASE, "--amg-smoother", "jacobi", "--pressure-solver", "pcg",
            "--amg-hierarchy", "cf", "--cf-interp", interp, "--cf-pmax", "8", "--amg-terminal", "100"])
    case = O.case_from_product(h)
    assert len(case.hier["P"]) >= 2
    U, p = _random_state(case, 11)
    _set_state(h, U, p)
    h.assemble(0.5)
    h.pressure_refresh()
    Af, fdiag = O.fine_refresh(case, h.get("rau", case.nv))
    vc = O.VCycle(case.hier, O.Smoother(kind="jacobi", jacobi_omega=0.7))
    vc.set_fine(Af, fdiag)
    r = np.random.default_rng(3).standard_normal(case.nc)
    assert rel(h.precond(r), vc.apply(r)) < 1e-12
    h.close()


def test_cf_hierarchy_converges_to_same_fields(mesh_dir):
    """The preconditioner does not change the converged SIMPLE fields: CF/PMIS
    (ext+i, pmax 8) and SA give the same velocity and pressure to 1e-6."""
    from nodals_amd._lib import H8

    args = ["--mesh", mesh_dir["small"], "--run-mode", "converge", "--max-outer", "3000", "--simple-tol", "1e-9",
            "--amg-smoother", "jacobi", "--pressure-solver", "fgmres", "--p-rtol", "1e-2", "--p-max", "200",
            "--verbose", "0"]
    fields = {}
    for hier in ("sa", "cf"):
        print(f"\n[gpu] SIMPLE converge small pipe amg-hierarchy={hier}", flush=True)
        extra = ["--amg-hierarchy", hier] + (["--cf-interp", "exti", "--cf-pmax", "8"] if hier == "cf" else [])
        h = H8(args + extra)
        r = h.run()
        assert r["converged"] == 1, (hier, r)
        fields[hier] = [h.get(f"u{d}", h.info["nv"]) for d in range(3)] + [h.get("p", h.info["cells"])]
        h.close()
    for a, b in zip(fields["cf"], fields["sa"]):
        assert rel(a, b) < 1e-6
