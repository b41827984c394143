"""GPU parity of the H8 options beyond the Jacobi/PCG defaults, and of
ten SIMPLE iterations at the pipe and cavity benchmark sizes.

* One AMG preconditioner application with the cheb2 smoother (the
  reference binary's default, h8_main_base.cu:417), L1-Jacobi and MCGS, on
  SA and CF/PMIS hierarchies, against the oracle V-cycle (pinned by the
  error-propagation formula in test_oracle_pins.py): 1e-12 relative.
* Ten SIMPLE iterations with those smoothers, with the outer Chebyshev
  pressure iteration (pressure_cheb_poly, h8_main_base.cu:670-726) and with
  the alternating momentum sweep altgs1 (h8_main_base.cu:1700-1715),
  against the oracle's SIMPLE loop: 1e-9 relative.
* Ten SIMPLE iterations on the 40,320-tet O-grid pipe (224-quad section,
  30 slabs) and on the 16^3 Kuhn cavity (24,576 tets), each against the
  oracle's loop, with the field errors printed: 1e-9 relative.
"""
import numpy as np
import pytest

from oracle import h8_oracle as O

pytestmark = pytest.mark.gpu

# p-max 200: with p-max 20 several of these smoothers miss the 0.5 target in
# the first iterations, and the solver then stops as the reference does
# (h8_main_base.cu:2101), so the cap is raised for oracle and GPU alike.
BASE = ["--max-outer", "10", "--alpha-u", "0.5", "--alpha-p", "0.5", "--p-rtol", "0.5", "--p-max", "200",
        "--verbose", "0"]
P_MAX = 200


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _smoother(kind, **kw):
    return O.Smoother(kind=kind, **kw)


PRECOND_CASES = [
    ("sa", "cheb2", [], _smoother("cheb2")),
    ("sa", "cheb2", ["--amg-cheb-degree", "3", "--amg-lambda-low-fraction", "0.1"],
     _smoother("cheb2", cheb_degree=3, low_fraction=0.1)),
    ("sa", "l1jacobi", ["--amg-jacobi-omega", "0.8", "--amg-jacobi-fine-sweeps", "2"],
     _smoother("l1jacobi", jacobi_omega=0.8, jacobi_fine_sweeps=2)),
    ("sa", "l1jacobi", [], _smoother("l1jacobi")),  # one sweep: the fused V-cycle kernels
    ("sa", "mcgs", [], _smoother("mcgs")),
    ("cf", "cheb2", ["--cf-interp", "exti", "--cf-pmax", "8"], _smoother("cheb2")),
    ("cf", "mcgs", ["--cf-interp", "exti", "--cf-pmax", "8", "--amg-mcgs-order", "forward", "--amg-mcgs-omega", "0.9"],
     _smoother("mcgs", mcgs_order="forward", mcgs_omega=0.9)),
]


@pytest.mark.parametrize("hier,kind,extra,opt", PRECOND_CASES,
                         ids=[f"{c[0]}-{c[1]}-{i}" for i, c in enumerate(PRECOND_CASES)])
def test_precond_smoother_matches_oracle(mesh_dir, hier, kind, extra, opt):
    from nodals_amd._lib import H8

    print(f"\n[gpu] precond small pipe hierarchy={hier} smoother={kind} {' '.join(extra)}", flush=True)
    h = H8(["--mesh", mesh_dir["small"], *BASE, "--amg-hierarchy", hier, "--amg-smoother", kind,
            "--amg-terminal", "100", "--pressure-solver", "fgmres", *extra])
    case = O.case_from_product(h)
    rng = np.random.default_rng(17)
    U = rng.standard_normal((3, case.nv)) * 0.3
    for d in range(3):
        h.set(f"u{d}", U[d])
    h.assemble(0.5)
    h.pressure_refresh()
    Af, fdiag = O.fine_refresh(case, h.get("rau", case.nv))
    vc = O.VCycle(case.hier, opt)
    vc.set_fine(Af, fdiag)
    for k in range(2):
        r = rng.standard_normal(case.nc)
        assert rel(h.precond(r), vc.apply(r)) < 1e-12, k
    h.close()


FIXED10_CASES = [
    # (label, native flags, oracle simple() keywords)
    ("cheb2+pcg (reference binary defaults)", ["--amg-smoother", "cheb2", "--pressure-solver", "pcg"],
     dict(p_solver="pcg", smoother=_smoother("cheb2"))),
    # L1-Jacobi inside FGMRES: inside PCG (17-29 iterations per solve here)
    # the loop amplifies rounding so much that a 1e-13 perturbation of the
    # start state moves the oracle's own ten-iteration fields by 6e-6, which
    # no cross-device comparison can resolve; inside FGMRES the same
    # perturbation moves them by 4e-10.
    ("l1jacobi+fgmres", ["--amg-smoother", "l1jacobi", "--pressure-solver", "fgmres"],
     dict(p_solver="fgmres", smoother=_smoother("l1jacobi"))),
    ("jacobi+outer cheb", ["--amg-smoother", "jacobi", "--pressure-solver", "cheb"],
     dict(p_solver="cheb", smoother=_smoother("jacobi"))),
    ("cheb2+outer cheb degree 4", ["--amg-smoother", "cheb2", "--pressure-solver", "cheb", "--pressure-cheb-degree", "4",
                                   "--pressure-power-its", "10", "--pressure-cheb-low-fraction", "0.1"],
     dict(p_solver="cheb", smoother=_smoother("cheb2"), cheb_degree=4, cheb_power_its=10, cheb_low_fraction=0.1)),
    ("mcgs+fgmres", ["--amg-smoother", "mcgs", "--pressure-solver", "fgmres"],
     dict(p_solver="fgmres", smoother=_smoother("mcgs"))),
    ("altgs1 momentum", ["--amg-smoother", "jacobi", "--pressure-solver", "pcg", "--momentum-work", "altgs1",
                         "--mom-omega", "0.9"],
     dict(p_solver="pcg", smoother=_smoother("jacobi"), momentum_work="altgs1", mom_omega=0.9)),
]


@pytest.mark.parametrize("label,flags,kw", FIXED10_CASES, ids=[c[0].split(" ")[0] for c in FIXED10_CASES])
def test_simple_fixed10_options_match_oracle(mesh_dir, label, flags, kw):
    from nodals_amd._lib import H8

    print(f"\n[gpu] SIMPLE fixed10 small pipe {label}", flush=True)
    h = H8(["--mesh", mesh_dir["small"], *BASE, *flags])
    case = O.case_from_product(h)
    r = h.run()
    assert r["pass_"] == 1 and r["outer"] == 10
    ref = O.simple(case, max_outer=10, p_max=P_MAX, **kw)
    for d in range(3):
        assert rel(h.get(f"u{d}", case.nv), ref["U"][d]) < 1e-9, (label, d)
    assert rel(h.get("p", case.nc), ref["p"]) < 1e-9, label
    assert r["avg_pressure_its"] == pytest.approx(np.mean([x["p_its"] for x in ref["hist"]]), abs=0.11)
    h.close()


@pytest.fixture(scope="module")
def bench_meshes(tmp_path_factory):
    from nodals_amd import meshgen

    d = tmp_path_factory.mktemp("benchsize")
    return {"pipe40k": str(meshgen.write_pipe_case(d / "pipe40k", core=8, ring=5, nslab=30)),
            "cavity16": str(meshgen.write_cavity_case(d / "cavity16", n=16))}


@pytest.mark.parametrize("which", ["pipe40k", "cavity16"])
def test_simple_fixed10_benchmark_sizes_match_oracle(bench_meshes, which):
    from nodals_amd._lib import H8

    if which == "pipe40k":
        flags = ["--amg-smoother", "cheb2", "--pressure-solver", "pcg"]
        kw = dict(p_solver="pcg", smoother=_smoother("cheb2"))
        ncells = 40320
    else:
        flags = ["--lid", "lid", "--wall", "walls", "--re", "100", "--amg-smoother", "jacobi",
                 "--pressure-solver", "fgmres"]
        kw = dict(p_solver="fgmres", smoother=_smoother("jacobi"))
        ncells = 24576
    print(f"\n[gpu] SIMPLE fixed10 {which} ({ncells} tets) {' '.join(flags)}", flush=True)
    h = H8(["--mesh", bench_meshes[which], *BASE, *flags])
    assert h.info["cells"] == ncells
    case = O.case_from_product(h)
    r = h.run()
    assert r["pass_"] == 1 and r["outer"] == 10
    ref = O.simple(case, max_outer=10, p_max=P_MAX, **kw)
    errs = {f"u{d}": rel(h.get(f"u{d}", case.nv), ref["U"][d]) for d in range(3)}
    errs["p"] = rel(h.get("p", case.nc), ref["p"])
    print(f"[gpu] {which} fixed10 relative L2 vs oracle: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()),
          flush=True)
    assert max(errs.values()) < 1e-9, errs
    h.close()
