"""Wall-function momentum path on the GPU against the wall oracle
(oracle/wall_oracle.py) on a small helix (4x4 core, 2 rings, 48 slabs:
13,824 tets, 1,536 wall faces) at Re = 5000:

* assembly at a perturbed plug state: the face fit (beta, u_tau, y+) against
  the reference's own compiled tangential fit, the effective viscosity at
  the wall points, the 3x3 block operator K = T^T M T entry by entry, the
  right-hand side T^T b and the component rAU, gamma-free rAU and delta;
* the twenty block Gauss-Seidel sweeps from that assembly;
* the constrained Schur operator and the velocity correction;
* ten SIMPLE iterations of the wall-function loop (k and omega included).
The runs start from the potential-flow start (pinned against the oracle on
the CPU).  The oracle fits with the product's own device fit
(nodals_wall_fit_tangential) on its own samples; that fit agrees with the
reference's compiled fit to the resolution of the reference's
golden-section fallback (a few 1e-4), which the first test reports and
bounds.  Tolerances are in the tests: 1e-11 relative for single stages,
1e-8 for the ten-iteration trajectory."""
import numpy as np
import pytest

from oracle import h8_oracle as H
from oracle import refbind
from oracle import turb_oracle as TO
from oracle import wall_oracle as WO

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refbind.sst_available(), reason="reference SST cores not built (oracle/_ref)")]

ARGS = ["--duct", "1", "--wall-model", "spalding", "--re", "5000", "--alpha-u", "0.7", "--alpha-p", "0.3",
        "--amg-hierarchy", "cf", "--cf-interp", "exti", "--amg-smoother", "mcgs", "--pressure-solver", "pcg",
        "--p-rtol", "1e-3", "--p-max", "600", "--verbose", "0"]


class DeviceFit:
    """The product's tangential fit on the GPU, with refbind.RefSst's
    fit_tangential signature."""

    @staticmethod
    def fit_tangential(samples, y, nrm, pic, nu, count, zero_limit):
        from nodals_amd import _lib

        return _lib.wall_fit_tangential(samples, y, nrm, pic, nu, count, zero_limit)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def helix(tmp_path_factory):
    from nodals_amd import meshgen

    return str(meshgen.write_helix_case(tmp_path_factory.mktemp("whelix") / "helix", core=4, ring=2, nslab=48))


def _setup(helix, extra=(), turb=(), params=None):
    from nodals_amd._lib import H8

    h = H8(["--mesh", helix, "--max-outer", "10", *ARGS, *extra])
    h.turb_init(list(turb))
    h.wall_init([])
    case = H.case_from_product(h)
    m = TO.mesh_from_product(h)
    # the omega wall fit takes the wall path's active point count (1 by default)
    st = TO.Step(m, case, h.get("pipe_geometry", 12), params or TO.Params(active_points=1), refbind.RefSst(),
                 wall_distance=h.get("wall_distance", m.nvert))
    wp = WO.WallPath(case, st, refbind.RefSst(), fitter=DeviceFit())
    return h, case, st, wp


def _perturbed_state(h, case, st, wp, seed=3):
    rng = np.random.default_rng(seed)
    U = h.get("wall_u0", 3 * case.nv).reshape(case.nv, 3).T + 0.05 * rng.standard_normal((3, case.nv))
    p = 0.1 * rng.standard_normal(case.nc)
    for d in range(3):
        h.set(f"u{d}", U[d])
    h.set("p", p)
    ns = st.m.ns
    st.k = st.k * (1.0 + 0.2 * rng.random(ns))
    st.w = st.w * (1.0 + 0.2 * rng.random(ns))
    h.set("k", st.k)
    h.set("omega", st.w)
    return U, p


def _k_entries(K, case):
    r = np.repeat(np.arange(case.nv), np.diff(case.mom_row))
    out = np.zeros((case.nnz, 9))
    for i in range(3):
        for j in range(3):
            out[:, 3 * i + j] = np.asarray(K[3 * r + i, 3 * case.mom_col + j]).ravel()
    return out


def test_wall_assembly_matches_oracle(helix):
    print("\n[gpu] wall assembly, small helix Re=5000, perturbed potential start", flush=True)
    h, case, st, wp = _setup(helix)
    U, p = _perturbed_state(h, case, st, wp)
    h.wall_assemble()
    nf = wp.W.nf
    fit = h.get("wall_fit", 3 * nf).reshape(nf, 3)
    Uc = wp.constrain(U)
    dev_fit, valid = wp.face_fit(Uc)
    wp.fitter = refbind.RefSst()
    ref_fit, rvalid = wp.face_fit(Uc)
    wp.fitter = DeviceFit()
    assert valid.all() and rvalid.all() and (fit[:, 0] > 0).all()
    e_dev, e_ref = rel(fit, dev_fit), rel(fit[:, :2], ref_fit[:, :2])
    print(f"[gpu]   fit: kernel vs device fit of the oracle's samples {e_dev:.2e}; vs the reference's compiled fit: "
          f"beta {rel(fit[:, 0], ref_fit[:, 0]):.2e} u_tau {rel(fit[:, 1], ref_fit[:, 1]):.2e}", flush=True)
    A = wp.assemble(U, p, fit=fit)
    nue = h.get("wall_nuq", 12 * nf).reshape(nf, 12)
    assert rel(nue, A["nue"]) < 1e-11
    Kg = h.get("wall_K", 9 * case.nnz).reshape(case.nnz, 9)
    Ko = _k_entries(A["K"], case)
    errs = {"K": rel(Kg, Ko),
            "rhs": rel(h.get("wall_rhs", 3 * case.nv), A["b"]),
            "rau3": rel(h.get("wall_rau3", 3 * case.nv).reshape(-1, 3), A["rau3"]),
            "r3": rel(h.get("wall_r3", 3 * case.nv).reshape(-1, 3), A["r3"]),
            "delta3": rel(h.get("wall_delta3", 3 * case.nv).reshape(-1, 3), A["delta3"])}
    print("[gpu]   " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert e_dev < 1e-9 and e_ref < 2e-3
    assert max(errs.values()) < 1e-11, errs
    # the constrained velocity the assembly ran at
    for d in range(3):
        assert rel(h.get(f"u{d}", case.nv), A["U"][d]) < 1e-13
    h.wall_momentum()
    Un = wp.momentum(A)
    e = max(rel(h.get(f"u{d}", case.nv), Un[d]) for d in range(3))
    print(f"[gpu]   after 20 block sweeps: u rel {e:.2e}", flush=True)
    assert e < 1e-11
    h.close()


def test_wall_schur_and_correction_match_oracle(helix):
    print("\n[gpu] wall constrained Schur operator and correction, small helix", flush=True)
    h, case, st, wp = _setup(helix)
    U, p = _perturbed_state(h, case, st, wp, seed=5)
    h.wall_assemble()
    nf = wp.W.nf
    A = wp.assemble(U, p, fit=h.get("wall_fit", 3 * nf).reshape(nf, 3))
    e_r3 = rel(h.get("wall_r3", 3 * case.nv).reshape(-1, 3), A["r3"])
    C, S = wp.schur_parts(A["r3"])
    rng = np.random.default_rng(11)
    xs = [rng.standard_normal(case.nc) for _ in range(2)]
    e_s = [rel(h.wall_schur(x), S @ x) for x in xs]
    print(f"[gpu]   r3 {e_r3:.2e} Schur {e_s[0]:.2e} {e_s[1]:.2e}", flush=True)
    assert e_r3 < 1e-11 and max(e_s) < 1e-12
    u0 = np.array([h.get(f"u{d}", case.nv) for d in range(3)])
    dp = rng.standard_normal(case.nc)
    h.wall_correct(dp)
    Uc = u0 + (C @ (wp.B.T @ dp)).reshape(case.nv, 3).T
    assert max(rel(h.get(f"u{d}", case.nv), Uc[d]) for d in range(3)) < 1e-12
    # the correction keeps every wall face's mean normal flux at zero
    Ucell = H.cell_values(case, np.array([h.get(f"u{d}", case.nv) for d in range(3)]))
    flux = np.einsum("fq,fqb,fbd,fd->f", wp.W.qw, wp.W.bw, Ucell[wp.W.cell], wp.W.n)
    assert np.abs(flux).max() < 1e-12 * max(1.0, np.abs(Uc).max() * wp.W.area.max())
    h.close()


@pytest.mark.timeout(400)
@pytest.mark.parametrize("target", ["sample_log", "trace_log"])
def test_wall_simple_fixed10_matches_oracle(helix, target):
    print(f"\n[gpu] wall-function SIMPLE fixed10, small helix Re=5000 (k, omega coupled, omega {target})", flush=True)
    h, case, st, wp = _setup(helix, turb=["--sst-omega-wall-target", target],
                             params=TO.Params(active_points=1, omega_target=target))
    r = h.run()
    assert r["outer"] == 10 and r["pass_"] == 1
    hist = []
    U0 = h.get("wall_u0", 3 * case.nv).reshape(case.nv, 3).T.copy()
    U, p = wp.simple(U0, np.zeros(case.nc), 10, alpha_p=0.3, p_rtol=1e-3, p_atol=1e-12, p_max=600,
                     smoother=H.Smoother(kind="mcgs"), hist=hist,
                     log=lambda it, its: print(f"[gpu]   oracle iteration {it}: {its} PCG iterations", flush=True))
    ns = st.m.ns
    errs = [rel(h.get(f"u{d}", case.nv), U[d]) for d in range(3)] + [rel(h.get("p", case.nc), p),
                                                                    rel(h.get("k", ns), st.k),
                                                                    rel(h.get("omega", ns), st.w)]
    print(f"[gpu]   rel errors u0 u1 u2 p k omega: {['%.1e' % e for e in errs]} "
          f"pressure its GPU {r['avg_pressure_its']:.2f} oracle {np.mean([x['p_its'] for x in hist]):.2f}", flush=True)
    assert max(errs) < 1e-8
    h.close()
