"""The H8 binary's optional momentum terms on the GPU (--mixing-length,
--supg, --dg-inlet; gpu/serial_cuda/h9_rans_mixlen.cuh,
h7_momentum_assembly.cuh, h10_dg_inlet.cuh):
fixed-10 SIMPLE on the small pipe against the oracle's SIMPLE loop with the
same cell blocks (oracle/rans_oracle.py, pinned in test_rans_oracle.py),
fields to 1e-9, and the mixing-length statistics of the last assembly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {
    "mixing-length": ["--mixing-length", "1", "--mixlen-scale", "1.0"],
    "supg": ["--supg", "1", "--supg-tau-scale", "0.05", "--supg-magic", "9.0"],
    "both": ["--mixing-length", "1", "--mixlen-scale", "0.7", "--supg", "1", "--supg-tau-scale", "0.1",
             "--supg-magic", "4.0"],
    "dg-inlet": ["--dg-inlet", "1"],
    "dg-inlet+mixing-length": ["--dg-inlet", "1", "--mixing-length", "1", "--mixlen-scale", "1.0"],
}


def _flag(args, name, default):
    return float(args[args.index(name) + 1]) if name in args else default


@pytest.mark.parametrize("which", list(CASES))
def test_rans_terms_fixed10_match_oracle(mesh_dir, which):
    from nodals_amd._lib import H8
    from oracle import h8_oracle as O
    from oracle import rans_oracle as R

    extra = CASES[which]
    print(f"\n[gpu] H8 fixed10 with {which}, small pipe, Re 200", flush=True)
    # p-max 200: at Re 200 the solves need up to about 30 PCG iterations
    # (19 without the terms), beyond the default cap of 20
    args = ["--mesh", mesh_dir["small"], "--max-outer", "10", "--amg-smoother", "jacobi", "--pressure-solver", "pcg",
            "--p-max", "200", "--re", "200", "--verbose", "0", *extra]
    h = H8(args)
    r = h.run()
    assert r["pass_"] == 1 and r["outer"] == 10
    case = O.case_from_product(h)
    pts, cx, cy, radius = R.product_geometry(h)
    nu = float(h.get("pipe", 5)[0])
    mixing = (_flag(extra, "--mixlen-scale", 1.0), radius, cx, cy) if "--mixing-length" in extra else None
    supg = (_flag(extra, "--supg-tau-scale", 0.05), _flag(extra, "--supg-magic", 9.0)) if "--supg" in extra else None
    seen = {}

    def keep(it, U, p):
        seen[it] = U.copy()

    blocks = R.extra_blocks_fn(case, pts, nu, mixing, supg) if (mixing or supg) else None
    rhs, U0 = None, None
    if "--dg-inlet" in extra:
        nf = int(h.get("dg_inlet_count", 1, np.int64)[0])
        faces = h.get("dg_inlet_faces", 2 * nf, np.int32).reshape(nf, 2)
        uh = h.get("pipe_geometry", 12)[6:9]
        dg_blocks, rhs = R.dg_inlet_fns(case, pts, faces, nu, uh, mixing)
        U0 = h.get("plug", 3 * case.nv).reshape(case.nv, 3).T.copy()
        if blocks is None:
            blocks = dg_blocks
        else:
            vol = blocks
            blocks = lambda U: vol(U) + dg_blocks(U)  # noqa: E731
    ref = O.simple(case, p_solver="pcg", p_max=200, smoother=O.Smoother(kind="jacobi"), max_outer=10,
                   extra_blocks=blocks, extra_rhs=rhs, U0=U0, callback=keep)
    worst = 0.0
    for d in range(3):
        u = h.get(f"u{d}", case.nv)
        worst = max(worst, np.linalg.norm(u - ref["U"][d]) / np.linalg.norm(ref["U"][d]))
    p = h.get("p", case.nc)
    worst = max(worst, np.linalg.norm(p - ref["p"]) / np.linalg.norm(ref["p"]))
    print(f"[gpu]   max relative L2 difference vs oracle {worst:.3e}", flush=True)
    assert worst < 1e-9
    # the terms change the solution (against the plain run)
    base = O.simple(case, p_solver="pcg", p_max=200, smoother=O.Smoother(kind="jacobi"), max_outer=10, U0=U0)
    assert np.linalg.norm(ref["U"][2] - base["U"][2]) > 1e-6 * np.linalg.norm(base["U"][2])
    if mixing is not None:
        stats = h.get("rans_stats", 3)
        _, want = R.blocks(pts, O.cell_values(case, seen[9]), nu, mixing, None)
        print(f"[gpu]   nu_t/nu mean {stats[0]:.4e} max {stats[1]:.4e}, strain max {stats[2]:.4e}", flush=True)
        assert np.allclose(stats, want, rtol=1e-9, atol=0)
    h.close()

