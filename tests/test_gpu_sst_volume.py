"""k-omega SST volume terms on the GPU (one wavefront per cell, Keast
24-point rule) against the oracle, which evaluates the SST-2003 local state
with the reference's own compiled core: element matrices and vectors agree
to 1e-10 relative (device transcendentals and FMA contraction)."""
import numpy as np
import pytest

from oracle import refbind
from oracle import sst_oracle as S

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refbind.sst_available(), reason="reference SST cores not built")]


@pytest.mark.parametrize("equation", [0, 1])
def test_sst_volume_matches_oracle(mesh_dir, equation):
    from nodals_amd._lib import H8
    from oracle import h8_oracle as O

    print(f"\n[gpu] SST volume terms small pipe equation={equation}", flush=True)
    h = H8(["--mesh", mesh_dir["small"], "--max-outer", "10", "--verbose", "0"])
    case = O.case_from_product(h, with_hierarchy=False)
    rng = np.random.default_rng(10 + equation)
    nv, nf = case.nv, h.info["nfixed"]
    U = rng.standard_normal((3, nv)) * 0.5
    for d in range(3):
        h.set(f"u{d}", U[d])
    k = 10.0 ** rng.uniform(-4, -1, nv + nf)
    w = 10.0 ** rng.uniform(0, 3, nv + nf)
    mat, rhs = h.sst_volume(equation, k[:nv], k[nv:], w[:nv], w[nv:])
    xy = h.get("cell_xy", case.nc * 8).reshape(case.nc, 4, 2)
    pipe = h.get("pipe", 5)
    mo, ro, _ = S.volume_terms(refbind.RefSst(), case.cell_ref, case.cell_invJ, case.cell_det, xy, U,
                               case.fixed_value, k[:nv], k[nv:], w[:nv], w[nv:], float(pipe[3]), float(pipe[0]),
                               equation)
    assert np.abs(mat - mo).max() <= 1e-10 * np.abs(mo).max()
    assert np.abs(rhs - ro).max() <= 1e-10 * np.abs(ro).max()
    h.close()
