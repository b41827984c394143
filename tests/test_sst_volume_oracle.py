"""Pins of the SST volume oracle: quadrature exactness, the basis, and the
constant null space of the diffusion part."""
import itertools

import numpy as np
import pytest

from oracle import sst_oracle as S


def test_keast24_integrates_degree6_monomials_exactly():
    L, W = S.keast24()
    assert L.shape == (24, 4) and np.allclose(L.sum(axis=1), 1.0, atol=1e-15)
    assert abs(W.sum() - 1.0 / 6.0) < 1e-15
    for deg in range(7):
        for e in itertools.product(range(deg + 1), repeat=4):
            if sum(e) != deg:
                continue
            q = float(np.sum(W * np.prod(L ** np.array(e), axis=1)))
            assert abs(q - S.monomial_exact(e)) <= 1e-15 + 1e-13 * S.monomial_exact(e), e


def test_basis_and_gradients():
    rng = np.random.default_rng(1)
    l = rng.dirichlet(np.ones(4), 50)
    phi, gr = S.basis(l)
    assert np.allclose(phi[:, :4].sum(axis=1), 1.0)
    # bubble i vanishes on the face opposite... no: it vanishes when any of its
    # three barycentrics is zero, and is 1 at the centroid of face i
    c = np.array([[0.0, 1 / 3, 1 / 3, 1 / 3]])
    assert np.isclose(S.basis(c)[0][0, 4], 1.0)
    # gradients w.r.t. (l1, l2, l3) with l0 = 1 - l1 - l2 - l3, by central differences
    h = 1e-6
    for d in range(3):
        lp, lm = l.copy(), l.copy()
        lp[:, 1 + d] += h
        lp[:, 0] -= h
        lm[:, 1 + d] -= h
        lm[:, 0] += h
        fd = (S.basis(lp)[0] - S.basis(lm)[0]) / (2 * h)
        assert np.allclose(fd, gr[:, :, d], atol=1e-8)


def _random_fields(h, seed=3):
    rng = np.random.default_rng(seed)
    nv, nf = h.info["nv"], h.info["nfixed"]
    U = rng.standard_normal((3, nv)) * 0.5
    k = 10.0 ** rng.uniform(-4, -1, nv + nf)
    w = 10.0 ** rng.uniform(0, 3, nv + nf)
    return U, k[:nv], k[nv:], w[:nv], w[nv:]


@pytest.mark.skipif(not __import__("oracle.refbind", fromlist=["x"]).sst_available(),
                    reason="reference SST cores not built")
def test_oracle_diffusion_annihilates_constants(mesh_dir):
    from nodals_amd._lib import H8
    from oracle import h8_oracle as O
    from oracle import refbind

    h = H8(["--mesh", mesh_dir["small"], "--max-outer", "10"], device=False)
    case = O.case_from_product(h, with_hierarchy=False)
    U, kf, kx, wf, wx = _random_fields(h)
    xy = h.get("cell_xy", case.nc * 8).reshape(case.nc, 4, 2)
    radius = float(h.get("pipe", 5)[3])
    mat, rhs, D = S.volume_terms(refbind.RefSst(), case.cell_ref, case.cell_invJ, case.cell_det, xy, U,
                                 case.fixed_value, kf, kx, wf, wx, radius, float(h.get("pipe", 5)[0]), 0)
    _, gr = S.basis(S.keast24()[0])
    gp = np.einsum("qaj,cjd->cqad", gr, case.cell_invJ)
    diff = np.einsum("cq,cqad,cqbd->cab", D, gp, gp)
    assert np.abs(diff[:, :, :4].sum(axis=2)).max() <= 1e-12 * np.abs(diff).max()
    assert np.allclose(mat, np.transpose(mat, (0, 2, 1)), rtol=0, atol=1e-12 * np.abs(mat).max())
    h.close()
