"""Pins of the mixing-length and SUPG oracle (oracle/rans_oracle.py):
the collapsed 64-point rule is exact for barycentric monomials up to degree
7 and equals the product's table; the P1+BF3 gradients and bubble
Laplacians match finite differences; the SUPG block for uniform advection
(a degree-4 integrand) and the mixing-length block for a uniform strain
equal adaptive cubature of their definitions; the mixing-length block
annihilates constants and is symmetric.  A planted sign error in the bubble
Laplacian fails the SUPG comparison."""
import math
import re
from pathlib import Path

import numpy as np
import pytest
from scipy import integrate

from oracle import rans_oracle as R

ROOT = Path(__file__).resolve().parents[1]


def test_rule_integrates_barycentric_monomials_exactly():
    lam, w = R.collapsed64()
    assert abs(w.sum() - 1 / 6) < 1e-16
    rng = np.random.default_rng(0)
    for _ in range(60):
        e = rng.integers(0, 4, size=4)
        while e.sum() > 7:
            e = rng.integers(0, 4, size=4)
        got = (w * np.prod(lam ** e, axis=1)).sum()
        exact = np.prod([math.factorial(int(k)) for k in e]) / math.factorial(int(e.sum()) + 3)
        assert abs(got - exact) <= 1e-14 * exact, e


def test_product_table_is_the_rule():
    text = (ROOT / "nodals_amd/csrc/device/rans_kernels.hpp").read_text()
    body = text[text.index("inline Collapsed64 collapsed64()"):]
    nums = [float(x) for x in re.findall(r"[-+]?\d+\.\d+(?:e[-+]?\d+)?", body[:body.index("}}") + 2])]
    assert len(nums) == 24
    from scipy.special import roots_jacobi, roots_legendre

    xr, wr = roots_jacobi(4, 2, 0)
    xs, ws = roots_jacobi(4, 1, 0)
    xt, wt = roots_legendre(4)
    want = np.concatenate([(xr + 1) / 2, wr / 8, (xs + 1) / 2, ws / 4, (xt + 1) / 2, wt / 2])
    assert np.allclose(nums, want, rtol=2e-16, atol=0)


def _tet(seed=1, scale=0.2):
    rng = np.random.default_rng(seed)
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * scale
    P = P + 0.2 * scale * rng.uniform(-1, 1, (4, 3)) + np.array([0.1, -0.05, 0.3])
    if np.linalg.det(np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)) < 0:
        P[[1, 2]] = P[[2, 1]]
    return P


def _at(P, GL, x):
    J = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
    xi = np.linalg.solve(J, x - P[0])
    lam = np.array([1 - xi.sum(), *xi])
    return R.basis(lam[None], GL[None])


def test_gradients_and_bubble_laplacians_match_finite_differences():
    P = _tet()
    GL, det, h2 = R.geometry(P[None])
    x0 = P.mean(axis=0) + np.array([0.003, -0.002, 0.001])
    _, g, lap = _at(P, GL[0], x0)
    eps = 1e-5
    fd_g = np.zeros((8, 3))
    fd_lap = np.zeros(8)
    for d in range(3):
        e = np.zeros(3)
        e[d] = eps
        fp = _at(P, GL[0], x0 + e)[0][0]
        fm = _at(P, GL[0], x0 - e)[0][0]
        f0 = _at(P, GL[0], x0)[0][0]
        fd_g[:, d] = (fp - fm) / (2 * eps)
        fd_lap += (fp - 2 * f0 + fm) / eps ** 2
    assert np.allclose(g[0], fd_g, rtol=0, atol=1e-7 * np.abs(fd_g).max())
    assert np.allclose(lap[0, 4:], fd_lap[4:], rtol=0, atol=1e-4 * np.abs(fd_lap).max())
    assert np.all(lap[0, :4] == 0)


def _cubature(P, f):
    """Adaptive cubature of f(lam) over the tetrahedron P."""
    J = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
    det = abs(np.linalg.det(J))
    val, _ = integrate.tplquad(lambda z, y, x: f(np.array([1 - x - y - z, x, y, z])), 0, 1, 0,
                               lambda x: 1 - x, 0, lambda x, y: 1 - x - y, epsabs=1e-14, epsrel=1e-12)
    return det * val


def _supg_entry(P, GL, h2, u, nu, tau_scale, magic, a, b, flip=1.0):
    adv = u[0]  # uniform velocity (all vertices equal, bubbles zero)
    speed2 = adv @ adv
    diff = 4 * nu / h2
    tau = tau_scale / math.sqrt(4 * speed2 / h2 + magic * diff * diff)

    def f(lam):
        _, g, lap = R.basis(lam[None], GL[None])
        st = g[0] @ adv
        return tau * st[a] * (st[b] - flip * nu * lap[0, b])

    return _cubature(P, f)


@pytest.mark.parametrize("ab", [(0, 0), (1, 5), (6, 2), (4, 7), (5, 5)])
def test_supg_block_for_uniform_advection_is_the_integral(ab):
    P = _tet(3)
    GL, det, h2 = R.geometry(P[None])
    u = np.zeros((1, 8, 3))
    u[0, :4] = [0.7, -0.2, 1.1]
    nu, ts, mg = 0.01, 0.05, 9.0
    blk, _ = R.blocks(P[None], u, nu, supg=(ts, mg))
    a, b = ab
    want = _supg_entry(P, GL[0], h2[0], u[0, :4], nu, ts, mg, a, b)
    assert abs(blk[0, a, b] - want) <= 1e-11 * max(abs(want), 1e-6)
    if b >= 4:  # a planted sign error in the bubble Laplacian is seen
        bad = _supg_entry(P, GL[0], h2[0], u[0, :4], nu, ts, mg, a, b, flip=-1.0)
        assert abs(blk[0, a, b] - bad) > 1e-6 * abs(want)


@pytest.mark.parametrize("ab", [(0, 0), (1, 3), (2, 6), (5, 7)])
def test_mixing_length_block_for_uniform_strain_is_the_integral(ab):
    """u = G x (a linear field: vertex values G x_v, bubbles zero) has the
    constant strain sqrt(2 S:S); nu_t then varies only through ell(eta)."""
    P = _tet(5, scale=0.15)
    GL, det, h2 = R.geometry(P[None])
    G = np.array([[0.0, 0.3, 1.2], [0.1, 0.0, -0.4], [2.0, 0.5, 0.0]])
    u = np.zeros((1, 8, 3))
    u[0, :4] = P @ G.T
    S = 0.5 * (G + G.T)
    strain = math.sqrt(2 * (S * S).sum())
    scale, radius, cx, cy = 1.0, 0.5, 0.02, -0.01
    blk, stats = R.blocks(P[None], u, 1e-3, mixing=(scale, radius, cx, cy))
    a, b = ab

    def f(lam):
        x = lam @ P
        eta = min(1.0, math.hypot(x[0] - cx, x[1] - cy) / radius)
        ell = radius * max(0.0, 0.14 - 0.08 * eta ** 2 - 0.06 * eta ** 4)
        _, g, _ = R.basis(lam[None], GL[None])
        return scale * ell * ell * strain * (g[0, a] @ g[0, b])

    want = _cubature(P, f)
    # ell^2 is degree 8 in x, y: an 8-point-per-direction rule (exact to
    # degree 15) integrates the block exactly; the 64-point rule of the
    # terms carries its truncation error
    exact, _ = R.blocks(P[None], u, 1e-3, mixing=(scale, radius, cx, cy), rule=R.collapsed(8))
    assert abs(exact[0, a, b] - want) <= 1e-11 * max(abs(want), 1e-8)
    assert abs(blk[0, a, b] - want) <= 2e-6 * max(abs(want), 1e-8)  # up to degree 12 (bubble pairs)
    # statistics: the volume-weighted mean of nu_t / nu, the max above it
    def ratio(lam):
        x = lam @ P
        eta = min(1.0, math.hypot(x[0] - cx, x[1] - cy) / radius)
        ell = radius * max(0.0, 0.14 - 0.08 * eta ** 2 - 0.06 * eta ** 4)
        return scale * ell * ell * strain / 1e-3

    vol = _cubature(P, lambda lam: 1.0)
    assert abs(stats[0] - _cubature(P, ratio) / vol) <= 1e-9 * stats[0]
    assert stats[1] >= stats[0] and abs(stats[2] - strain) <= 1e-12 * strain


def test_mixing_length_block_symmetric_and_annihilates_constants():
    rng = np.random.default_rng(7)
    P = np.stack([_tet(s, 0.1) for s in range(5)])
    u = rng.standard_normal((5, 8, 3))
    blk, _ = R.blocks(P, u, 1e-3, mixing=(1.0, 0.5, 0.0, 0.0))
    assert np.allclose(blk, np.swapaxes(blk, 1, 2), rtol=0, atol=1e-15 * np.abs(blk).max())
    one = np.array([1, 1, 1, 1, 0, 0, 0, 0.0])
    assert np.abs(blk @ one).max() < 1e-13 * np.abs(blk).max()
    # uniform flow: no strain (to rounding), no eddy viscosity
    uu = np.zeros((5, 8, 3))
    uu[:, :4] = [0.3, 0.1, 1.0]
    assert np.abs(R.blocks(P, uu, 1e-3, mixing=(1.0, 0.5, 0.0, 0.0))[0]).max() < 1e-12 * np.abs(blk).max()


def test_rans_options_are_checked(mesh_dir):
    """The reference binary's checks of the optional terms
    (h8_main_base.cu:1853-1865), on a host-only context."""
    from nodals_amd._lib import H8, NodalsError

    base = ["--mesh", mesh_dir["small"], "--max-outer", "10", "--verbose", "0"]
    for bad, msg in ((["--supg", "1", "--supg-form", "explicit"], "form=implicit"),
                     (["--supg", "1", "--supg-quad-points", "24"], "quad-points=64"),
                     (["--supg", "1", "--supg-tau-scale", "0"], "tau scale"),
                     (["--mixing-length", "1", "--mixlen-scale", "-1"], "mixlen-scale")):
        with pytest.raises(NodalsError, match=msg):
            H8(base + bad, device=False)


def test_triangle_rule_is_degree_six():
    T = R.TRI12
    assert abs(T[:, 3].sum() - 1) < 1e-14
    for e in [(a, b, c) for a in range(7) for b in range(7) for c in range(7) if a + b + c <= 6]:
        got = (T[:, 3] * T[:, 0] ** e[0] * T[:, 1] ** e[1] * T[:, 2] ** e[2]).sum()
        exact = 2 * math.factorial(e[0]) * math.factorial(e[1]) * math.factorial(e[2]) / math.factorial(sum(e) + 2)
        assert abs(got - exact) <= 1e-12 * exact, e


def test_dg_inlet_terms_vanish_at_the_prescribed_plug():
    """u = uh (vertices, zero bubbles) satisfies the weak inlet condition:
    the inflow and Nitsche terms cancel, G U = R."""
    P = np.stack([_tet(s, 0.2) for s in range(6)])
    faces = np.array([[c, c % 4] for c in range(6)])
    uh = np.array([0.2, -0.1, 1.3])
    Uc = np.zeros((6, 8, 3))
    Uc[:, :4] = uh
    for mixing in (None, (1.0, 0.5, 0.0, 0.0)):
        G, Rr = R.dg_inlet(P, faces, Uc, 0.01, uh, mixing)
        assert np.abs(np.einsum("fab,fbd->fad", G, Uc) - Rr).max() < 1e-14 * np.abs(Rr).max()


def test_dg_inlet_entries_are_face_integrals():
    """G and R against adaptive cubature over the face (the integrands have
    degree <= 6, which the 12-point rule integrates exactly); a flipped
    Nitsche sign fails."""
    P = _tet(9, 0.3)
    GL, _, _ = R.geometry(P[None])
    lf = 2
    area, n = R.face_frame(P[None], np.array([lf]))
    V = P[[j for j in range(4) if j != lf]]
    uh = np.array([0.3, 0.1, -0.9])
    beta = max(0.0, -(n[0] @ uh))
    nu = 0.02
    rng = np.random.default_rng(2)
    Uc = rng.standard_normal((1, 8, 3))
    G, Rr = R.dg_inlet(P[None], np.array([[0, lf]]), Uc, nu, uh)

    def face_int(f):
        def g(y, x):
            lam = np.zeros(4)
            fv = [j for j in range(4) if j != lf]
            lam[fv[0]], lam[fv[1]], lam[fv[2]] = 1 - x - y, x, y
            return f(lam)
        val, _ = integrate.dblquad(g, 0, 1, 0, lambda x: 1 - x, epsabs=1e-14, epsrel=1e-12)
        return 2 * area[0] * val

    def at(lam):
        phi, grad, _ = R.basis(lam[None], GL)
        return phi[0], grad[0] @ n[0]

    for a, b in ((0, 5), (6, 1), (4, 4), (7, 6)):
        want = face_int(lambda lam: beta * at(lam)[0][a] * at(lam)[0][b]
                        + nu * (at(lam)[1][a] * at(lam)[0][b] - at(lam)[0][a] * at(lam)[1][b]))
        assert abs(G[0, a, b] - want) <= 1e-11 * max(abs(want), 1e-8)
        bad = face_int(lambda lam: beta * at(lam)[0][a] * at(lam)[0][b]
                       - nu * (at(lam)[1][a] * at(lam)[0][b] - at(lam)[0][a] * at(lam)[1][b]))
        if abs(want - bad) > 1e-8:
            assert abs(G[0, a, b] - bad) > 1e-9
    a = 5
    want = face_int(lambda lam: beta * at(lam)[0][a] + nu * at(lam)[1][a])
    assert np.allclose(Rr[0, a], want * uh, rtol=1e-11, atol=1e-14)
