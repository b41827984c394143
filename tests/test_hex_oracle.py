"""Pins of the Q1+BF2 hexahedron oracle (oracle/hex_oracle.py), on the CPU:

* the Q1 functions are a partition of unity and Kronecker at the vertices;
  each BF2 bubble is 1 at its own face centre, 0 at the other five and 1/2
  at the cube centre (the reference's bf2_audit, tools/prepare_hxt1.py);
* every basis gradient equals a central difference of the basis (so a
  wrong sign or factor in a derivative fails);
* on affine hexahedra the 3-point Gauss rule equals a 6-point rule for the
  diffusion block and, with a Q1 velocity, for the convection block (the
  integrands are polynomials of degree <= 5 per variable there), so the
  rule's points and weights are pinned;
* on curved hexahedra both blocks annihilate a constant field, and a linear
  field (nodal values c . x_a, zero bubbles) has diffusion energy
  nu |c|^2 volume;
* the convection block of a constant carrier a applied to a linear field is
  (a . c) times the integral of each basis function.
"""
import numpy as np

from oracle import hex_oracle as X


def test_basis_values():
    rng = np.random.default_rng(1)
    for p in rng.uniform(-1, 1, (20, 3)):
        phi, _ = X.basis(*p)
        assert abs(phi[:8].sum() - 1.0) < 1e-14
    corners = np.stack([X.SX, X.SY, X.SZ], axis=1)
    for a, v in enumerate(corners):
        phi, _ = X.basis(*v)
        assert np.allclose(phi[:8], np.eye(8)[a], atol=1e-15)
    centres = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0)]
    M = np.array([X.basis(*c)[0][8:] for c in centres])
    assert np.abs(M - np.eye(6)).max() < 1e-15
    assert np.abs(X.basis(0, 0, 0)[0][8:] - 0.5).max() < 1e-15


def test_gradients_are_derivatives():
    rng = np.random.default_rng(2)
    h = 1e-6
    for p in rng.uniform(-0.9, 0.9, (10, 3)):
        _, gr = X.basis(*p)
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            fd = (X.basis(*(p + e))[0] - X.basis(*(p - e))[0]) / (2 * h)
            assert np.abs(fd - gr[:, j]).max() < 1e-8


def test_three_point_rule_is_exact_on_affine_hexahedra():
    hexes = X.random_hexes(4, seed=3, affine=True)
    rng = np.random.default_rng(4)
    for c in range(4):
        u = np.zeros((14, 3))
        u[:8] = rng.standard_normal((8, 3))            # Q1 velocity
        D3, C3 = X.element_blocks(hexes[c], u, 0.7, gauss=3)
        D6, C6 = X.element_blocks(hexes[c], u, 0.7, gauss=6)
        assert np.abs(D3 - D6).max() < 1e-12 * np.abs(D6).max()
        assert np.abs(C3 - C6).max() < 1e-12 * np.abs(C6).max()
    # a planted wrong weight (two-point rule) is not exact
    D2, _ = X.element_blocks(hexes[0], np.zeros((14, 3)), 0.7, gauss=2)
    D6, _ = X.element_blocks(hexes[0], np.zeros((14, 3)), 0.7, gauss=6)
    assert np.abs(D2 - D6).max() > 1e-6 * np.abs(D6).max()


def test_constant_and_linear_fields_on_curved_hexahedra():
    hexes = X.random_hexes(6, seed=5)
    rng = np.random.default_rng(6)
    nu = 0.3
    for c in range(6):
        P = hexes[c]
        u = rng.standard_normal((14, 3))
        D, C = X.element_blocks(P, u, nu)
        one = np.r_[np.ones(8), np.zeros(6)]
        assert np.abs(D @ one).max() < 1e-12 * np.abs(D).max()
        assert np.abs(C @ one).max() < 1e-12 * np.abs(C).max()
        cvec = rng.standard_normal(3)
        lin = np.r_[P @ cvec, np.zeros(6)]
        xg, wg = X.rule(3)
        vol = 0.0
        for ix in range(3):
            for iy in range(3):
                for iz in range(3):
                    _, gr = X.basis(xg[ix], xg[iy], xg[iz])
                    vol += np.linalg.det(P.T @ gr[:8]) * wg[ix] * wg[iy] * wg[iz]
        assert abs(lin @ D @ lin - nu * (cvec @ cvec) * vol) < 1e-12 * nu * (cvec @ cvec) * vol


def test_constant_carrier_convection_of_a_linear_field():
    hexes = X.random_hexes(3, seed=7, affine=True)
    rng = np.random.default_rng(8)
    for c in range(3):
        P = hexes[c]
        a = rng.standard_normal(3)
        u = np.zeros((14, 3))
        u[:8] = a
        cvec = rng.standard_normal(3)
        lin = np.r_[P @ cvec, np.zeros(6)]
        _, C = X.element_blocks(P, u, 1.0)
        # integral of each basis function with the same rule
        xg, wg = X.rule(3)
        ints = np.zeros(14)
        for ix in range(3):
            for iy in range(3):
                for iz in range(3):
                    phi, gr = X.basis(xg[ix], xg[iy], xg[iz])
                    ints += phi * np.linalg.det(P.T @ gr[:8]) * wg[ix] * wg[iy] * wg[iz]
        assert np.abs(C @ lin - (a @ cvec) * ints).max() < 1e-12 * np.abs(ints).max() * abs(a @ cvec) + 1e-15
