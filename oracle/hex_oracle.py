"""CPU oracle of the Q1+BF2 hexahedral element blocks.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

Restates the reference's 14-DOF hexahedron of the SST application
(turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/foundation/
hxt2_foundation.inc): the trilinear Q1 basis on [-1, 1]^3 with vertex signs
(hex_q1_ref, :166-176), the six BF2 face bubbles
0.5 (1 -+ s) (1 - t^2)(1 - r^2) on faces z-, z+, y-, x+, y+, x- in that
order (hex_basis_ref, :178-189), the isoparametric Q1 map and its inverse
Jacobian (hex_metric, inv3_dev, :191-196, :159-166), the physical gradients
(phys_grad14, :198-200) and the 3 x 3 x 3 Gauss rule (c_g3_x / c_g3_w,
:52-58).  Element blocks, per hexahedron:

  diffusion  D_ab = sum_q nu det_q w_q grad(phi_a) . grad(phi_b)
             (hex_volume_kernel, :211-225, as a matrix);
  convection C_ab = sum_q det_q w_q phi_a (u_q . grad(phi_b)),
             u_q = sum_m u_m phi_m over all 14 DOFs
             (hex_conv_assemble_cached, modules/foundation/
             hxt4b_foundation.inc:246-345).

``gauss`` selects the rule; n > 3 is used by the pins as the exact
integral on affine hexahedra.
"""
from __future__ import annotations

import numpy as np

SX = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
SY = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
SZ = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)


def basis(x, y, z):
    """(phi[14], grad[14][3]) on the reference cube at one point."""
    phi = np.empty(14)
    gr = np.empty((14, 3))
    ax, ay, az = 1 + SX * x, 1 + SY * y, 1 + SZ * z
    phi[:8] = 0.125 * ax * ay * az
    gr[:8, 0] = 0.125 * SX * ay * az
    gr[:8, 1] = 0.125 * SY * ax * az
    gr[:8, 2] = 0.125 * SZ * ax * ay
    bx, by, bz = 1 - x * x, 1 - y * y, 1 - z * z
    dx, dy, dz = -2 * x, -2 * y, -2 * z
    # (sign, normal axis, the two tangential axes' bubble factors)
    faces = [(-1, 2), (1, 2), (-1, 1), (1, 0), (1, 1), (-1, 0)]
    b = (bx, by, bz)
    db = (dx, dy, dz)
    xyz = (x, y, z)
    for f, (sgn, ax_) in enumerate(faces):
        s = 0.5 * (1 + sgn * xyz[ax_])
        t = [i for i in range(3) if i != ax_]
        phi[8 + f] = s * b[t[0]] * b[t[1]]
        g = np.zeros(3)
        g[ax_] = 0.5 * sgn * b[t[0]] * b[t[1]]
        g[t[0]] = s * db[t[0]] * b[t[1]]
        g[t[1]] = s * b[t[0]] * db[t[1]]
        gr[8 + f] = g
    return phi, gr


def rule(n=3):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def element_blocks(coords, u, nu, gauss=3):
    """coords (8,3), u (14,3) -> (D (14,14), C (14,14))."""
    coords = np.asarray(coords, dtype=float)
    u = np.asarray(u, dtype=float)
    xg, wg = rule(gauss)
    D = np.zeros((14, 14))
    C = np.zeros((14, 14))
    for ix in range(gauss):
        for iy in range(gauss):
            for iz in range(gauss):
                x, y, z = xg[ix], xg[iy], xg[iz]
                w = wg[ix] * wg[iy] * wg[iz]
                phi, gr = basis(x, y, z)
                J = coords.T @ gr[:8]                      # J[i][j] = sum_a x_a,i dphi_a/dxi_j
                det = np.linalg.det(J)
                if not det > 0:
                    raise ValueError("non-positive hexahedron Jacobian")
                inv = np.linalg.inv(J)                     # inv[j][d] = dxi_j / dx_d
                gp = gr @ inv                              # gp[a][d] = sum_j gr[a][j] inv[j][d]
                D += nu * det * w * (gp @ gp.T)
                uq = phi @ u                               # (3,)
                C += det * w * np.outer(phi, gp @ uq)
    return D, C


def batch(coords, u, nu, gauss=3):
    n = coords.shape[0]
    D = np.empty((n, 14, 14))
    C = np.empty((n, 14, 14))
    for c in range(n):
        D[c], C[c] = element_blocks(coords[c], u[c], nu, gauss)
    return D, C


def random_hexes(n, seed=0, affine=False, scale=0.15):
    """Unit cubes at random offsets, stretched/sheared by a random matrix,
    with (unless affine) random vertex perturbations."""
    rng = np.random.default_rng(seed)
    ref = np.stack([SX, SY, SZ], axis=1) * 0.5
    out = np.empty((n, 8, 3))
    for c in range(n):
        A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        while np.linalg.det(A) < 0.3:
            A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        p = ref @ A.T + rng.standard_normal(3)
        if not affine:
            p = p + scale * 0.5 * rng.uniform(-1, 1, (8, 3)) * np.abs(np.linalg.det(A)) ** (1 / 3)
        out[c] = p
    return out
