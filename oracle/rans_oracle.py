"""CPU oracle of the H8 binary's optional momentum terms: the Nikuradse
mixing-length eddy viscosity and the implicit SUPG stabilisation.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

Restates gpu/serial_cuda/h9_rans_mixlen.cuh:8-16, 51-190 (mixing length:
nu_t = scale ell^2 sqrt(2 S:S), ell / R = max(0, 0.14 - 0.08 eta^2 - 0.06
eta^4), block int nu_t grad(phi_a).grad(phi_b)) and
gpu/serial_cuda/h7_momentum_assembly.cuh:101-251 (SUPG: tau = tauScale /
sqrt(4 |a|^2 / h^2 + magic (4 nu / h^2)^2), block sum_q tau w (a.grad
phi_a)(a.grad phi_b - nu lap phi_b)), both with the lagged velocity and the
64-point collapsed rule of the tetrahedron (h7_momentum_assembly.cuh:88-93,
the quadrature loops at :141-149), assembled like the convection
(h8_main_base.cu:1137-1148: added to the momentum matrix, Dirichlet columns
to the right-hand side).  The rule here comes from scipy's Gauss-Jacobi
and Gauss-Legendre roots.
"""
from __future__ import annotations

import numpy as np
from scipy.special import roots_jacobi, roots_legendre


def collapsed(n=4):
    """(lam (n^3,4), w (n^3,)): the collapsed rule with n points per
    direction (exact to degree 2n - 1): lambda = ((1-r)(1-s)(1-t), r,
    (1-r)s, (1-r)(1-s)t), weights summing to 1/6, point index
    n^2 ir + n is + it."""
    xr, wr = roots_jacobi(n, 2, 0)
    xs, ws = roots_jacobi(n, 1, 0)
    xt, wt = roots_legendre(n)
    r, s, t = (xr + 1) / 2, (xs + 1) / 2, (xt + 1) / 2
    wr, ws, wt = wr / 8, ws / 4, wt / 2
    lam, w = [], []
    for ir in range(n):
        for is_ in range(n):
            for it in range(n):
                R, S, T = r[ir], s[is_], t[it]
                lam.append([(1 - R) * (1 - S) * (1 - T), R, (1 - R) * S, (1 - R) * (1 - S) * T])
                w.append(wr[ir] * ws[is_] * wt[it])
    return np.array(lam), np.array(w)


def collapsed64():
    """The 64-point rule of both terms."""
    return collapsed(4)


_OPP = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]


def basis(lam, GL):
    """P1+BF3 values (...,8), physical gradients (...,8,3) and Laplacians
    (...,8) at barycentric points lam (...,4) of cells whose barycentric
    gradients are GL (...,4,3)."""
    phi = np.zeros(lam.shape[:-1] + (8,))
    grad = np.zeros(lam.shape[:-1] + (8, 3))
    lap = np.zeros(lam.shape[:-1] + (8,))
    phi[..., :4] = lam
    grad[..., :4, :] = GL
    for o, (i, j, k) in enumerate(_OPP):
        li, lj, lk = lam[..., i, None], lam[..., j, None], lam[..., k, None]
        phi[..., 4 + o] = 27 * (li * lj * lk)[..., 0]
        grad[..., 4 + o, :] = 27 * (lj * lk * GL[..., i, :] + li * lk * GL[..., j, :] + li * lj * GL[..., k, :])
        dij = (GL[..., i, :] * GL[..., j, :]).sum(-1)
        dik = (GL[..., i, :] * GL[..., k, :]).sum(-1)
        djk = (GL[..., j, :] * GL[..., k, :]).sum(-1)
        lap[..., 4 + o] = 54 * (lam[..., k] * dij + lam[..., j] * dik + lam[..., i] * djk)
    return phi, grad, lap


def geometry(points):
    """points (nc,4,3) -> barycentric gradients (nc,4,3), |det J| (nc,),
    largest squared edge (nc,)."""
    J = np.stack([points[:, 1] - points[:, 0], points[:, 2] - points[:, 0], points[:, 3] - points[:, 0]], axis=2)
    inv = np.linalg.inv(J)                        # rows: grad lambda_1..3
    GL = np.concatenate([-inv.sum(axis=1, keepdims=True), inv], axis=1)
    det = np.abs(np.linalg.det(J))
    h2 = np.max([((points[:, a] - points[:, b]) ** 2).sum(-1) for a in range(4) for b in range(a + 1, 4)], axis=0)
    return GL, det, h2


def blocks(points, Uc, nu, mixing=None, supg=None, rule=None):
    """Cell blocks (nc,8,8) of the optional terms for cell velocities Uc
    (nc,8,3).  mixing = (scale, radius, cx, cy) or None; supg = (tauScale,
    magic) or None.  Also returns the mixing-length statistics (weighted
    mean and max nu_t/nu, max strain)."""
    lam, w = rule or collapsed64()
    GL, det, h2 = geometry(points)
    nc = points.shape[0]
    out = np.zeros((nc, 8, 8))
    rw = ws = 0.0
    rmax = smax = 0.0
    for q in range(len(w)):
        lq = np.broadcast_to(lam[q], (nc, 4))
        phi, grad, lap = basis(lq, GL)                       # (nc,8), (nc,8,3), (nc,8)
        wq = w[q] * det
        if mixing is not None:
            scale, R, cx, cy = mixing
            gu = np.einsum("cmi,cmj->cij", Uc, grad)          # gu[c,i,j] = d u_i / d x_j
            S = 0.5 * (gu + np.swapaxes(gu, 1, 2))
            strain = np.sqrt(np.maximum(0.0, 2 * (S * S).sum((1, 2))))
            xq = (lq * points[:, :, 0]).sum(-1)
            yq = (lq * points[:, :, 1]).sum(-1)
            eta = np.clip(np.hypot(xq - cx, yq - cy) / R, 0.0, 1.0)
            ell = R * np.maximum(0.0, 0.14 - 0.08 * eta ** 2 - 0.06 * eta ** 4)
            nut = scale * ell * ell * strain
            out += (nut * wq)[:, None, None] * np.einsum("cad,cbd->cab", grad, grad)
            rw += float((nut / nu * wq).sum())
            ws += float(wq.sum())
            rmax = max(rmax, float((nut / nu).max()))
            smax = max(smax, float(strain.max()))
        if supg is not None:
            tau_scale, magic = supg
            adv = np.einsum("cm,cmd->cd", phi, Uc)
            speed2 = (adv * adv).sum(-1)
            diff = 4 * nu / h2
            den = np.maximum(4 * speed2 / h2 + magic * diff * diff, 1e-30)
            tau = tau_scale / np.sqrt(den)
            stream = np.einsum("cd,cad->ca", adv, grad)
            strong = stream - nu * lap
            out += (tau * wq)[:, None, None] * stream[:, :, None] * strong[:, None, :]
    return out, (rw / ws if ws > 0 else 0.0, rmax, smax)


def product_geometry(h):
    """Cell vertex coordinates (nc,4,3) and the pipe axis / radius of an H8
    context (vertex order of the product's cells)."""
    nc = h.info["cells"]
    pts = h.get("cell_points", nc * 12).reshape(nc, 4, 3)
    g = h.get("pipe_geometry", 12)
    return pts, float(g[0]), float(g[1]), float(g[2])


def extra_blocks_fn(case, points, nu, mixing=None, supg=None):
    """extra_blocks(U) for h8_oracle.simple: the blocks at the cell values of U."""
    from . import h8_oracle as O

    def f(U):
        return blocks(points, O.cell_values(case, U), nu, mixing, supg)[0]

    return f
