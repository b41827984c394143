"""CPU oracle of the H8 binary's optional momentum terms: the Nikuradse
mixing-length eddy viscosity, the implicit SUPG stabilisation and the DG
weak inlet.

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
and Gauss-Legendre roots.  The DG inlet restates
gpu/serial_cuda/h10_dg_inlet.cuh:94-133: per inlet face, beta = max(0,
-uh.n), nu_e = nu + nu_t, G_ab = sum_q w (beta phi_a phi_b + nu_e (dn phi_a
phi_b - phi_a dn phi_b)), R_a = sum_q w (beta phi_a + nu_e dn phi_a) uh on
Dunavant's 12-point degree-6 triangle rule.
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


# Dunavant's 12-point degree-6 triangle rule: (l0, l1, l2, weight), weights
# summing to 1 (Dunavant 1985, Int. J. Numer. Meth. Eng. 21, table entry 6).
TRI12 = np.array([
    [0.063089014491502, 0.063089014491502, 0.873821971016996, 0.050844906370207],
    [0.063089014491502, 0.873821971016996, 0.063089014491502, 0.050844906370207],
    [0.873821971016996, 0.063089014491502, 0.063089014491502, 0.050844906370207],
    [0.249286745170910, 0.249286745170910, 0.501426509658180, 0.116786275726379],
    [0.249286745170910, 0.501426509658180, 0.249286745170910, 0.116786275726379],
    [0.501426509658180, 0.249286745170910, 0.249286745170910, 0.116786275726379],
    [0.053145049844816, 0.310352451033785, 0.636502499121399, 0.082851075618374],
    [0.053145049844816, 0.636502499121399, 0.310352451033785, 0.082851075618374],
    [0.310352451033785, 0.053145049844816, 0.636502499121399, 0.082851075618374],
    [0.310352451033785, 0.636502499121399, 0.053145049844816, 0.082851075618374],
    [0.636502499121399, 0.053145049844816, 0.310352451033785, 0.082851075618374],
    [0.636502499121399, 0.310352451033785, 0.053145049844816, 0.082851075618374]])


def face_frame(points, lf):
    """Area (nf,) and outward unit normal (nf,3) of local face lf (opposite
    vertex lf) of cells with vertex coordinates points (nf,4,3)."""
    idx = np.array([[j for j in range(4) if j != o] for o in range(4)])[lf]    # (nf,3)
    V = np.take_along_axis(points, idx[:, :, None], axis=1)
    sf = 0.5 * np.cross(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0])
    opp = points[np.arange(len(lf)), lf]
    flip = np.einsum("fd,fd->f", sf, V[:, 0] - opp) < 0
    sf[flip] *= -1
    area = np.linalg.norm(sf, axis=1)
    return area, sf / area[:, None]


def dg_inlet(points, faces, Uc, nu, uh, mixing=None, rule=TRI12):
    """G (nf,8,8) and R (nf,8,3) of the DG inlet faces (cell, lf) with the
    cells' velocities Uc (nf,8,3) and vertex coordinates points (nf,4,3)."""
    lf = np.asarray(faces)[:, 1]
    nf = len(lf)
    GL, _, _ = geometry(points)
    area, n = face_frame(points, lf)
    uh = np.asarray(uh, float)
    inflow = np.maximum(0.0, -(n @ uh))
    fv = np.array([[j for j in range(4) if j != o] for o in range(4)])[lf]
    G = np.zeros((nf, 8, 8))
    R = np.zeros((nf, 8, 3))
    for q in range(len(rule)):
        lam = np.zeros((nf, 4))
        for j in range(3):
            lam[np.arange(nf), fv[:, j]] = rule[q, j]
        phi, grad, _ = basis(lam, GL)
        dn = np.einsum("fad,fd->fa", grad, n)
        nue = np.full(nf, nu)
        if mixing is not None:
            scale, Rr, cx, cy = mixing
            gu = np.einsum("fmi,fmj->fij", Uc, grad)
            S = 0.5 * (gu + np.swapaxes(gu, 1, 2))
            strain = np.sqrt(np.maximum(0.0, 2 * (S * S).sum((1, 2))))
            xq = (lam * points[:, :, 0]).sum(-1)
            yq = (lam * points[:, :, 1]).sum(-1)
            eta = np.clip(np.hypot(xq - cx, yq - cy) / Rr, 0.0, 1.0)
            ell = Rr * np.maximum(0.0, 0.14 - 0.08 * eta ** 2 - 0.06 * eta ** 4)
            nue = nue + scale * ell * ell * strain
        w = rule[q, 3] * area
        G += w[:, None, None] * (inflow[:, None, None] * phi[:, :, None] * phi[:, None, :]
                                 + nue[:, None, None] * (dn[:, :, None] * phi[:, None, :] - phi[:, :, None] * dn[:, None, :]))
        R += (w * inflow)[:, None, None] * phi[:, :, None] * uh[None, None, :] \
            + (w * nue)[:, None, None] * dn[:, :, None] * uh[None, None, :]
    return G, R


def dg_inlet_fns(case, points_all, faces, nu, uh, mixing=None):
    """(extra_blocks(U), extra_rhs(U)) for h8_oracle.simple: the DG inlet's
    cell blocks and its right-hand side on the free rows."""
    from . import h8_oracle as O

    faces = np.asarray(faces)
    cells = faces[:, 0]

    def terms(U):
        Uc = O.cell_values(case, U)[cells]
        return dg_inlet(points_all[cells], faces, Uc, nu, uh, mixing)

    def blocks_fn(U):
        G, _ = terms(U)
        out = np.zeros((case.nc, 8, 8))
        np.add.at(out, cells, G)
        return out

    def rhs_fn(U):
        _, R = terms(U)
        b = np.zeros((3, case.nv))
        ref = case.cell_ref[cells]                     # (nf,8)
        free = ref >= 0
        for d in range(3):
            np.add.at(b[d], ref[free], R[:, :, d][free])
        return b

    return blocks_fn, rhs_fn
