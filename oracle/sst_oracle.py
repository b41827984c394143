"""CPU oracle of the k-omega SST volume terms of a tetrahedron (P1+BF3).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

Restates turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
modules/transport/tet_volume_kernel.inc:1-30 (sst_tet_volume) with
volume_coefficients (modules/transport/portable/transport_core.hpp:14-25):
at each point of Keast's 24-point rule, velocity gradient, k, omega and
their gradients from the element's 8 basis functions; the SST-2003 local
state from the REFERENCE's own compiled core (oracle/_ref, via
refbind.RefSst); D = sigma nu_t, R = reaction rate, S = source, weighted by
det * w; A_ab = sum D grad phi_a . grad phi_b + R phi_a phi_b, b_a = sum S
phi_a.

Pins (tests/test_sst_volume_oracle.py): the quadrature table integrates
every barycentric monomial of degree <= 6 exactly (closed form
a! b! c! d! / (a + b + c + d + 3)!); the basis matches the exact
P1 + face-bubble definition and its gradients finite differences; the
diffusion part of each element matrix annihilates constants (sum over the
vertex columns).
"""
from __future__ import annotations

import itertools
from math import factorial

import numpy as np

ORBITS4 = [(0.214602871259151673, 0.356191386222544926, 0.00665379170969464511, False),
           (0.0406739585346113375, 0.87797812439616596, 0.00167953517588677629, False),
           (0.322337890142275652, 0.0329863295731730594, 0.00922619692394239903, True)]
ORBIT12 = ((0.0636610018750175255, 0.269672331458315873, 0.603005664791649076), 0.00803571428571428319)


def keast24():
    """(24,4) barycentric points and (24,) weights, in the reference order."""
    L, W = [], []
    for a, b, w, b_first in ORBITS4:
        for i in range(4):
            l = [a] * 4
            l[i if b_first else 3 - i] = b
            L.append(l)
            W.append(w)
    vals, w = ORBIT12
    for perm in sorted(set(itertools.permutations((0, 0, 1, 2)))):
        L.append([vals[r] for r in perm])
        W.append(w)
    return np.array(L), np.array(W)


def monomial_exact(e):
    return np.prod([factorial(k) for k in e]) / factorial(sum(e) + 3)


GL = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def basis(l):
    """phi (n,8) and reference gradients (n,8,3) at barycentric points l (n,4)."""
    l = np.atleast_2d(l)
    n = l.shape[0]
    phi = np.zeros((n, 8))
    gr = np.zeros((n, 8, 3))
    phi[:, :4] = l
    gr[:, :4, :] = GL[None]
    for i in range(4):
        js = [j for j in range(4) if j != i]
        phi[:, 4 + i] = 27.0 * l[:, js[0]] * l[:, js[1]] * l[:, js[2]]
        gr[:, 4 + i, :] = 27.0 * (GL[js[0]][None] * (l[:, js[1]] * l[:, js[2]])[:, None]
                                  + GL[js[1]][None] * (l[:, js[0]] * l[:, js[2]])[:, None]
                                  + GL[js[2]][None] * (l[:, js[0]] * l[:, js[1]])[:, None])
    return phi, gr


def volume_terms(ref_sst, cell_ref, cell_invJ, cell_det, cell_xy, U_free, U_fixed, k_free, k_fixed, w_free, w_fixed,
                 radius, nu, equation, k_floor=1e-8, omega_floor=1e-3, deviatoric=False):
    """Element matrices (nc,8,8) and vectors (nc,8).  U_free (3,nv),
    U_fixed (nfixed,3); cell_xy (nc,4,2) vertex coordinates about the axis."""
    ref = cell_ref
    free = ref >= 0
    slot = np.where(free, 0, -ref - 1)
    idx = np.where(free, ref, 0)
    Uc = np.where(free[..., None], U_free.T[idx], U_fixed[slot])      # (nc,8,3)
    Kc = np.where(free, k_free[idx], k_fixed[slot])
    Wc = np.where(free, w_free[idx], w_fixed[slot])
    return volume_terms_cells(ref_sst, Uc, Kc, Wc, cell_invJ, cell_det, cell_xy, radius, nu, equation, k_floor,
                              omega_floor, deviatoric)


def volume_terms_cells(ref_sst, Uc, Kc, Wc, cell_invJ, cell_det, cell_xy, radius, nu, equation, k_floor=1e-8,
                       omega_floor=1e-3, deviatoric=False, cell_vdist=None):
    """volume_terms from per-cell basis coefficients: Uc (nc,8,3), Kc and Wc
    (nc,8).  cell_vdist (nc,4), when given, is the wall distance at the four
    vertices: the distance at a point is its P1 interpolation floored at zero
    (modules/geometry/wall_distance_device.cuh:4-10) instead of the pipe's
    analytic distance."""
    nc = Uc.shape[0]
    Lq, Wq = keast24()
    phi, gr = basis(Lq)                                   # (24,8), (24,8,3)
    gp = np.matmul(gr[None], cell_invJ[:, None])                      # (nc,24,8,3): sum_j gr[q,a,j] invJ[c,j,d]
    gu = np.matmul(Uc.transpose(0, 2, 1)[:, None], gp).reshape(nc, 24, 9)  # du_i/dx_d row-major
    kv = Kc @ phi.T                                                   # (nc,24)
    wv = Wc @ phi.T
    gk = np.einsum("ca,cqad->cqd", Kc, gp)
    gw = np.einsum("ca,cqad->cqd", Wc, gp)
    if cell_vdist is not None:
        dq = np.maximum(0.0, np.einsum("qa,ca->cq", Lq, cell_vdist))
        out, _ = ref_sst.sst_local(gu.reshape(-1, 9), kv.ravel(), wv.ravel(), gk.reshape(-1, 3), gw.reshape(-1, 3),
                                   nu, radius, dist=dq.ravel(), k_floor=k_floor, omega_floor=omega_floor,
                                   deviatoric=deviatoric)
    else:
        xy = np.einsum("qa,cak->cqk", Lq, cell_xy)                    # (nc,24,2)
        out, _ = ref_sst.sst_local(gu.reshape(-1, 9), kv.ravel(), wv.ravel(), gk.reshape(-1, 3), gw.reshape(-1, 3),
                                   nu, radius, pipe_xy=xy.reshape(-1, 2), k_floor=k_floor, omega_floor=omega_floor,
                                   deviatoric=deviatoric)
    out = out.reshape(nc, 24, 14)
    wt = cell_det[:, None] * Wq[None, :]
    if equation == 2:  # momentum form: eddy-viscosity diffusion only
        D, R, S = wt * out[..., 0], np.zeros_like(wt), np.zeros_like(wt)
    elif equation == 0:
        D, R, S = wt * out[..., 3] * out[..., 0], wt * out[..., 10], wt * out[..., 8]
    else:
        D, R, S = wt * out[..., 4] * out[..., 0], wt * out[..., 11], wt * out[..., 9]
    G = gp.transpose(0, 1, 3, 2).reshape(nc, 24 * 3, 8)                # [c, (q, d), a]
    DG = (np.repeat(D, 3, axis=1))[:, :, None] * G
    mat = np.matmul(DG.transpose(0, 2, 1), G) + np.einsum("cq,qa,qb->cab", R, phi, phi, optimize=True)
    rhs = np.einsum("cq,qa->ca", S, phi)
    return mat, rhs, D
