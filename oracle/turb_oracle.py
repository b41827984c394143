"""CPU oracle of the k-omega SST transport step in the masked-P1 state space.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

Restates, for a tetrahedral pipe, the reference's step_turbulence
(turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/coupling/
g2_step_initialization.inc, g2_step_convection.inc, g2_step_k.inc,
g2_step_omega.inc) on the FULL scalar system, as the reference holds it:
one scalar per mesh entity (vertices, then faces), face-bubble rows fixed at
zero (g2_state.inc:12-27):

- matrix = base (nu diffusion, exact tensor) + Picard convection
  (hxt4b_foundation.inc:474, the H8 central tensor) + volume terms
  (tet_volume_kernel.inc, via sst_oracle) + inlet flux (tet_inlet_flux.inc)
  + omega wall condition (modules/wall/of_auto_point3_kernel.inc:19-76), combined per entry
  (hxt5c_sst_g1.cuh:438 combine_scalar) by sparse summation;
- row-L1 relaxation over every stored entry of each row
  (scalar_relaxation.inc:8-33, transport_core.hpp:26-31);
- "fgs-fixed" sweeps (scalar_linear_solver.inc:33-67): per sweep, colours
  ascending, x_i += w ((b_i - sum_{j != i} a_ij x_j) / a_ii - x_i) on free
  (vertex) rows, fixed rows set to zero (hxt5b_boundary.cuh:1820);
- floor of the vertex values (scalar_relaxation.inc:176).

The pointwise SST state, the Spalding fit over the active points
(active::fit_values, on the speed tangential to the wall face,
spalding_profile5_tet_fit.inc:40-95) and the OF-auto omega target come from
the REFERENCE's own compiled cores (refbind.RefSst).  The omega wall
condition is one of the reference's three (STAGE41_OMEGA_MODE,
sst_application.inc:88-139): of_auto (blended target) or sample_log (log
target) as a rank-one penalty at the third sample
(of_auto_point3_rank1_penalty.inc), or trace_log, the log target imposed on
the wall face by symmetric Nitsche (of_auto_trace_log_nitsche.inc).  Wall
samples are located by the barycentric adjacency walk of
modules/wall/tet_wall_match_cache.inc:3-67.

Pins (tests/test_turb_oracle.py): the twelve-point triangle rule integrates
every monomial of degree <= 6 exactly; the walk reconstructs each sample
point; the inlet block reproduces its right-hand side on the constant inlet
field (the flux terms cancel exactly there); the wall block vanishes on a
field equal to the target at the sample; the trace block is symmetric and
its residual on a linear field equal to the target on the face, and on a
constant field, are the closed-form face integrals of the basis and of its
normal derivatives (with h = 3 volume / area from the vertices); row-L1 relaxation keeps the fixed
point of the unrelaxed system; many sweeps converge to a direct solve; the
colouring is proper.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.sparse as sp

from oracle import h8_oracle as H
from oracle import sst_oracle as S

TET_FACE = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]])
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
ORIGINAL_HEIGHTS = np.array([1.4694676501540146e-4, 2.9389353003080293e-4, 4.4084029504620439e-4,
                             5.8778706006160585e-4, 7.3473382507700737e-4])
C_MU = 0.09


@dataclass
class Mesh:
    points: np.ndarray      # (np,3)
    tets: np.ndarray        # (nc,4)
    opp: np.ndarray         # (nc,4) face opposite each local vertex
    faces: np.ndarray       # (nf,3)
    owner: np.ndarray       # (nf,)
    neighbour: np.ndarray   # (ni,)
    face_patch: np.ndarray  # (nf,) -1 internal

    @property
    def nvert(self):
        return self.points.shape[0]

    @property
    def ns(self):
        return self.points.shape[0] + self.faces.shape[0]


def mesh_from_product(h) -> Mesh:
    npt, nf, ni, nc = (int(v) for v in h.get("mesh_sizes", 4))
    g = lambda n, c, t=np.int32: h.get(n, c, t)
    return Mesh(points=h.get("mesh_points", npt * 3).reshape(npt, 3), tets=g("mesh_tets", nc * 4).reshape(nc, 4),
                opp=g("mesh_opp", nc * 4).reshape(nc, 4), faces=g("mesh_faces", nf * 3).reshape(nf, 3),
                owner=g("mesh_owner", nf), neighbour=g("mesh_neighbour", ni), face_patch=g("mesh_face_patch", nf))


@dataclass
class Params:
    intensity: float = 0.05
    length_ratio: float = 0.07
    alpha_k: float = 0.3
    alpha_omega: float = 0.3
    k_sweeps: int = 20
    omega_sweeps: int = 20
    gs_omega: float = 0.7
    k_floor: float = 1e-8
    omega_floor: float = 1e-3
    deviatoric: bool = False
    wall: bool = True
    gamma: float = 1e6
    heights: Optional[np.ndarray] = None
    sample_scale: float = 0.5
    of_kappa: float = 0.41
    of_E: float = 9.8
    of_beta1: float = 0.075
    of_re_blend: float = 11.0
    of_prod_limit: float = 10.0
    active_points: int = 5          # Spalding fit points (the wall path's count when it runs)
    omega_target: str = "sample_log"  # of_auto | sample_log | trace_log
    trace_gamma: float = 50.0


def cell_entities(m: Mesh) -> np.ndarray:
    return np.concatenate([m.tets, m.nvert + m.opp], axis=1)


def face_plan(m: Mesh, f: int):
    """(cell, lf, outward unit normal, area) of boundary face f."""
    c = int(m.owner[f])
    lf = int(np.flatnonzero(m.opp[c] == f)[0])
    p = m.points[m.faces[f]]
    n = np.cross(p[1] - p[0], p[2] - p[0])
    ln = np.linalg.norm(n)
    opp = m.points[m.tets[c, lf]]
    s = 1.0 if np.dot(n, p[0] - opp) > 0 else -1.0
    return c, lf, s * n / ln, 0.5 * ln


def adjacency(m: Mesh) -> np.ndarray:
    """adj[c][j]: the cell across the face opposite local vertex j (-1 on the boundary)."""
    ni = m.neighbour.size
    adj = np.full(m.opp.shape, -1, dtype=np.int64)
    for c in range(m.tets.shape[0]):
        for j in range(4):
            f = m.opp[c, j]
            if m.owner[f] != c:
                adj[c, j] = m.owner[f]
            elif f < ni:
                adj[c, j] = m.neighbour[f]
    return adj


def barycentric(m: Mesh, c: int, p):
    v = m.points[m.tets[c]]
    J = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)
    q = np.linalg.solve(J, p - v[0])
    return np.array([1.0 - q.sum(), q[0], q[1], q[2]])


def locate(m: Mesh, adj: np.ndarray, start: int, p):
    """Barycentric adjacency walk from cell `start` (tet_wall_match_cache.inc:35-52)."""
    c, seen = start, set()
    while c >= 0 and c not in seen and len(seen) < 4096:
        seen.add(c)
        l = barycentric(m, c, p)
        neg = int(np.argmin(l))
        if l[neg] >= -1e-8:
            return c, l
        c = int(adj[c, neg])
    raise RuntimeError("sample outside the mesh")


class Step:
    """The transport step on one context's data."""

    def __init__(self, m: Mesh, case: H.H8Case, pipe: np.ndarray, P: Params, ref_sst, wall_distance=None):
        """pipe: cx cy R D bulk nu inletVelocity(3) inletPatch wallPatch outletPatch.
        wall_distance (np,): vertex distance field of a duct (None: the
        straight pipe's analytic distance about (cx, cy))."""
        self.m, self.case, self.P, self.ref = m, case, P, ref_sst
        self.vdist = None if wall_distance is None else np.asarray(wall_distance, dtype=float)
        self.cx, self.cy, self.R, self.D, self.bulk, self.nu = pipe[:6]
        self.inlet_velocity = pipe[6:9]
        inlet_patch, wall_patch = int(pipe[9]), int(pipe[10])
        self.ent = cell_entities(m)
        self.k_in = 1.5 * (P.intensity * self.bulk) ** 2
        self.omega_in = np.sqrt(self.k_in) / (C_MU ** 0.25 * P.length_ratio * self.D)
        self.heights = (np.asarray(P.heights, dtype=float) if P.heights is not None
                        else ORIGINAL_HEIGHTS * self.D / 0.004 * P.sample_scale)
        bf = np.arange(m.neighbour.size, m.faces.shape[0])
        self.inlet = [face_plan(m, f) for f in bf if m.face_patch[f] == inlet_patch]
        self.wall = [face_plan(m, f) for f in bf if P.wall and m.face_patch[f] == wall_patch]
        adj = adjacency(m)
        nw = len(self.wall)
        self.s_cell = np.zeros((nw, 12, 5), dtype=np.int64)
        self.s_l = np.zeros((nw, 12, 5, 4))
        for fi, (c, lf, n, _) in enumerate(self.wall):
            X = TRI12[:, :3] @ m.points[m.tets[c, TET_FACE[lf]]]
            for q in range(12):
                for k, y in enumerate(self.heights):
                    self.s_cell[fi, q, k], self.s_l[fi, q, k] = locate(m, adj, c, X[q] - y * n)
        T_diff, self.T_cent = H.tensors()
        G = np.einsum("cjd,ckd->cjk", case.cell_invJ, case.cell_invJ)
        self.diff = np.einsum("abjk,cjk->cab", T_diff, G) * case.cell_det[:, None, None]
        xy = m.points[m.tets][:, :, :2] - np.array([self.cx, self.cy])
        self.cell_xy = xy
        nv = m.nvert
        self.k = np.zeros(m.ns)
        self.w = np.zeros(m.ns)
        self.k[:nv], self.w[:nv] = self.k_in, self.omega_in
        # vertex-row colouring of the coupling graph (greedy, natural order)
        rows = np.repeat(self.ent[:, :4], 8, axis=1).ravel()
        cols = np.tile(self.ent, (1, 4)).ravel()
        vv = cols < nv
        G = sp.csr_matrix((np.ones(vv.sum()), (rows[vv], cols[vv])), shape=(nv, nv))
        G.sum_duplicates()
        self.colors = H.greedy_colors(G)

    # -- pointwise state at barycentric points of given cells -------------------
    def _state(self, cells, L, Uc_all, k, w):
        """Local SST outputs (n,14), phi (n,8), grad phi (n,8,3), velocity (n,3), k, omega at points."""
        phi, gr = S.basis(L)
        inv = self.case.cell_invJ[cells]
        gp = np.einsum("naj,njd->nad", gr, inv)
        Uc = Uc_all[cells]
        Kc, Wc = k[self.ent[cells]], w[self.ent[cells]]
        gu = np.einsum("nai,nad->nid", Uc, gp).reshape(-1, 9)
        kv, wv = (Kc * phi).sum(1), (Wc * phi).sum(1)
        gk, gw = np.einsum("na,nad->nd", Kc, gp), np.einsum("na,nad->nd", Wc, gp)
        if self.vdist is not None:
            d = np.maximum(0.0, np.einsum("na,na->n", L, self.vdist[self.m.tets[cells]]))
            out, _ = self.ref.sst_local(gu, kv, wv, gk, gw, self.nu, self.R, dist=d, k_floor=self.P.k_floor,
                                        omega_floor=self.P.omega_floor, deviatoric=self.P.deviatoric)
        else:
            xy = np.einsum("na,nak->nk", L, self.cell_xy[cells])
            out, _ = self.ref.sst_local(gu, kv, wv, gk, gw, self.nu, self.R, pipe_xy=xy, k_floor=self.P.k_floor,
                                        omega_floor=self.P.omega_floor, deviatoric=self.P.deviatoric)
        us = np.einsum("nai,na->ni", Uc, phi)
        return out, phi, gp, us, kv, wv

    def assemble(self, eq: int, U: np.ndarray):
        """Full scalar matrix (nS x nS, CSR) and right-hand side of equation eq
        (0: k with the step's k and omega, 1: omega with the new k)."""
        m, case, P = self.m, self.case, self.P
        nc = case.nc
        Uc = H.cell_values(case, U)
        ur = np.einsum("cmd,cjd->cmj", Uc, case.cell_invJ)
        conv = np.einsum("cmj,ambj->cab", ur, self.T_cent) * case.cell_det[:, None, None]
        Kc, Wc = self.k[self.ent], self.w[self.ent]
        vol, vrhs, _ = S.volume_terms_cells(self.ref, Uc, Kc, Wc, case.cell_invJ, case.cell_det, self.cell_xy, self.R,
                                            self.nu, eq, P.k_floor, P.omega_floor, P.deviatoric, self._cell_vdist())
        blocks = [self.nu * self.diff + conv + vol]
        brhs = [vrhs]
        bcell = [np.arange(nc)]
        # inlet flux
        ib, ir = [], []
        for face in self.inlet:
            A, b = self.inlet_block(face, eq, Uc)
            ib.append(A)
            ir.append(b)
            bcell.append(np.array([face[0]]))
        if ib:
            blocks.append(np.array(ib))
            brhs.append(np.array(ir))
        # omega wall condition
        if eq == 1 and self.wall:
            wb, wr, wc = self.wall_blocks(Uc)
            blocks.append(wb)
            brhs.append(wr)
            bcell.append(wc)
        Bk = np.concatenate(blocks)
        Br = np.concatenate(brhs)
        cells = np.concatenate(bcell)
        E = self.ent[cells]
        rows = np.repeat(E, 8, axis=1).ravel()
        cols = np.tile(E, (1, 8)).ravel()
        A = sp.csr_matrix((Bk.ravel(), (rows, cols)), shape=(m.ns, m.ns))
        A.sum_duplicates()
        rhs = np.bincount(E.ravel(), weights=Br.ravel(), minlength=m.ns)
        return A, rhs

    def inlet_block(self, face, eq: int, Uc):
        """8x8 block and 8-vector of the inlet flux of one face
        (tet_inlet_flux.inc:1-23): sum_q w_q [qin phi_a phi_b
        + nu_e (-phi_a dn phi_b + dn phi_a phi_b)] and sum_q w_q [qin phi_a
        + nu_e dn phi_a] phi_in."""
        c, lf, n, area = face
        phi_in = self.k_in if eq == 0 else self.omega_in
        L = np.zeros((12, 4))
        L[:, TET_FACE[lf]] = TRI12[:, :3]
        out, phi, gp, _, _, _ = self._state(np.full(12, c), L, Uc, self.k, self.w)
        nue = self.nu + (out[:, 3] if eq == 0 else out[:, 4]) * out[:, 0]
        wq = TRI12[:, 3] * area
        dn = gp @ n
        inflow = max(0.0, -float(np.dot(self.inlet_velocity, n)))
        A = np.einsum("q,qa,qb->ab", wq * inflow, phi, phi) + \
            np.einsum("q,qab->ab", wq * nue, -phi[:, :, None] * dn[:, None, :] + dn[:, :, None] * phi[:, None, :])
        b = np.einsum("q,qa->a", wq * inflow * phi_in, phi) + np.einsum("q,qa->a", wq * nue * phi_in, dn)
        return A, b

    def wall_data(self, Uc, with_nue=False):
        """Per wall quadrature point: penalty coefficient, omega target,
        basis at the third sample, its cell and the invalid count (and nu_e
        at the third sample)."""
        P = self.P
        nw = len(self.wall)
        cells = self.s_cell.reshape(-1)
        L = self.s_l.reshape(-1, 4)
        phi, _ = S.basis(L)
        uv = np.einsum("nad,na->nd", Uc[cells], phi).reshape(nw * 12, 5, 3)
        nrm = np.repeat(np.array([f[2] for f in self.wall]), 12, axis=0)
        tang = uv - np.einsum("nmd,nd->nm", uv, nrm)[:, :, None] * nrm[:, None, :]
        speed = np.linalg.norm(tang, axis=2)
        self.last_speed = speed
        y = np.broadcast_to(self.heights, (nw * 12, 5))
        fit, valid = self.ref.fit_values(speed, y, self.nu, P.active_points)
        c3 = self.s_cell[:, :, 2].reshape(-1)
        L3 = self.s_l[:, :, 2].reshape(-1, 4)
        out, phi3, _, us, kv, wv = self._state(c3, L3, Uc, self.k, self.w)
        n = np.repeat(np.array([f[2] for f in self.wall]), 12, axis=0)
        area = np.repeat(np.array([f[3] for f in self.wall]), 12)
        normal = (us * n).sum(1)
        Ut = np.sqrt(np.maximum(0.0, (us * us).sum(1) - normal ** 2))
        q = nw * 12
        inp = np.column_stack([np.full(q, self.heights[2]), np.full(q, self.nu), Ut, fit[:, 0], kv, wv,
                               np.full(q, P.k_floor), np.full(q, P.omega_floor), np.full(q, P.of_kappa),
                               np.full(q, P.of_E), np.full(q, P.of_beta1), np.full(q, P.of_re_blend),
                               np.full(q, P.of_prod_limit), out[:, 8], out[:, 0]])
        oa = self.ref.of_auto(inp, np.ones(q, dtype=np.int32))
        tgt = oa[:, 0] if P.omega_target == "of_auto" else oa[:, 2]
        ell = self.heights[4] - self.heights[0]
        nue = self.nu + out[:, 4] * out[:, 0]
        vw = np.tile(TRI12[:, 3], nw) * area * ell
        coef = P.gamma * nue / ell * vw / ell
        ok = (valid != 0) & np.isfinite(nue) & (nue > 0) & np.isfinite(coef) & (coef > 0) & \
            np.isfinite(tgt) & (tgt > 0)
        res = np.where(ok, coef, 0.0), np.where(ok, tgt, 0.0), phi3, c3, int((~ok).sum())
        return res + (np.where(ok, nue, 0.0),) if with_nue else res

    def trace_block(self, face, q: int, mu: float, g: float):
        """8x8 block and 8-vector of the trace_log condition at point q of a
        wall face on the face's cell (of_auto_trace_log_nitsche.inc:6-45):
        w [-mu dn phi_b phi_a - mu dn phi_a phi_b + tau phi_a phi_b] and
        w [-mu dn phi_a + tau phi_a] g, w = w_q area, tau = gamma mu / h,
        h = det / (2 area)."""
        c, lf, n, area = face
        L = np.zeros((1, 4))
        L[0, TET_FACE[lf]] = TRI12[q, :3]
        phi, gr = S.basis(L)
        dn = (gr[0] @ self.case.cell_invJ[c]) @ n
        phi = phi[0]
        h = self.case.cell_det[c] / (2.0 * area)
        tau = self.P.trace_gamma * mu / h
        w = TRI12[q, 3] * area
        A = w * (-mu * np.outer(phi, dn) - mu * np.outer(dn, phi) + tau * np.outer(phi, phi))
        b = w * (-mu * dn + tau * phi) * g
        ok = all(np.isfinite(v) and v > 0 for v in (h, mu, tau, w, g))
        return A, b, ok

    def wall_blocks(self, Uc):
        if self.P.omega_target != "trace_log":
            coef, target, phi3, c3, _ = self.wall_data(Uc)
            wb = coef[:, None, None] * phi3[:, :, None] * phi3[:, None, :]
            wr = (coef * target)[:, None] * phi3
            return wb, wr, c3
        coef, target, _, _, _, nue = self.wall_data(Uc, with_nue=True)
        nq = coef.size
        wb, wr = np.zeros((nq, 8, 8)), np.zeros((nq, 8))
        for i in range(nq):
            if coef[i] > 0:
                A, b, ok = self.trace_block(self.wall[i // 12], i % 12, nue[i], target[i])
                if ok:
                    wb[i], wr[i] = A, b
        return wb, wr, np.repeat(np.array([f[0] for f in self.wall]), 12)

    @staticmethod
    def relax(A: sp.csr_matrix, rhs: np.ndarray, x: np.ndarray, alpha: float):
        """Row-L1 relaxation over every stored entry; returns the relaxed copy."""
        A = A.tocsr(copy=True)
        l1 = np.asarray(abs(A).sum(axis=1)).ravel()
        delta = (1.0 / alpha - 1.0) * l1
        d = A.diagonal() + delta
        A.setdiag(d)
        return A, rhs + delta * x, d

    def sweeps(self, A: sp.csr_matrix, b: np.ndarray, x: np.ndarray, n: int, omega: float):
        nv = self.m.nvert
        x = x.copy()
        x[nv:] = 0.0
        d = A.diagonal()
        order = [np.flatnonzero(self.colors == c) for c in range(int(self.colors.max()) + 1)]
        Av = A[:nv].tocsr()
        for _ in range(n):
            for rows in order:
                s = Av[rows] @ x - d[rows] * x[rows]
                x[rows] += omega * ((b[rows] - s) / d[rows] - x[rows])
        return x

    def equation(self, eq: int, U: np.ndarray):
        P = self.P
        x = self.k if eq == 0 else self.w
        A, rhs = self.assemble(eq, U)
        Ar, br, d = self.relax(A, rhs, x, P.alpha_k if eq == 0 else P.alpha_omega)
        self.last = (Ar, br, d)
        xn = self.sweeps(Ar, br, x, P.k_sweeps if eq == 0 else P.omega_sweeps, P.gs_omega)
        nv = self.m.nvert
        fl = P.k_floor if eq == 0 else P.omega_floor
        bad = ~(xn[:nv] >= fl) | ~np.isfinite(xn[:nv])
        xn[:nv][bad] = fl
        if eq == 0:
            self.k = xn
        else:
            self.w = xn
        return int(bad.sum())

    def _cell_vdist(self):
        return None if self.vdist is None else self.vdist[self.m.tets]

    def eddy_blocks(self, U: np.ndarray) -> np.ndarray:
        """(nc,8,8) eddy-viscosity diffusion of the momentum equation at the
        current k, omega: sum_q w_q det nu_t grad phi_a . grad phi_b (the
        isotropic part of sst_mom_tet, hxt5d_sst_g2.cuh:1220)."""
        P, case = self.P, self.case
        Uc = H.cell_values(case, U)
        blocks, _, _ = S.volume_terms_cells(self.ref, Uc, self.k[self.ent], self.w[self.ent], case.cell_invJ,
                                            case.cell_det, self.cell_xy, self.R, self.nu, 2, P.k_floor,
                                            P.omega_floor, P.deviatoric, self._cell_vdist())
        return blocks

    def step(self, U: np.ndarray):
        """One k then omega update at velocity U (3, nv free, reference order)."""
        ck = self.equation(0, U)
        cw = self.equation(1, U)
        return ck, cw
