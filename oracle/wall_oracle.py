"""CPU oracle of the wall-function momentum path of the k-omega SST solver.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

Restates, as sparse linear algebra, the reference's spiral-case momentum
and pressure coupling (turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/):

- wall quadrature caches (modules/wall/tet_wall_types.inc:27-61): twelve
  Dunavant points per wall face, the P1+BF3 wall trace and the normal-column
  Picard functional (4-point Gauss-Legendre mean over 90% of the inward
  normal chord through the wall cell);
- face-averaged Spalding fit (modules/wall/stage22d4b_nitsche_live.inc:25-96)
  with the REFERENCE's own compiled active-point fit
  (modules/wall/portable/spalding_active.hpp, refbind.RefSst.fit_tangential);
- Nitsche face blocks in BF3 flux mode (stage22d4b_nitsche_live.inc:99-335)
  and the frozen cross-component coupling
  (stage22d4d_coupled_matrixfree.inc:68-105);
- skew-symmetric convection (modules/foundation/skew_picard_tet.inc:9-49),
  implicit deviatoric volume tensor (modules/coupling/
  implicit_dev_volume_tet.inc:1-68);
- component row-L1 relaxation (modules/wall/segregated_vector_tet_22c.inc:81-104)
  and the gamma-free rAU (modules/coupling/stage22d4i3a_schur.cuh:11-41);
- the BF3 wall-flux constraint as explicit sparse matrices T and Q
  (modules/wall/stage22d4i_constraint.inc) and K = T^T M T
  (stage22d4d_coupled_matrixfree.inc:934-966, here a sparse triple product);
- block Gauss-Seidel in colour order with the tangential projection of wall
  bubble rows (:843-889), vectorised per colour;
- the constrained Schur operator B T Q R Q T^T B^T and velocity correction
  (stage22d4i3a_schur.cuh:55-75), PCG with the oracle AMG on
  B diag(rAU_x, rAU_y, rAU_z) B^T (hxt5b_boundary.cuh:1864-1883).

Unknowns are interleaved: row 3 i + c is component c of free DOF i
(reference free numbering).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from oracle import h8_oracle as H
from oracle import sst_oracle as S
from oracle import turb_oracle as TO

GAUSS4 = np.array([[-0.8611363115940525752, 0.3478548451374538574],
                   [-0.3399810435848562648, 0.6521451548625461426],
                   [0.3399810435848562648, 0.6521451548625461426],
                   [0.8611363115940525752, 0.3478548451374538574]])


def basis8(L):
    phi, _ = S.basis(np.atleast_2d(L))
    return phi


class Wall:
    """Wall faces of a wall-function duct with their quadrature caches and
    DOFs, computed from the mesh (the faces are the k-omega step's wall
    faces, in that order)."""

    def __init__(self, step: TO.Step, case: H.H8Case):
        m = step.m
        self.faces = step.wall
        nf = len(self.faces)
        self.nf = nf
        self.cell = np.array([f[0] for f in self.faces], dtype=np.int64)
        self.lf = np.array([f[1] for f in self.faces], dtype=np.int64)
        self.n = np.array([f[2] for f in self.faces]).reshape(nf, 3)
        self.area = np.array([f[3] for f in self.faces])
        self.h = 3.0 * np.abs(case.cell_det[self.cell]) / 6.0 / self.area
        self.qw = self.area[:, None] * TO.TRI12[None, :, 3]
        self.ql = np.zeros((nf, 12, 4))
        self.bw = np.zeros((nf, 12, 8))
        self.bp = np.zeros((nf, 12, 8))
        for f in range(nf):
            c, lf = self.cell[f], self.lf[f]
            L = np.zeros((12, 4))
            L[:, TO.TET_FACE[lf]] = TO.TRI12[:, :3]
            self.ql[f] = L
            self.bw[f] = basis8(L)
            P = m.points[m.tets[c]]
            E = (P[1:] - P[0]).T                                  # columns e1 e2 e3
            dl = np.linalg.solve(E, -self.n[f])
            dr = np.concatenate([[-dl.sum()], dl])
            for q in range(12):
                neg = dr < -1e-15
                exit_ = np.min(-L[q, neg] / dr[neg])
                for x, w in GAUSS4:
                    s = 0.45 * exit_ * (1.0 + x)
                    self.bp[f, q] += 0.5 * w * basis8(L[q] + s * dr)[0]
        ref = case.cell_ref
        self.bub = ref[self.cell, 4 + self.lf]
        self.vert = ref[self.cell[:, None], TO.TET_FACE[self.lf]]
        assert (self.bub >= 0).all() and (self.vert >= 0).all()
        nv = case.nv
        # T and Q as sparse maps on interleaved vectors
        T = sp.lil_matrix((3 * nv, 3 * nv))
        Q = sp.lil_matrix((3 * nv, 3 * nv))
        T.setdiag(1.0)
        Q.setdiag(1.0)
        for f in range(nf):
            n, b = self.n[f], self.bub[f]
            for i in range(3):
                for j in range(3):
                    T[3 * b + i, 3 * b + j] -= n[i] * n[j]
                    Q[3 * b + i, 3 * b + j] -= n[i] * n[j]
                for v in self.vert[f]:
                    for j in range(3):
                        T[3 * b + i, 3 * v + j] -= (20.0 / 27.0) * n[i] * n[j]
        self.T = T.tocsr()
        self.Q = Q.tocsr()
        self.bub_face = np.full(nv, -1, dtype=np.int64)
        self.bub_face[self.bub] = np.arange(nf)


def skew_tensor():
    _, Tc = H.tensors()
    return Tc + 0.5 * np.transpose(Tc, (0, 2, 1, 3))


def b_matrix(case: H.H8Case) -> sp.csr_matrix:
    """B: cells x interleaved free velocity (fixed DOFs dropped)."""
    ref = case.cell_ref
    rows, cols, vals = [], [], []
    for a in range(8):
        free = ref[:, a] >= 0
        c = np.flatnonzero(free)
        for d in range(3):
            rows.append(c)
            cols.append(3 * ref[c, a] + d)
            vals.append(case.bcoef[c, a, d])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(case.nc, 3 * case.nv))


def potential_start(m: TO.Mesh, pipe_geometry, g2free, nv: int, tol=None):
    """Potential-flow start (stage27_potential_host.hpp:44-167): the P1
    Laplace problem  int grad(phi).grad(v) = int_inlet v (U_in . n)  with
    phi = 0 on the outlet vertices and Neumann walls (solved here directly,
    scipy's sparse LU, where the reference runs Jacobi PCG to 1e-10), the
    mass-lumped projection of the cellwise grad(phi) onto the vertices, and
    one normal BF3 bubble per boundary face that sets its mean normal
    velocity: U_in . n on the inlet, zero on the wall, and on the outlet the
    projected value plus one uniform correction that balances the inlet
    flux.  Returns (U (3, nv) in free numbering -- fixed DOFs left out --,
    every entity's value (n_entities, 3), report dict)."""
    from scipy.sparse.linalg import spsolve

    pts, tets = m.points, m.tets
    npt = pts.shape[0]
    uin = np.asarray(pipe_geometry[6:9], dtype=np.float64)
    inlet, wall, outlet = (int(pipe_geometry[9]), int(pipe_geometry[10]), int(pipe_geometry[11]))
    P = pts[tets]                                                       # (nc,4,3)
    J = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], axis=2)   # columns
    det = np.linalg.det(J)
    vol = np.abs(det) / 6.0
    G = np.zeros((tets.shape[0], 4, 3))
    G[:, 1:] = np.linalg.inv(J)                                         # rows: grad of lambda_1..3
    G[:, 0] = -G[:, 1:].sum(1)
    Kl = np.einsum("c,cai,cbi->cab", vol, G, G)
    rows = np.repeat(tets, 4, axis=1).ravel()
    cols = np.tile(tets, (1, 4)).ravel()
    K = sp.csr_matrix((Kl.ravel(), (rows, cols)), shape=(npt, npt))
    rhs = np.zeros(npt)
    bfaces = np.flatnonzero(m.face_patch >= 0)
    plans = {int(f): TO.face_plan(m, int(f)) for f in bfaces}
    outlet_v = np.zeros(npt, dtype=bool)
    for f in bfaces:
        if m.face_patch[f] == outlet:
            outlet_v[m.faces[f]] = True
        elif m.face_patch[f] == inlet:
            _, _, n, area = plans[int(f)]
            rhs[m.faces[f]] += float(uin @ n) * area / 3.0
    free = ~outlet_v
    phi = np.zeros(npt)
    phi[free] = spsolve(K[free][:, free].tocsc(), rhs[free])
    gphi = np.einsum("ca,cai->ci", phi[tets], G)
    mass = np.bincount(tets.ravel(), weights=np.repeat(vol / 4.0, 4), minlength=npt)
    u = np.zeros((npt, 3))
    for d in range(3):
        u[:, d] = np.bincount(tets.ravel(), weights=np.repeat(vol / 4.0 * gphi[:, d], 4), minlength=npt) / mass
    ent = np.zeros((m.ns, 3))
    ent[:npt] = u
    in_flux = sum(plans[int(f)][3] * float(uin @ plans[int(f)][2]) for f in bfaces if m.face_patch[f] == inlet)
    out_f = [int(f) for f in bfaces if m.face_patch[f] == outlet]
    out_area = sum(plans[f][3] for f in out_f)
    out_proj = sum(plans[f][3] * float(u[m.faces[f]].mean(0) @ plans[f][2]) for f in out_f)
    corr_out = (-in_flux - out_proj) / out_area
    for f in bfaces:
        _, _, n, area = plans[int(f)]
        before = float(u[m.faces[f]].mean(0) @ n)
        kind = m.face_patch[f]
        goal = float(uin @ n) if kind == inlet else (before + corr_out if kind == outlet else 0.0)
        ent[npt + f] = (goal - before) / 0.45 * n
    U = np.zeros((3, nv))
    g2 = np.asarray(g2free)
    e = np.flatnonzero(g2 >= 0)
    U[:, g2[e]] = ent[e].T
    return U, ent, dict(phi=phi, inlet_flux=in_flux, outlet_area=out_area, outlet_projected=out_proj)


class WallPath:
    """Assembly, momentum sweeps, Schur operator and SIMPLE loop."""

    def __init__(self, case: H.H8Case, step: TO.Step, ref_sst, gamma_n=1e-6, active=1, zero_limit=True, skew=True,
                 deviatoric=True, m3_sweeps=20, m3_omega=0.7, symmetric=False, rau_scale=2.0, alpha_u=0.7,
                 fitter=None):
        """fitter: object with the fit_tangential method of
        refbind.RefSst (the default, the reference's compiled fit)."""
        self.case, self.step, self.ref = case, step, ref_sst
        self.fitter = fitter if fitter is not None else ref_sst
        self.W = Wall(step, case)
        self.gamma_n, self.active, self.zero_limit = gamma_n, active, zero_limit
        self.skew, self.deviatoric = skew, deviatoric
        self.m3_sweeps, self.m3_omega, self.symmetric = m3_sweeps, m3_omega, symmetric
        self.rau_scale, self.alpha_u = rau_scale, alpha_u
        self.B = b_matrix(case)
        self.Tt = skew_tensor() if skew else H.tensors()[1]
        rows = np.repeat(np.arange(case.nv), np.diff(case.mom_row))
        self.rows = rows
        self.dp = case.diag_pos()

    # -- wall fit and blocks -------------------------------------------------
    def face_fit(self, U):
        W, st, case = self.W, self.step, self.case
        Uc = H.cell_values(case, U)
        mask = [0x04, 0x11, 0x15, 0x1b, 0x1f][self.active - 1]
        act = [m for m in range(5) if mask & (1 << m)]
        pic = np.einsum("fq,fqb,fbd->fd", W.qw, W.bp, Uc[W.cell]) / W.qw.sum(1)[:, None]
        samp = np.zeros((W.nf, 5, 3))
        yy = np.broadcast_to(st.heights, (W.nf, 5)).copy()
        for m in act:
            cells = st.s_cell[:, :, m]
            phi = basis8(st.s_l[:, :, m].reshape(-1, 4)).reshape(W.nf, 12, 8)
            vq = np.einsum("fqb,fqbd->fqd", phi, Uc[cells])
            samp[:, m] = np.einsum("fq,fqd->fd", W.qw, vq) / W.qw.sum(1)[:, None]
            yy[:, m] = (W.qw * st.heights[m]).sum(1) / W.qw.sum(1)
        out, valid = self.fitter.fit_tangential(samp, yy, W.n, pic, st.nu, self.active, self.zero_limit)
        return out, valid

    def wall_nue(self, U):
        W, st = self.W, self.step
        Uc = H.cell_values(self.case, U)
        cells = np.repeat(W.cell, 12)
        out, *_ = st._state(cells, W.ql.reshape(-1, 4), Uc, st.k, st.w)
        return (st.nu + out[:, 0]).reshape(W.nf, 12)

    def nitsche(self, beta, nue):
        W = self.W
        g = self.gamma_n / W.h
        pp = np.einsum("fqa,fqb->fqab", W.bw, W.bw)                          # (nf,12,8,8)
        normalQ = (W.qw * nue * g[:, None])[:, :, None, None] * pp
        robin = (W.qw * beta[:, None])[:, :, None, None] * pp
        n2 = W.n ** 2
        blk = np.einsum("fd,fab->fabd", n2, normalQ.sum(1)) + np.einsum("fd,fab->fabd", 1 - n2, robin.sum(1))
        pen = np.einsum("fd,fab->fabd", n2, normalQ.sum(1))
        cpl = (normalQ - robin).sum(1)
        return blk, pen, cpl

    def dev_blocks(self, U):
        """(nc,8,8,3,3) implicit deviatoric tensors."""
        case, st = self.case, self.step
        Uc = H.cell_values(case, U)
        _, _, D = S.volume_terms_cells(self.ref, Uc, st.k[st.ent], st.w[st.ent], case.cell_invJ, case.cell_det,
                                       st.cell_xy, st.R, st.nu, 2, st.P.k_floor, st.P.omega_floor, st.P.deviatoric,
                                       st._cell_vdist())
        Lq, Wq = S.keast24()
        _, gr = S.basis(Lq)
        gp = np.einsum("qaj,cjd->cqad", gr, case.cell_invJ)
        NW = D + case.cell_det[:, None] * Wq[None, :] * st.nu
        nc = gp.shape[0]
        G = gp.reshape(nc, gp.shape[1], 24)
        Mx = np.matmul((NW[:, :, None] * G).transpose(0, 2, 1), G).reshape(nc, 8, 3, 8, 3)  # [c, a, x, b, y]
        # nu_eff (d_j phi_a d_i phi_b - 2/3 d_i phi_a d_j phi_b) at block (a, b), entry (i, j)
        return Mx.transpose(0, 1, 3, 4, 2) - (2.0 / 3.0) * Mx.transpose(0, 1, 3, 2, 4)

    # -- assembly -------------------------------------------------------------
    def plug(self):
        """Inlet velocity on every free vertex, zero face bubbles (plug_initial,
        hxt5b_boundary.cuh:698-700)."""
        case = self.case
        U = np.zeros((3, case.nv))
        vert = case.cell_ref[:, :4].ravel()
        vert = np.unique(vert[vert >= 0])
        U[:, vert] = np.asarray(self.step.inlet_velocity)[:, None]
        return U

    def constrain(self, U):
        return (self.W.T @ U.T.reshape(-1)).reshape(self.case.nv, 3).T.copy()

    def assemble(self, U, p, fit=None):
        """K (3nv x 3nv, CSR, = T^T M T), right-hand side T^T b, rau3, r3,
        delta3 (nv,3), the fit and the constrained velocity "U" the assembly
        ran at (T U: simple_momentum_assembly.inc:1-5).  fit (nf,3), when
        given, replaces the face fit."""
        case, W = self.case, self.W
        nv, nnz = case.nv, case.nnz
        U = self.constrain(U)
        eddy = self.step.eddy_blocks(U)
        aval, conv, C = H.assemble_central(case, U, self.Tt, eddy)
        if fit is None:
            fit, valid = self.face_fit(U)
            assert valid.all(), "oracle face fit failed"
        nue = self.wall_nue(U)
        blk, pen, cpl = self.nitsche(fit[:, 0], nue)
        ref = case.cell_ref
        slot = case.cell_slot.reshape(case.nc, 8, 8).astype(np.int64)
        # wall blocks into per-component entry values
        wv = np.zeros((nnz, 3))
        pv = np.zeros((nnz, 3))
        cross = np.zeros((nnz, 3, 3))
        rc = ref[W.cell]                                                 # (nf,8)
        both = (rc[:, :, None] >= 0) & (rc[:, None, :] >= 0)              # (nf,8,8)
        kpos = case.mom_row[np.where(rc >= 0, rc, 0)][:, :, None] + slot[W.cell]
        nn = np.einsum("fi,fj->fij", W.n, W.n) * (1.0 - np.eye(3))[None]
        fidx = np.broadcast_to(np.arange(W.nf)[:, None, None], both.shape)[both]
        kk = kpos[both]                                                  # face-major, (a, b) row-major
        np.add.at(wv, kk, blk[both])
        np.add.at(pv, kk, pen[both])
        np.add.at(cross, kk, nn[fidx] * cpl[both][:, None, None])
        vals = aval[:, None] + wv                                        # (nnz,3)
        fac = 1.0 / self.alpha_u - 1.0
        l1 = np.zeros((nv, 3))
        l1p = np.zeros((nv, 3))
        for d in range(3):
            l1[:, d] = np.bincount(self.rows, weights=np.abs(vals[:, d]), minlength=nv)
            l1p[:, d] = np.bincount(self.rows, weights=np.abs(vals[:, d] - pv[:, d]), minlength=nv)
        delta = fac * l1
        diag = vals[self.dp] + delta
        rau3 = self.rau_scale / diag
        r3 = self.rau_scale / (vals[self.dp] - pv[self.dp] + fac * l1p)
        rel = vals.copy()
        rel[self.dp] += delta
        # 3x3 entries: relaxed component values, cross coupling, deviatoric
        blocks = cross.copy()
        for d in range(3):
            blocks[:, d, d] += rel[:, d]
        rhs_dev = np.zeros((3, nv))
        if self.deviatoric:
            dev = self.dev_blocks(U)
            ra = np.repeat(ref[:, :, None], 8, axis=2)
            rb = np.repeat(ref[:, None, :], 8, axis=1)
            both = (ra >= 0) & (rb >= 0)
            pos = case.mom_row[np.where(ra >= 0, ra, 0)] + slot
            np.add.at(blocks, pos[both], dev[both])
            fixb = (ra >= 0) & (rb < 0)
            fv = case.fixed_value[np.where(rb < 0, -rb - 1, 0)]            # (nc,8,8,3)
            contrib = -np.einsum("nij,nj->ni", dev[fixb], fv[fixb])
            for d in range(3):
                np.add.at(rhs_dev[d], ra[fixb], contrib[:, d])
        r = np.repeat(np.arange(nv), np.diff(case.mom_row))
        I = (3 * r[:, None, None] + np.arange(3)[None, :, None]) + np.zeros((1, 1, 3), dtype=np.int64)
        J = (3 * case.mom_col[:, None, None] + np.arange(3)[None, None, :]) + np.zeros((1, 3, 1), dtype=np.int64)
        M = sp.csr_matrix((blocks.ravel(), (I.ravel(), J.ravel())), shape=(3 * nv, 3 * nv))
        K = (W.T.T @ M @ W.T).tocsr()
        bt = H.bt_apply(case, p)
        b = case.static_rhs + conv + rhs_dev + bt + delta.T * U
        bT = W.T.T @ b.T.reshape(-1)
        return dict(K=K, M=M, b=bT, rau3=rau3, r3=r3, delta3=delta, fit=fit, nue=nue, blk=blk, pen=pen, cpl=cpl, U=U)

    # -- block sweeps ---------------------------------------------------------
    def momentum(self, A):
        """Q, m3 sweeps, T on the assembly's velocity A["U"]; returns the new U."""
        case, W = self.case, self.W
        nv = case.nv
        K = A["K"]
        b = A["b"]
        x = W.Q @ A["U"].T.reshape(-1)
        ncol = int(case.color.max()) + 1
        rows_of = [np.flatnonzero(case.color == c) for c in range(ncol)]
        Kd = np.zeros((nv, 3, 3))
        for i in range(3):
            for j in range(3):
                Kd[:, i, j] = K[3 * np.arange(nv) + i, 3 * np.arange(nv) + j].A1
        fb = W.bub_face

        def colour(rows):
            idx = (3 * rows[:, None] + np.arange(3)).ravel()
            r = (b[idx] - K[idx] @ x).reshape(-1, 3)
            a = Kd[rows].copy()
            f = fb[rows]
            wb = f >= 0
            if wb.any():
                n = W.n[f[wb]]
                sc = np.maximum(1e-20, np.max(np.abs(np.diagonal(a[wb], axis1=1, axis2=2)), axis=1))
                a[wb] += sc[:, None, None] * np.einsum("ri,rj->rij", n, n)
                r[wb] -= np.einsum("ri,ri->r", n, r[wb])[:, None] * n
            sol = np.linalg.solve(a, r[:, :, None])[:, :, 0]
            if wb.any():
                n = W.n[f[wb]]
                sol[wb] -= np.einsum("ri,ri->r", n, sol[wb])[:, None] * n
            x[idx] += self.m3_omega * sol.ravel()

        for _ in range(self.m3_sweeps):
            for rows in rows_of:
                colour(rows)
            if self.symmetric:
                for rows in rows_of[::-1]:
                    colour(rows)
        return (W.T @ x).reshape(nv, 3).T.copy()

    # -- pressure -------------------------------------------------------------
    def schur_parts(self, r3):
        W = self.W
        R = sp.diags(r3.reshape(-1))
        C = W.T @ W.Q @ R @ W.Q @ W.T.T
        return C, (self.B @ C @ self.B.T).tocsr()

    def fine_aniso(self, rau3):
        """B diag(rAU3) B^T with its diagonal (the AMG fine operator)."""
        A = (self.B @ sp.diags(rau3.reshape(-1)) @ self.B.T).tocsr()
        return A, A.diagonal()

    def simple(self, U0, p0, max_outer, alpha_p=0.3, p_rtol=1e-3, p_atol=1e-12, p_max=600, smoother=None,
               p_solver="pcg", restart=30, hist=None, log=None):
        case = self.case
        U, p = U0.copy(), p0.copy()
        vc = H.VCycle(case.hier, smoother or H.Smoother())
        for it in range(1, max_outer + 1):
            A = self.assemble(U, p)
            U = self.momentum(A)
            cont = -H.continuity(case, U)
            Cm, S_op = self.schur_parts(A["r3"])
            Af, fd = self.fine_aniso(A["rau3"])
            vc.set_fine(Af, fd)
            if p_solver == "fgmres":
                dp, pr = H.fgmres(S_op, vc.apply, cont, p_rtol, p_atol, p_max, restart)
            else:
                dp, pr = H.pcg(S_op, vc.apply, cont, p_rtol, p_atol, p_max)
            p = p + alpha_p * dp
            U = U + (Cm @ (self.B.T @ dp)).reshape(case.nv, 3).T
            self.step.step(U)
            if hist is not None:
                hist.append({"it": it, "p_its": pr["its"], "cont_post": float(np.linalg.norm(H.continuity(case, U)))})
            if log is not None:
                log(it, pr["its"])
        return U, p
