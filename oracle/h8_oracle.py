"""CPU FP64 oracle of the H8 SIMPLE hot path.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; never by the
product package.

This module restates, in plain numpy/scipy, the reference algorithms the
HIP kernels implement, each function citing the reference file:line it
follows.  Setup data (DOF numbering, cell plans, momentum CSR, colouring,
Dirichlet values, AMG hierarchy) enter as arrays: they are pinned bit-exactly
against the reference's own host builders compiled in ``oracle/_ref``
(tests/test_setup_vs_reference.py).  The oracle's own restatements of the
reference tensors and of the fine Schur topology are pinned by closed-form
integrals, by a brute-force pattern check and by action parity against the
reference's matrix-free ``cpu_schur_apply`` (tests/test_oracle.py).

Every function is pinned; none is "parity unpinned".
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import scipy.sparse as sp

# ---------------------------------------------------------------------------
# reference tensors
# ---------------------------------------------------------------------------


def duffy125() -> Tuple[np.ndarray, np.ndarray]:
    """125-point collapsed Gauss-Jacobi rule on the reference tetrahedron.

    gpu/serial_cuda/g3_host.hpp:58-71 (nodes/weights of the published rule).
    Returns barycentric coordinates (125,4) and weights (125,).
    """
    rn = [0.034578939918215090, 0.17348032077169567, 0.38988638706551931, 0.63433347263088680, 0.85105421294701644]
    rw = [0.081764784285771011, 0.12619896189991137, 0.089200161221590066, 0.032055600722961895, 0.0041138252030990035]
    sn = [0.039809857051468722, 0.19801341787360821, 0.43797481024738616, 0.69546427335363614, 0.90146491420117358]
    sw = [0.096781590226651476, 0.16717463809436969, 0.14638698708466985, 0.073908870072616678, 0.015747914521692299]
    tn = [0.046910077030668018, 0.23076534494715845, 0.50000000000000000, 0.76923465505284150, 0.95308992296933193]
    tw = [0.11846344252809449, 0.23931433524968326, 0.28444444444444450, 0.23931433524968326, 0.11846344252809449]
    lam, w = [], []
    for ir in range(5):
        for i_s in range(5):
            for it in range(5):
                r, s, t = rn[ir], sn[i_s], tn[it]
                omr, oms = 1 - r, 1 - s
                lam.append((omr * oms * (1 - t), r, omr * s, omr * oms * t))
                w.append(rw[ir] * sw[i_s] * tw[it])
    return np.array(lam), np.array(w)


_GL = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def p1bf3_basis(lam: np.ndarray):
    """Values (8,) and reference gradients (8,3) of P1 + face bubbles.

    Vertex functions lambda_i; bubble 4+i = 27 lambda_j lambda_k lambda_l on
    the face opposite vertex i (gpu/serial_cuda/g3_host.hpp:73-85,
    g4_host.hpp:93-95).
    """
    val = np.zeros(8)
    gr = np.zeros((8, 3))
    val[:4] = lam
    gr[:4] = _GL
    for i in range(4):
        js = [j for j in range(4) if j != i]
        val[4 + i] = 27.0 * lam[js[0]] * lam[js[1]] * lam[js[2]]
        for a in range(3):
            j, o1, o2 = js[a], js[(a + 1) % 3], js[(a + 2) % 3]
            gr[4 + i] += 27.0 * lam[o1] * lam[o2] * _GL[j]
    return val, gr


def tensors() -> Tuple[np.ndarray, np.ndarray]:
    """Diffusion T[a][b][j][k] = int grad_j phi_a grad_k phi_b and central
    convection C[a][m][b][j] = int phi_a phi_m d_j phi_b on the reference
    tetrahedron (gpu/serial_cuda/g3_host.hpp:87-100, g4_host.hpp:92-98)."""
    lam, w = duffy125()
    D = np.zeros((8, 8, 3, 3))
    Cc = np.zeros((8, 8, 8, 3))
    for q in range(lam.shape[0]):
        val, gr = p1bf3_basis(lam[q])
        D += np.einsum("aj,bk->abjk", gr, gr) * w[q]
        Cc += np.einsum("a,m,bj->ambj", val, val, gr) * w[q]
    return D, Cc


# ---------------------------------------------------------------------------
# case data
# ---------------------------------------------------------------------------


@dataclass
class H8Case:
    """Setup arrays of an H8 case, reference numbering (free DOFs)."""
    nc: int
    nv: int
    cell_ref: np.ndarray      # (nc,8) int: >=0 free DOF, <0 -(fixed slot+1)
    cell_slot: np.ndarray     # (nc,64) uint8 row slots
    cell_det: np.ndarray      # (nc,)
    cell_invJ: np.ndarray     # (nc,3,3)
    bcoef: np.ndarray         # (nc,8,3)
    fixed_value: np.ndarray   # (nfixed,3)
    mom_row: np.ndarray       # (nv+1,) int64
    mom_col: np.ndarray       # (nnz,) int32
    mom_diff: np.ndarray      # (nnz,) diffusion at nu, alpha=1
    color: np.ndarray         # (nv,) greedy colours
    static_rhs: np.ndarray    # (3,nv)
    fixed_div: np.ndarray     # (nc,)
    init_rau: np.ndarray      # (nv,)
    cell_z: Optional[np.ndarray] = None
    zspan: float = 1.0
    pin: int = -1             # pressure row pinned in a closed domain (-1: none)
    hier: Optional[dict] = None
    _fine: Optional[dict] = field(default=None, repr=False)

    @property
    def nnz(self) -> int:
        return int(self.mom_row[-1])

    def diag_pos(self) -> np.ndarray:
        rows = np.repeat(np.arange(self.nv), np.diff(self.mom_row))
        return np.flatnonzero(self.mom_col == rows)


def case_from_product(h, with_hierarchy: bool = True) -> H8Case:
    """Collect setup arrays from a product H8 context (pinned separately
    against the reference builders)."""
    nc, nv = h.info["cells"], h.info["nv"]
    nnz, nfix = h.info["mom_nnz"], h.info["nfixed"]
    c = H8Case(
        nc=nc, nv=nv,
        cell_ref=h.get("cell_ref", nc * 8).reshape(nc, 8).astype(np.int64),
        cell_slot=h.get("cell_slot", nc * 64).reshape(nc, 64),
        cell_det=h.get("cell_det", nc),
        cell_invJ=h.get("cell_invJ", nc * 9).reshape(nc, 3, 3),
        bcoef=h.get("bcoef", nc * 24).reshape(nc, 8, 3),
        fixed_value=h.get("fixed_value", nfix * 3).reshape(nfix, 3),
        mom_row=h.get("mom_row", nv + 1),
        mom_col=h.get("mom_col", nnz),
        mom_diff=h.get("mom_diff", nnz),
        color=h.get("color", nv),
        static_rhs=h.get("static_rhs", 3 * nv).reshape(3, nv),
        fixed_div=h.get("fixed_div", nc),
        init_rau=h.get("init_rau", nv),
        cell_z=h.get("cell_z", nc),
    )
    pipe = h.get("pipe", 5)
    c.zspan = float(pipe[4])
    c.pin = int(h.get("pin_cell", 1)[0])
    if with_hierarchy:
        c.hier = hierarchy_from_product(h)
    return c


def hierarchy_from_product(h) -> dict:
    nl = h.info["amg_levels"]
    H = {"P": [], "A": []}
    nfine = h.info["cells"]
    for l in range(nl):
        pr = h.get(f"P{l}_row", nfine + 1, np.int64)
        pc = h.get(f"P{l}_col", int(pr[-1]), np.int32)
        pv = h.get(f"P{l}_val", int(pr[-1]))
        ncoarse = int(pc.max()) + 1 if pc.size else 0
        ar = h.get(f"A{l}_row", ncoarse + 1, np.int64)
        ac = h.get(f"A{l}_col", int(ar[-1]), np.int32)
        av = h.get(f"A{l}_val", int(ar[-1]))
        H["P"].append(sp.csr_matrix((pv, pc, pr), shape=(nfine, ncoarse)))
        H["A"].append(sp.csr_matrix((av, ac, ar), shape=(ncoarse, ncoarse)))
        nfine = ncoarse
    H["Tinv"] = h.get("terminal_inv", nfine * nfine).reshape(nfine, nfine)
    H["fine_lambda"] = float(h.get("fine_lambda", 1)[0])
    H["level_lambda"] = h.get("level_lambda", nl)
    return H


# ---------------------------------------------------------------------------
# momentum
# ---------------------------------------------------------------------------


def cell_values(case: H8Case, U: np.ndarray) -> np.ndarray:
    """(nc,8,3) velocity at every cell basis: free DOFs from U (3,nv), fixed
    from the Dirichlet table (g5e_fp32_kernels.cuh:55-59)."""
    ref = case.cell_ref
    out = np.empty((case.nc, 8, 3))
    free = ref >= 0
    for d in range(3):
        out[:, :, d] = np.where(free, U[d][np.where(free, ref, 0)], case.fixed_value[np.where(free, 0, -ref - 1), d])
    return out


def assemble_central(case: H8Case, U: np.ndarray, T_cent: np.ndarray, extra: Optional[np.ndarray] = None):
    """Diffusion CSR + Picard central convection and its Dirichlet RHS.

    host_assemble_central_g4 (gpu/serial_cuda/g4_host.hpp:317-320) /
    h4_assemble_convection_only_kernel (h4_kernels.cuh:10-56):
    ur[m][j] = sum_d u_d(m) invJ[j][d], C_ab = det sum_{m,j} ur[m][j] T[a][m][b][j];
    free columns add into the row slot, fixed columns move to the RHS.
    extra (nc,8,8), when given, is added to every cell's block first (the
    eddy-viscosity diffusion of the coupled SST mode).
    Returns aval (nnz,), conv (3,nv), C (nc,8,8).
    """
    cf = cell_values(case, U)                                     # (nc,8,3)
    ur = np.einsum("cmd,cjd->cmj", cf, case.cell_invJ)            # (nc,8,3)
    C = np.einsum("cmj,ambj->cab", ur, T_cent) * case.cell_det[:, None, None]
    if extra is not None:
        C = C + extra
    aval = case.mom_diff.copy()
    conv = np.zeros((3, case.nv))
    ref = case.cell_ref
    ra = np.repeat(ref[:, :, None], 8, axis=2)                    # row dof of (c,a,b)
    rb = np.repeat(ref[:, None, :], 8, axis=1)                    # col dof
    both = (ra >= 0) & (rb >= 0)
    slot = case.cell_slot.reshape(case.nc, 8, 8).astype(np.int64)
    pos = case.mom_row[np.where(ra >= 0, ra, 0)] + slot
    np.add.at(aval, pos[both], C[both])
    fixed_b = (ra >= 0) & (rb < 0)
    fv = case.fixed_value[np.where(rb < 0, -rb - 1, 0)]           # (nc,8,8,3)
    for d in range(3):
        np.add.at(conv[d], ra[fixed_b], -C[fixed_b] * fv[fixed_b][:, d])
    return aval, conv, C


def finalize_relax(case: H8Case, aval: np.ndarray, alpha_u: float):
    """Row-L1 under-relaxation: delta_i = (1/alpha-1) sum_k |a_ik|, a_ii += delta_i,
    rAU_i = 1/a_ii (g4_finalize_relax_kernel, g5e_fp32_kernels.cuh:78-85;
    host_finalize_relax_g4, g4_host.hpp:422).  Modifies aval in place."""
    rows = np.repeat(np.arange(case.nv), np.diff(case.mom_row))
    m = np.bincount(rows, weights=np.abs(aval), minlength=case.nv)
    delta = (1.0 / alpha_u - 1.0) * m
    dp = case.diag_pos()
    aval[dp] += delta
    diag = aval[dp].copy()
    return delta, diag, 1.0 / diag


def bt_apply(case: H8Case, p: np.ndarray) -> np.ndarray:
    """Pressure gradient v_d[g] = sum_{(c,a): ref=g} B[c][a][d] p_c
    (h8_bt3_kernel, h8_precomputed_b.cuh:93-116)."""
    ref = case.cell_ref
    free = ref >= 0
    v = np.zeros((3, case.nv))
    contrib = case.bcoef * p[:, None, None]
    for d in range(3):
        np.add.at(v[d], ref[free], contrib[:, :, d][free])
    return v


def momentum_rhs(case: H8Case, conv, bt, delta, U):
    """b = s + c + B^T p + delta u (g4_momentum_rhs_kernel, g5e_fp32_kernels.cuh:97-104)."""
    return case.static_rhs + conv + bt + delta[None, :] * U


def mom_matrix(case: H8Case, aval: np.ndarray) -> sp.csr_matrix:
    return sp.csr_matrix((aval, case.mom_col, case.mom_row), shape=(case.nv, case.nv))


def gs_sweep(case: H8Case, A: sp.csr_matrix, diag: np.ndarray, b: np.ndarray, U: np.ndarray, omega: float,
             reverse: bool = False) -> np.ndarray:
    """One multicolour Gauss-Seidel sweep on all three components sharing A,
    colour by colour in ascending (or descending) order
    (forward_mcgs_fixed / backward_mcgs_fixed, h8_main_base.cu:1678-1699;
    g4_mcgs_color_kernel, g5e_fp32_kernels.cuh:106-116)."""
    U = U.copy()
    off = A.copy()
    off.setdiag(0.0)
    off.eliminate_zeros()
    ncol = int(case.color.max()) + 1
    order = range(ncol - 1, -1, -1) if reverse else range(ncol)
    for c in order:
        rows = np.flatnonzero(case.color == c)
        sub = off[rows]
        for d in range(3):
            s = sub @ U[d]
            z = (b[d, rows] - s) * (1.0 / diag[rows])
            U[d, rows] = (1.0 - omega) * U[d, rows] + omega * z
    return U


def momentum_audit(A: sp.csr_matrix, b: np.ndarray, U: np.ndarray) -> np.ndarray:
    """||b_d - (A u)_d|| / ||b_d|| (g4_mom_norms_kernel + h8_main_base.cu:1716-1726)."""
    out = np.zeros(3)
    for d in range(3):
        bn = np.linalg.norm(b[d]) or 1.0
        out[d] = np.linalg.norm(b[d] - A @ U[d]) / bn
    return out


def continuity(case: H8Case, U: np.ndarray) -> np.ndarray:
    """r_c = fixedDiv_c + sum_a B[c][a] . u(ref a) (h8_continuity_kernel,
    h8_precomputed_b.cuh:140-163)."""
    ref = case.cell_ref
    free = ref >= 0
    ug = np.zeros((case.nc, 8, 3))
    for d in range(3):
        ug[:, :, d] = np.where(free, U[d][np.where(free, ref, 0)], 0.0)
    return case.fixed_div + np.einsum("cad,cad->c", case.bcoef, ug)


# ---------------------------------------------------------------------------
# fine pressure Schur operator
# ---------------------------------------------------------------------------


def fine_topology(case: H8Case) -> dict:
    """P0 Schur pattern and refresh map (build_h2_fine_csr_host_reference,
    gpu/serial_cuda/h2b_fine_csr.cuh:57-149): free-velocity incidence in
    ascending (cell, basis) order, per-row directed contributions in
    (basis, incidence) order, sorted unique columns and uint8 slots."""
    if case._fine is not None:
        return case._fine
    ref = case.cell_ref
    cells = np.repeat(np.arange(case.nc), 8)
    basis = np.tile(np.arange(8), case.nc)
    g = ref.ravel()
    keep = g >= 0
    order = np.lexsort((basis[keep], cells[keep], g[keep]))  # by g, then cell, then basis
    inc_g = g[keep][order]
    inc_c = cells[keep][order]
    inc_a = basis[keep][order]
    inc_off = np.zeros(case.nv + 1, dtype=np.int64)
    np.add.at(inc_off, inc_g + 1, 1)
    inc_off = np.cumsum(inc_off)
    row = [0]
    cols: List[np.ndarray] = []
    slots: List[np.ndarray] = []
    contrib_off = [0]
    for c in range(case.nc):
        tgt = []
        for a in range(8):
            gg = ref[c, a]
            if gg < 0:
                continue
            tgt.append(inc_c[inc_off[gg]:inc_off[gg + 1]])
        t = np.concatenate(tgt)
        u = np.unique(t)
        cols.append(u)
        slots.append(np.searchsorted(u, t))
        row.append(row[-1] + u.size)
        contrib_off.append(contrib_off[-1] + t.size)
    topo = {
        "row": np.array(row, dtype=np.int64), "col": np.concatenate(cols).astype(np.int64),
        "inc_off": inc_off, "inc_c": inc_c, "inc_a": inc_a,
        "contrib_off": np.array(contrib_off, dtype=np.int64), "slot": np.concatenate(slots).astype(np.int64),
    }
    rows = np.repeat(np.arange(case.nc), np.diff(topo["row"]))
    topo["diag_pos"] = np.flatnonzero(topo["col"] == rows)
    case._fine = topo
    return topo


def fine_refresh(case: H8Case, rau: np.ndarray) -> Tuple[sp.csr_matrix, np.ndarray]:
    """Numeric B diag(rAU) B^T: val[row c, slot] += rAU_g B[c][a].B[c2][b]
    over every free DOF g of c and every (c2,b) incident to g
    (h2_fine_csr_refresh_kernel, h2b_fine_csr.cuh:465-501)."""
    T = fine_topology(case)
    ref = case.cell_ref
    val = np.zeros(T["col"].size)
    # directed contributions in (row c, basis a, incidence k) order
    rows_c, rows_a, ks = [], [], []
    for a in range(8):
        gg = ref[:, a]
        free = np.flatnonzero(gg >= 0)
        lens = T["inc_off"][gg[free] + 1] - T["inc_off"][gg[free]]
        rc = np.repeat(free, lens)
        starts = np.repeat(T["inc_off"][gg[free]], lens)
        within = np.arange(lens.sum()) - np.repeat(np.cumsum(lens) - lens, lens)
        rows_c.append(rc)
        rows_a.append(np.full(rc.size, a))
        ks.append(starts + within)
    rc = np.concatenate(rows_c)
    ra = np.concatenate(rows_a)
    kk = np.concatenate(ks)
    order = np.lexsort((np.arange(rc.size), rc))   # row-major, (a, k) kept
    rc, ra, kk = rc[order], ra[order], kk[order]
    c2 = T["inc_c"][kk]
    b = T["inc_a"][kk]
    g = ref[rc, ra]
    dot = np.einsum("nd,nd->n", case.bcoef[rc, ra], case.bcoef[c2, b])
    pos = T["row"][rc] + T["slot"]
    np.add.at(val, pos, rau[g] * dot)
    if case.pin >= 0:
        # closed domain: decouple the pinned cell's row and column, keep its
        # diagonal (the product's pressure pin)
        rows = np.repeat(np.arange(case.nc), np.diff(T["row"]))
        off = (rows != T["col"]) & ((rows == case.pin) | (T["col"] == case.pin))
        val[off] = 0.0
    A = sp.csr_matrix((val, T["col"], T["row"]), shape=(case.nc, case.nc))
    return A, val[T["diag_pos"]].copy()


def schur_matrix_free(case: H8Case, rau: np.ndarray, x: np.ndarray) -> np.ndarray:
    """y = B diag(rAU) B^T x (cpu_schur_apply, g2_host.hpp:77)."""
    v = bt_apply(case, x) * rau[None, :]
    return continuity(H8Case(**{**case.__dict__, "fixed_div": np.zeros(case.nc)}), v)


# ---------------------------------------------------------------------------
# AMG V-cycle and Krylov
# ---------------------------------------------------------------------------


@dataclass
class Smoother:
    kind: str = "jacobi"
    jacobi_omega: float = 0.7
    jacobi_fine_sweeps: int = 1
    jacobi_coarse_sweeps: int = 1
    cheb_degree: int = 2
    low_fraction: float = 0.05
    mcgs_omega: float = 1.0
    mcgs_fine_sweeps: int = 1
    mcgs_coarse_sweeps: int = 1
    mcgs_order: str = "symmetric"


def greedy_colors(A: sp.csr_matrix) -> np.ndarray:
    """Natural-order greedy colouring (g4_build_greedy_coloring,
    h8_main_base.cu:129-195).  Pure Python: small operators only."""
    n = A.shape[0]
    color = -np.ones(n, dtype=np.int64)
    ip, ix = A.indptr, A.indices
    for i in range(n):
        used = {int(color[j]) for j in ix[ip[i]:ip[i + 1]] if j != i and color[j] >= 0}
        c = 0
        while c in used:
            c += 1
        color[i] = c
    return color


def mcgs_plan(A: sp.csr_matrix, colors: np.ndarray):
    """Per-colour (rows, off-diagonal rows of A) of the multicolour sweep."""
    off = A.tocsr(copy=True)
    off.setdiag(0.0)
    ncol = int(colors.max()) + 1
    out = []
    for c in range(ncol):
        rows = np.flatnonzero(colors == c)
        out.append((rows, off[rows]))
    return out


def _smooth(A: sp.csr_matrix, diag, l1, b, opt: Smoother, lam: float, jac_sweeps: int, mcgs_sweeps: int,
            colors=None, plan=None) -> np.ndarray:
    """Smoother from x = 0 (G4AMG::jacobi_*/cheb_*/mcgs_*, h8_main_base.cu:429-519)."""
    if opt.kind in ("jacobi", "l1jacobi"):
        den = l1 if opt.kind == "l1jacobi" else diag
        x = opt.jacobi_omega * b / den
        for _ in range(1, jac_sweeps):
            x = x + opt.jacobi_omega * (b - A @ x) / den
        return x
    if opt.kind == "cheb2":
        lo = opt.low_fraction * lam
        centre, radius = 0.5 * (lam + lo), 0.5 * (lam - lo)
        root = lambda k: centre - radius * math.cos(math.pi * (2 * k + 1) / (2 * opt.cheb_degree))
        x = (1.0 / root(0)) * b / diag
        for k in range(1, opt.cheb_degree):
            x = x + (1.0 / root(k)) * (b - A @ x) / diag
        return x
    if opt.kind == "mcgs":
        x = np.zeros_like(b)
        if plan is None:
            plan = mcgs_plan(A, colors)
        passes = {"forward": [False], "backward": [True], "symmetric": [False, True]}[opt.mcgs_order]
        for _ in range(mcgs_sweeps):
            for rev in passes:
                for rows, off in (plan[::-1] if rev else plan):
                    gs = (b[rows] - off @ x) / diag[rows]
                    x[rows] = x[rows] + opt.mcgs_omega * (gs - x[rows])
        return x
    raise ValueError(opt.kind)


class VCycle:
    """G4AMG::apply / explicit_vcycle (gpu/serial_cuda/h8_main_base.cu:544-564)."""

    def __init__(self, hier: dict, opt: Smoother):
        self.H = hier
        self.opt = opt
        self.fine = None
        self.coarse_colors = []
        self.l1 = [np.asarray(abs(A).sum(axis=1)).ravel() for A in hier["A"]]
        self.coarse_plans = []
        if opt.kind == "mcgs":
            self.coarse_colors = [greedy_colors(A) for A in hier["A"][:-1]]
            self.coarse_plans = [mcgs_plan(A, c) for A, c in zip(hier["A"][:-1], self.coarse_colors)]

    def set_fine(self, A: sp.csr_matrix, diag: np.ndarray):
        self.fine = A
        self.fdiag = diag
        self.fl1 = np.asarray(abs(A).sum(axis=1)).ravel()
        self.fcolors = greedy_colors(A) if self.opt.kind == "mcgs" else None
        self.fplan = mcgs_plan(A, self.fcolors) if self.opt.kind == "mcgs" else None

    def _level(self, l, b):
        H, o = self.H, self.opt
        A = H["A"][l]
        if l == len(H["A"]) - 1:
            return H["Tinv"] @ b
        diag = A.diagonal()
        lam = float(H["level_lambda"][l]) if o.kind == "cheb2" else 1.0
        cols = self.coarse_colors[l] if o.kind == "mcgs" else None
        plan = self.coarse_plans[l] if o.kind == "mcgs" else None
        x = _smooth(A, diag, self.l1[l], b, o, lam, o.jacobi_coarse_sweeps, o.mcgs_coarse_sweeps, cols, plan)
        r = b - A @ x
        P = H["P"][l + 1]
        xc = self._level(l + 1, P.T @ r)
        x = x + P @ xc
        r = b - A @ x
        x = x + _smooth(A, diag, self.l1[l], r, o, lam, o.jacobi_coarse_sweeps, o.mcgs_coarse_sweeps, cols, plan)
        return x

    def apply(self, b: np.ndarray) -> np.ndarray:
        A, o = self.fine, self.opt
        lam = self.H["fine_lambda"]
        z = _smooth(A, self.fdiag, self.fl1, b, o, lam, o.jacobi_fine_sweeps, o.mcgs_fine_sweeps, self.fcolors,
                    self.fplan)
        r = b - A @ z
        P = self.H["P"][0]
        z = z + P @ self._level(0, P.T @ r)
        r = b - A @ z
        return z + _smooth(A, self.fdiag, self.fl1, r, o, lam, o.jacobi_fine_sweeps, o.mcgs_fine_sweeps, self.fcolors,
                           self.fplan)


def pcg(A, M, b, rtol, atol, maxit):
    """pressure_pcg (h8_main_base.cu:569-573) from x = 0."""
    x = np.zeros_like(b)
    r = b.copy()
    r0 = float(np.linalg.norm(r))
    if r0 <= atol:
        return x, {"its": 0, "rel": 0.0, "ok": True, "r0": r0}
    z = M(r)
    p = z.copy()
    rho = float(r @ z)
    target = max(atol, rtol * r0)
    out = {"its": 0, "rel": 1.0, "ok": False, "r0": r0}
    for k in range(maxit):
        q = A @ p
        alpha = rho / float(p @ q)
        x += alpha * p
        r -= alpha * q
        rn = float(np.linalg.norm(r))
        out.update(its=k + 1, rel=rn / r0)
        if rn <= target:
            out["ok"] = True
            break
        z = M(r)
        rhon = float(r @ z)
        p = z + (rhon / rho) * p
        rho = rhon
    return x, out


def richardson(A, M, b, rtol, atol, maxit, omega=1.0):
    """pressure_richardson (h8_main_base.cu:575-611)."""
    x = np.zeros_like(b)
    r = b.copy()
    r0 = float(np.linalg.norm(r))
    if r0 <= atol:
        return x, {"its": 0, "rel": 0.0, "ok": True, "r0": r0}
    target = max(atol, rtol * r0)
    out = {"its": 0, "rel": 1.0, "ok": False, "r0": r0}
    for k in range(maxit):
        z = M(r)
        x += omega * z
        r -= omega * (A @ z)
        rn = float(np.linalg.norm(r))
        out.update(its=k + 1, rel=rn / r0)
        if rn <= target:
            out["ok"] = True
            break
    return x, out


def precond_power(A, M, its, safety):
    """pressure_precond_power (h8_main_base.cu:613-668): power iteration on
    M A from v_i = sin(0.731 (i+1)) + 0.27 cos(1.117 (i+1))
    (g5_power_init_kernel, g5e_fp32_kernels.cuh:151-157), normalised each
    step; the estimate is the largest of |v.z| and |z| seen, times safety."""
    n = A.shape[0]
    g = np.arange(1, n + 1, dtype=np.float64)
    v = np.sin(0.731 * g) + 0.27 * np.cos(1.117 * g)
    v = v / math.sqrt(float(v @ v))
    est = 0.0
    for _ in range(its):
        z = M(A @ v)
        zn = math.sqrt(float(z @ z))
        ray = float(v @ z)
        est = max(est, max(abs(ray), zn))
        v = z / zn
    return safety * est


def cheb_roots(lam_max, low_fraction, degree):
    """Chebyshev points of [low_fraction * lam_max, lam_max] in the
    reference's order (cheb_root, h8_main_base.cu:424-428)."""
    lo = low_fraction * lam_max
    centre, radius = 0.5 * (lam_max + lo), 0.5 * (lam_max - lo)
    return [centre - radius * math.cos(math.pi * (2 * k + 1) / (2 * degree)) for k in range(degree)]


def cheb_poly(A, M, b, rtol, atol, maxit, lam_max, low_fraction, degree):
    """pressure_cheb_poly (h8_main_base.cu:670-726): preconditioned
    Richardson x += z / theta_k, r -= A z / theta_k with z = M r, the step
    sizes cycling through the Chebyshev points theta_k of
    [low_fraction lam_max, lam_max]."""
    x = np.zeros_like(b)
    r = b.copy()
    r0 = float(np.linalg.norm(r))
    if r0 <= atol:
        return x, {"its": 0, "rel": 0.0, "ok": True, "r0": r0}
    target = max(atol, rtol * r0)
    roots = cheb_roots(lam_max, low_fraction, degree)
    out = {"its": 0, "rel": 1.0, "ok": False, "r0": r0}
    for k in range(maxit):
        w = 1.0 / roots[k % degree]
        z = M(r)
        x += w * z
        r -= w * (A @ z)
        rn = float(np.linalg.norm(r))
        out.update(its=k + 1, rel=rn / r0)
        if rn <= target:
            out["ok"] = True
            break
    return x, out


def fgmres(A, M, b, rtol, atol, maxit, restart=30):
    """Right-preconditioned flexible GMRES(m) (textbook FGMRES, Saad 1993).
    A cycle whose least-squares residual reaches the target ends the solve:
    that residual equals |b - A x| up to FP64 rounding, because w = A z_j
    and its orthogonalisation are FP64 whatever M's precision.  Otherwise
    the true residual is recomputed for the restart."""
    n = b.size
    x = np.zeros(n)
    r = b.copy()
    beta = float(np.linalg.norm(r))
    r0 = beta
    out = {"its": 0, "rel": 1.0, "ok": False, "r0": r0}
    if beta <= atol:
        out.update(rel=0.0, ok=True)
        return x, out
    target = max(atol, rtol * r0)
    its = 0
    while True:
        V = np.zeros((restart + 1, n))
        Z = np.zeros((restart, n))
        H = np.zeros((restart + 1, restart))
        V[0] = r / beta
        k = 0
        est = None
        for j in range(restart):
            if its >= maxit:
                break
            Z[j] = M(V[j])
            w = A @ Z[j]
            for _ in range(2):
                h = V[:j + 1] @ w
                w = w - h @ V[:j + 1]
                H[:j + 1, j] += h
            H[j + 1, j] = np.linalg.norm(w)
            its += 1
            k = j + 1
            e1 = np.zeros(k + 1)
            e1[0] = beta
            y, *_ = np.linalg.lstsq(H[:k + 1, :k], e1, rcond=None)
            res = np.linalg.norm(e1 - H[:k + 1, :k] @ y)
            if res <= target or H[j + 1, j] == 0.0:
                est = res
                break
            V[j + 1] = w / H[j + 1, j]
        e1 = np.zeros(k + 1)
        e1[0] = beta
        y, *_ = np.linalg.lstsq(H[:k + 1, :k], e1, rcond=None)
        x = x + y @ Z[:k]
        if est is not None and est <= target:
            out.update(its=its, rel=est / r0, ok=True)
            return x, out
        r = b - A @ x
        beta = float(np.linalg.norm(r))
        out.update(its=its, rel=beta / r0)
        if beta <= target:
            out["ok"] = True
            return x, out
        if its >= maxit:
            return x, out


# ---------------------------------------------------------------------------
# SIMPLE loop
# ---------------------------------------------------------------------------


def pressure_drop_fit(case: H8Case, p: np.ndarray) -> float:
    """-slope * (zmax - zmin) of the least-squares fit of p over cell-centroid
    z (pressure_drop_fit_g4, g4_host.hpp:424)."""
    z = case.cell_z
    n = p.size
    sz, sp_, szz, szp = z.sum(), p.sum(), (z * z).sum(), (z * p).sum()
    den = n * szz - sz * sz
    slope = (n * szp - sz * sp_) / den
    return float(-slope * case.zspan)


def simple(case: H8Case, alpha_u=0.5, alpha_p=0.5, mom_omega=1.0, momentum_work="fgs1", p_solver="pcg",
           p_rtol=0.5, p_atol=1e-12, p_max=20, smoother: Optional[Smoother] = None, run_mode="fixed10",
           max_outer=10, simple_tol=1e-6, restart=30, U0=None, p0=None, callback=None, extra_blocks=None,
           mom_sweeps=1, cheb_degree=3, cheb_power_its=6, cheb_low_fraction=0.05, cheb_power_safety=1.15,
           extra_rhs=None):
    """H8 SIMPLE loop (gpu/serial_cuda/h8_main_base.cu:2046-2119): assemble,
    B^T p + RHS, fixed momentum work, continuity, fine refresh, pressure
    Krylov + AMG, p += alpha_p dp; converge mode audits the momentum residual
    once the continuity target is met.  extra_blocks(U), when given, adds
    cell blocks to the momentum assembly and extra_rhs(U) a (3,nv) term to
    its right-hand side (the DG inlet); callback(it, U, p) runs after the
    pressure update (the coupled SST mode advances k and omega there).
    mom_sweeps repeats the momentum work unit that many times per
    iteration, as the SST path's fixed M3 sweep count
    (NODALS_M3_SWEEPS, turbulence/sst_gpu_v2/source/
    puretet_profile5_eq5_hxt5b/modules/app/sst_application.inc:702;
    its sweeps are m3_fixed_coupled_csr, simple_pressure_predictor.inc:171)."""
    smoother = smoother or Smoother()
    _, T_cent = tensors()
    U = np.zeros((3, case.nv)) if U0 is None else U0.copy()
    p = np.zeros(case.nc) if p0 is None else p0.copy()
    vc = VCycle(case.hier, smoother)
    cheb_lam = None
    if p_solver == "cheb":
        # lambda_max of M A once, on the setup operator (h8_main_base.cu:1986-1990)
        Af0, fdiag0 = fine_refresh(case, case.init_rau)
        vc.set_fine(Af0, fdiag0)
        cheb_lam = precond_power(Af0, vc.apply, cheb_power_its, cheb_power_safety)
    cont0, cont_rel = None, 1.0
    hist = []
    converged = False
    for it in range(1, max_outer + 1):
        aval, conv, _ = assemble_central(case, U, T_cent, None if extra_blocks is None else extra_blocks(U))
        delta, diag, rau = finalize_relax(case, aval, alpha_u)
        bt = bt_apply(case, p)
        b = momentum_rhs(case, conv, bt, delta, U)
        if extra_rhs is not None:
            b = b + extra_rhs(U)
        A = mom_matrix(case, aval)
        audit = None
        if run_mode == "converge" and it > 1 and cont_rel <= simple_tol:
            audit = momentum_audit(A, b, U)
        for _ in range(mom_sweeps):
            if momentum_work == "fgs1":
                U = gs_sweep(case, A, diag, b, U, mom_omega)
            elif momentum_work == "sgs1":
                U = gs_sweep(case, A, diag, b, U, mom_omega)
                U = gs_sweep(case, A, diag, b, U, mom_omega, reverse=True)
            elif momentum_work == "altgs1":
                U = gs_sweep(case, A, diag, b, U, mom_omega, reverse=(it % 2 == 0))
        r = continuity(case, U)
        cn = float(np.linalg.norm(r))
        if cont0 is None:
            cont0 = cn
        cont_rel = cn / max(cont0, 1e-300)
        Af, fdiag = fine_refresh(case, rau)
        vc.set_fine(Af, fdiag)
        rhs = -r
        if case.pin >= 0:
            rhs[case.pin] = 0.0
        if p_solver == "pcg":
            dp, pr = pcg(Af, vc.apply, rhs, p_rtol, p_atol, p_max)
        elif p_solver == "richardson":
            dp, pr = richardson(Af, vc.apply, rhs, p_rtol, p_atol, p_max)
        elif p_solver == "fgmres":
            dp, pr = fgmres(Af, vc.apply, rhs, p_rtol, p_atol, p_max, restart)
        elif p_solver == "cheb":
            dp, pr = cheb_poly(Af, vc.apply, rhs, p_rtol, p_atol, p_max, cheb_lam, cheb_low_fraction, cheb_degree)
        else:
            raise ValueError(p_solver)
        p = p + alpha_p * dp
        hist.append({"it": it, "cont_rel": cont_rel, "p_its": pr["its"], "p_rel": pr["rel"]})
        if callback:
            callback(it, U, p)
        if audit is not None and cont_rel <= simple_tol and np.all(audit <= simple_tol):
            converged = True
            break
    return {"U": U, "p": p, "hist": hist, "converged": converged, "outer": it}
