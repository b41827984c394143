"""Pins of the wall-function oracle (oracle/wall_oracle.py) and of the
product's host wall setup, on a small helix (host setup only, no GPU):

* the product's restatement of the tangential Spalding fit, built for the
  host, reproduces the reference's compiled fit (vector3::fit_active and
  active::fit_values) bit for bit for 1..5 active points, zero limit on and
  off, including stationary samples;
* the product's wall quadrature caches equal the oracle's independent
  computation;
* T makes every wall face's mean normal flux vanish, Q is a projection and
  T Q = T (what T is for, so a wrong 20/27 or sign fails);
* the Nitsche blocks' quadratic form is the integral of
  beta |u_t|^2 + gamma nu_eff / h (u.n)^2 over the wall, point by point;
* the implicit deviatoric tensor plus the scalar diffusion annihilates
  rigid rotations and uniform dilations (2 nu dev(sym grad u) = 0);
* the skew convection block satisfies C + C^T = boundary flux integral of
  (a.n) phi_a phi_b over the cell's faces;
* many block Gauss-Seidel sweeps (omega 0.7) reach the direct solve of K on
  the constrained space;
* the potential-flow start is the plug on a straight pipe (a linear
  potential, exact in P1), and the product's host solve equals the oracle's
  direct solve on the helix.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import h8_oracle as H
from oracle import refbind
from oracle import sst_oracle as S
from oracle import turb_oracle as TO
from oracle import wall_oracle as WO

pytestmark = pytest.mark.skipif(not (refbind.sst_available() and refbind.port_available()),
                                reason="reference SST cores / product host port not built (oracle/_ref)")


@pytest.fixture(scope="module")
def wcase(tmp_path_factory):
    from nodals_amd import meshgen
    from nodals_amd._lib import H8

    mesh = meshgen.write_helix_case(tmp_path_factory.mktemp("wo") / "helix", core=4, ring=2, nslab=12)
    h = H8(["--mesh", str(mesh), "--duct", "1", "--wall-model", "spalding", "--re", "5000", "--max-outer", "10",
            "--verbose", "0"], device=False)
    h.mesh_path = str(mesh)
    h.turb_init([])
    h.wall_init([])
    case = H.case_from_product(h)
    m = TO.mesh_from_product(h)
    st = TO.Step(m, case, h.get("pipe_geometry", 12), TO.Params(), refbind.RefSst(),
                 wall_distance=h.get("wall_distance", m.nvert))
    wp = WO.WallPath(case, st, refbind.RefSst())
    return h, case, st, wp


def _fit_inputs(n=400, seed=2):
    rng = np.random.default_rng(seed)
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    base = rng.standard_normal((n, 3)) * np.exp(rng.uniform(-6, 2, (n, 1)))
    samples = base[:, None, :] * (1 + 0.2 * rng.random((n, 5, 1))) * np.linspace(0.6, 1.4, 5)[None, :, None]
    samples[:20] = 0.0                      # stationary wall-parallel samples: u_tau = 0
    y = np.sort(rng.uniform(1e-5, 1e-2, (n, 5)), axis=1)
    pic = samples[:, 2] * 0.8 + 0.01 * rng.standard_normal((n, 3))
    return samples, y, nrm, pic, 1.46e-5


@pytest.mark.parametrize("count", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("zero_limit", [True, False])
def test_tangential_fit_port_is_reference_bitwise(count, zero_limit):
    s, y, n, pic, nu = _fit_inputs()
    ref, rv = refbind.RefSst().fit_tangential(s, y, n, pic, nu, count, zero_limit)
    got, gv = refbind.PortSst().fit_tangential(s, y, n, pic, nu, count, zero_limit)
    assert np.array_equal(rv, gv)
    m = rv.astype(bool)
    assert m.sum() > 0.9 * m.size
    assert np.array_equal(ref[m].view(np.int64), got[m].view(np.int64))


def test_host_wall_caches_equal_oracle(wcase):
    h, case, st, wp = wcase
    nf = wp.W.nf
    for name, arr, shape in [("wall_bw", wp.W.bw, (nf, 12, 8)), ("wall_bp", wp.W.bp, (nf, 12, 8)),
                             ("wall_qw", wp.W.qw, (nf, 12)), ("wall_h", wp.W.h, (nf,)),
                             ("wall_normal", wp.W.n, (nf, 3)), ("wall_ql", wp.W.ql, (nf, 12, 4))]:
        g = h.get(name, int(np.prod(shape))).reshape(shape)
        assert np.abs(g - arr).max() <= 1e-14 * max(np.abs(arr).max(), 1.0), name
    assert np.array_equal(h.get("wall_bub", nf).astype(np.int64), wp.W.bub)
    assert np.array_equal(h.get("wall_vert", 3 * nf).astype(np.int64).reshape(nf, 3), wp.W.vert)
    # the Picard functional and the trace are partitions of unity on the P1 part
    assert np.abs(wp.W.bw[:, :, :4].sum(2) - 1).max() < 1e-14
    assert np.abs(wp.W.bp[:, :, :4].sum(2) - 1).max() < 1e-14


def test_constraint_kills_wall_flux(wcase):
    h, case, st, wp = wcase
    W = wp.W
    rng = np.random.default_rng(4)
    x = rng.standard_normal(3 * case.nv)
    y = W.T @ x
    flux = np.einsum("fq,fqb,fbd,fd->f", W.qw, W.bw, H.cell_values(case, y.reshape(-1, 3).T)[W.cell], W.n)
    assert np.abs(flux).max() < 1e-13 * np.abs(x).max()
    assert abs(W.Q @ W.Q - W.Q).max() < 1e-15
    assert abs(W.T @ W.Q - W.T).max() < 1e-15
    # a planted mistake fails: 9/20 in place of 20/27
    Tbad = W.T.tolil()
    f = 0
    n, b = W.n[f], W.bub[f]
    for i in range(3):
        for j in range(3):
            for v in W.vert[f]:
                Tbad[3 * b + i, 3 * v + j] += (20 / 27 - 9 / 20) * n[i] * n[j]
    yb = Tbad.tocsr() @ x
    fb = np.einsum("q,qb,bd,d->", W.qw[f], W.bw[f], H.cell_values(case, yb.reshape(-1, 3).T)[W.cell[f]], W.n[f])
    assert abs(fb) > 1e-6 * np.abs(x).max()


def test_nitsche_blocks_quadratic_form(wcase):
    h, case, st, wp = wcase
    W = wp.W
    rng = np.random.default_rng(6)
    beta = np.exp(rng.uniform(-6, -2, W.nf))
    nue = np.exp(rng.uniform(-9, -4, (W.nf, 12)))
    wp.gamma_n = 3.0
    blk, pen, cpl = wp.nitsche(beta, nue)
    wp.gamma_n = 1e-6
    u = rng.standard_normal((W.nf, 8, 3))
    # block quadratic form: diagonal components plus n_i n_j cross coupling
    q = np.einsum("fad,fabd,fbd->f", u, blk, u)
    nn = np.einsum("fi,fj->fij", W.n, W.n) * (1 - np.eye(3))[None]
    q += np.einsum("fai,fab,fij,fbj->f", u, cpl, nn, u)
    # direct quadrature: beta |u_t|^2 + (3 nu_eff / h) (u.n)^2 at the wall points
    uq = np.einsum("fqa,fad->fqd", W.bw, u)
    un = np.einsum("fqd,fd->fq", uq, W.n)
    ut2 = (uq ** 2).sum(2) - un ** 2
    direct = (W.qw * (beta[:, None] * ut2 + 3.0 * nue / W.h[:, None] * un ** 2)).sum(1)
    assert np.abs(q - direct).max() < 1e-12 * np.abs(direct).max()
    # the penalty part is the component-diagonal share of the normal term,
    # sum_i n_i^2 u_i^2 (stage22d4b_nitsche_live.inc:294-300)
    qp = np.einsum("fad,fabd,fbd->f", u, pen, u)
    pd = np.einsum("fqd,fd->fq", uq ** 2, W.n ** 2)
    assert np.abs(qp - (W.qw * 3.0 * nue / W.h[:, None] * pd).sum(1)).max() < 1e-12 * np.abs(direct).max()


def test_deviatoric_tensor_annihilates_rigid_motion_and_dilation(wcase):
    h, case, st, wp = wcase
    U = wp.plug()
    dev = wp.dev_blocks(U)                                     # (nc,8,8,3,3)
    # scalar diffusion part with the same nu_eff: sum_q NW grad phi_a . grad phi_b
    Uc = H.cell_values(case, U)
    _, _, D = S.volume_terms_cells(wp.ref, Uc, st.k[st.ent], st.w[st.ent], case.cell_invJ, case.cell_det, st.cell_xy,
                                   st.R, st.nu, 2, st.P.k_floor, st.P.omega_floor, st.P.deviatoric, st._cell_vdist())
    Lq, Wq = S.keast24()
    _, gr = S.basis(Lq)
    gp = np.einsum("qaj,cjd->cqad", gr, case.cell_invJ)
    NW = D + case.cell_det[:, None] * Wq[None, :] * st.nu
    lap = np.einsum("cq,cqad,cqbd->cab", NW, gp, gp)
    full = dev + lap[:, :, :, None, None] * np.eye(3)[None, None, None]
    X = st.m.points[st.m.tets]                                 # (nc,4,3) vertex coordinates
    Om = np.array([[0.0, -0.3, 0.7], [0.3, 0.0, -1.1], [-0.7, 1.1, 0.0]])
    for field in (lambda x: x @ Om.T, lambda x: 0.9 * x):       # rotation, dilation
        ub = np.zeros((case.nc, 8, 3))
        ub[:, :4] = field(X)                                   # P1 coefficients; bubbles zero
        r = np.einsum("cabij,cbj->cai", full, ub)
        assert np.abs(r).max() < 1e-10 * np.abs(full).max() * np.abs(ub).max()
    # a planted mistake fails: dropping the transpose term
    bad = -(2.0 / 3.0) * np.einsum("cq,cqai,cqbj->cabij", NW, gp, gp) + lap[:, :, :, None, None] * np.eye(3)
    ub[:, :4] = X @ Om.T
    assert np.abs(np.einsum("cabij,cbj->cai", bad, ub)).max() > 1e-3 * np.abs(full).max() * np.abs(ub).max()


def test_skew_convection_is_skew_up_to_boundary_flux(wcase):
    h, case, st, wp = wcase
    rng = np.random.default_rng(8)
    U = rng.standard_normal((3, case.nv))
    Uc = H.cell_values(case, U)
    ur = np.einsum("cmd,cjd->cmj", Uc, case.cell_invJ)
    C = np.einsum("cmj,ambj->cab", ur, wp.Tt) * case.cell_det[:, None, None]
    sym = C + np.transpose(C, (0, 2, 1))
    # boundary integral of (a.n) phi_a phi_b over the four faces: degree 9 in
    # the face, integrated exactly by a collapsed 8x8 Gauss rule (degree 14)
    g, gw = np.polynomial.legendre.leggauss(8)
    g, gw = 0.5 * (g + 1), 0.5 * gw
    X, Y = np.meshgrid(g, g, indexing="ij")
    WX, WY = np.meshgrid(gw, gw, indexing="ij")
    tx, ty = X.ravel(), (Y * (1 - X)).ravel()
    tw = (WX * WY * (1 - X)).ravel()                           # sums to 1/2
    bary = np.column_stack([1 - tx - ty, tx, ty])
    m = st.m
    cells = np.arange(0, case.nc, 97)
    for c in cells:
        P = m.points[m.tets[c]]
        ref = np.zeros((8, 8))
        for lf in range(4):
            L = np.zeros((bary.shape[0], 4))
            L[:, TO.TET_FACE[lf]] = bary
            phi, _ = S.basis(L)
            v = P[TO.TET_FACE[lf]]
            nrm = np.cross(v[1] - v[0], v[2] - v[0])
            area = 0.5 * np.linalg.norm(nrm)
            nrm /= 2 * area
            if np.dot(nrm, v[0] - P[lf]) < 0:
                nrm = -nrm
            a = phi @ Uc[c]                                    # carrier at the points
            ref += np.einsum("q,qa,qb->ab", 2 * tw * area * (a @ nrm), phi, phi)
        assert np.abs(sym[c] - ref).max() < 1e-12 * max(1.0, np.abs(C[c]).max())


def test_block_sweeps_reach_constrained_direct_solve(wcase):
    h, case, st, wp = wcase
    U = wp.plug()
    A = wp.assemble(U, np.zeros(case.nc))
    W = wp.W
    # reduced space: Q's range (every coordinate but the wall bubbles' normal parts)
    keep = sp.csr_matrix(W.Q)
    K, b = A["K"], A["b"]
    # direct: solve Q K Q x = Q b on the range of Q (add the projector onto null(Q) to make it regular)
    N = sp.identity(3 * case.nv, format="csr") - keep
    x_direct = sp.linalg.spsolve((keep @ K @ keep + N).tocsc(), keep @ b)
    # (omega 0.7 as the reference runs it; at omega 1 the block sweep of this
    # non-symmetric operator diverges)
    wp2 = WO.WallPath(case, st, wp.ref, m3_sweeps=300, m3_omega=0.7)
    A2 = dict(A)
    A2["U"] = np.zeros((3, case.nv))
    Ux = wp2.momentum(A2)
    x_direct_phys = (W.T @ x_direct).reshape(case.nv, 3).T
    assert max(np.linalg.norm(Ux[d] - x_direct_phys[d]) / np.linalg.norm(x_direct_phys[d]) for d in range(3)) < 1e-8


def test_potential_start_straight_pipe_is_the_plug(tmp_path):
    """On a straight pipe the potential is linear along the axis, which P1
    represents exactly: the start is the plug on every free vertex, with
    zero face bubbles (closed form).  The product's host solve and the
    oracle's direct solve both give it; a reversed inlet flux gives the
    negated plug, so the test sees the sign of the source."""
    from nodals_amd import meshgen
    from nodals_amd._lib import H8

    mesh = meshgen.write_pipe_case(tmp_path / "pipe", core=4, ring=2, nslab=10)
    h = H8(["--mesh", str(mesh), "--duct", "1", "--wall-model", "spalding", "--re", "5000", "--max-outer", "10",
            "--verbose", "0"],
           device=False)
    h.turb_init([])
    h.wall_init(["--potential-tol", "1e-13"])
    case = H.case_from_product(h, with_hierarchy=False)
    m = TO.mesh_from_product(h)
    geo = h.get("pipe_geometry", 12)
    g2 = h.get("g2free", m.ns, np.int32)
    U, ent, rep = WO.potential_start(m, geo, g2, case.nv)
    plug = np.asarray(geo[6:9])
    npt = m.nvert
    vfree = g2[:npt] >= 0
    assert np.abs(ent[:npt] - plug[None, :]).max() < 1e-9
    assert np.abs(ent[npt:]).max() < 1e-8
    got = h.get("wall_u0", 3 * case.nv).reshape(case.nv, 3)
    assert np.abs(got[g2[:npt][vfree]] - plug[None, :]).max() < 1e-9
    start = h.get("wall_start", 7)
    assert start[0] > 0 and start[1] < 1e-13 and start[4] < 1e-10 and start[5] < 1e-10
    flipped = geo.copy()
    flipped[6:9] = -flipped[6:9]
    _, ent2, _ = WO.potential_start(m, flipped, g2, case.nv)
    assert np.abs(ent2[:npt] + plug[None, :]).max() < 1e-9


def test_potential_start_product_matches_oracle(wcase):
    """The product's host potential start (Jacobi PCG to 1e-10) against the
    oracle's direct solve on the helix; every wall face's mean normal
    velocity is zero (T leaves it unchanged) and the outlet carries the
    inlet flux."""
    h, case, st, wp = wcase
    m = st.m
    U, ent, rep = WO.potential_start(m, h.get("pipe_geometry", 12), h.get("g2free", m.ns, np.int32), case.nv)
    got = h.get("wall_u0", 3 * case.nv).reshape(case.nv, 3).T
    assert np.linalg.norm(got - U) / np.linalg.norm(U) < 1e-7
    assert np.linalg.norm(wp.constrain(U) - U) < 1e-12 * np.linalg.norm(U)
    start = h.get("wall_start", 7)
    assert start[1] < 1e-10 and start[4] < 1e-10 and start[5] < 1e-10
    assert abs(start[2] - rep["inlet_flux"]) < 1e-12 * abs(rep["inlet_flux"])


def test_wall_divergence_drops_the_wall_normal_flux(wcase):
    """With wall-function walls the product's B is B_N = B - int_wall (n.u):
    its coefficients equal the no-slip context's volume B on the same mesh
    minus the oracle's 12-point quadrature of every wall face's traces times
    the outward normal; the face's own bubble then drops out of B_N, and on
    velocities that satisfy the wall-flux constraint B_N = B."""
    from nodals_amd._lib import H8

    h, case, st, wp = wcase
    W = wp.W
    h0 = H8(["--mesh", h.mesh_path, "--duct", "1", "--re", "5000", "--max-outer", "10", "--verbose", "0"], device=False)
    nc = case.nc
    b_vol = h0.get("bcoef", nc * 24).reshape(nc, 8, 3)
    h0.close()
    I = np.einsum("fq,fqa->fa", W.qw, W.bw)                             # (nf, 8): integral of each trace
    patch = np.zeros_like(b_vol)
    np.add.at(patch, W.cell, I[:, :, None] * W.n[:, None, :])
    assert np.abs(case.bcoef - (b_vol - patch)).max() < 1e-14 * np.abs(b_vol).max()
    assert np.abs(case.bcoef - (b_vol + patch)).max() > 1e-3 * np.abs(b_vol).max()     # a flipped sign shows
    own_bubble = case.bcoef[W.cell, 4 + W.lf]
    assert np.abs(own_bubble).max() < 1e-14 * np.abs(b_vol).max()
    # on constrained velocities the two divergences agree
    rng = np.random.default_rng(4)
    x = wp.constrain(rng.standard_normal((3, case.nv)))
    ug = H.cell_values(case, x)
    div_n = np.einsum("cad,cad->c", case.bcoef, ug)
    div_v = np.einsum("cad,cad->c", b_vol, ug)
    assert np.abs(div_n - div_v).max() < 1e-13 * np.abs(div_v).max()
