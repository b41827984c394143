"""Pins of the k-omega transport-step oracle, and the product's host
topology of the step against it (no GPU).

The oracle's pieces are checked against what the mathematics fixes: the
twelve-point triangle rule against exact monomial integrals, the sample walk
by reconstructing every sample point, the inlet flux block by its exact
cancellation on the constant inlet field, the wall penalty by vanishing on a
field equal to its target, the row-L1 relaxation by keeping the fixed point
of the unrelaxed system, the sweeps by converging to a direct solve, the
colouring by properness.  A flipped flux sign, a transposed block, a wrong
relaxation delta or a wrong sweep update fails one of them.
"""
from math import factorial

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import h8_oracle as H
from oracle import refbind
from oracle import turb_oracle as T

pytestmark = pytest.mark.skipif(not refbind.sst_available(), reason="reference SST cores not built (oracle/_ref)")

BASE = ["--run-mode", "converge", "--verbose", "0"]


@pytest.fixture(scope="module")
def small(mesh_dir):
    from nodals_amd._lib import H8

    h = H8(["--mesh", mesh_dir["small"]] + BASE, device=False)
    h.turb_init([])
    m = T.mesh_from_product(h)
    case = H.case_from_product(h, with_hierarchy=False)
    st = T.Step(m, case, h.get("pipe_geometry", 12), T.Params(), refbind.RefSst())
    yield h, st
    h.close()


def plug(case):
    U = np.zeros((3, case.nv))
    U[2] = 1.0
    return U


def test_tri12_integrates_degree_6_exactly():
    w = T.TRI12[:, 3]
    L = T.TRI12[:, :3]
    assert abs(w.sum() - 1.0) < 1e-14
    for a in range(7):
        for b in range(7 - a):
            for c in range(7 - a - b):
                exact = 2.0 * factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 2)
                got = float((w * L[:, 0] ** a * L[:, 1] ** b * L[:, 2] ** c).sum())
                assert abs(got - exact) <= 1e-13 * max(exact, 1e-3), (a, b, c)
    # a degree-7 monomial is not integrated exactly (the rule is degree 6)
    got = float((w * L[:, 0] ** 7).sum())
    assert abs(got - 2.0 * factorial(7) / factorial(9)) > 1e-8


def test_walk_reconstructs_samples(small):
    _, st = small
    m = st.m
    for fi, (c, lf, n, _) in enumerate(st.wall):
        X = T.TRI12[:, :3] @ m.points[m.tets[c, T.TET_FACE[lf]]]
        for q in range(12):
            for k, y in enumerate(st.heights):
                cell, l = st.s_cell[fi, q, k], st.s_l[fi, q, k]
                assert l.min() >= -1e-8 and abs(l.sum() - 1.0) < 1e-12
                p = l @ m.points[m.tets[cell]]
                assert np.linalg.norm(p - (X[q] - y * n)) < 1e-12


def test_inlet_block_reproduces_rhs_on_inlet_state(small):
    _, st = small
    Uc = H.cell_values(st.case, plug(st.case))
    for eq, val in ((0, st.k_in), (1, st.omega_in)):
        for face in st.inlet[:8]:
            A, b = st.inlet_block(face, eq, Uc)
            x = np.r_[np.full(4, val), np.zeros(4)]
            assert np.abs(A @ x - b).max() <= 1e-12 * np.abs(b).max()
            # and the outflow direction alone does not cancel: a different value leaves a residual
            assert np.abs(A @ (1.5 * x) - b).max() > 1e-6 * np.abs(b).max()


def test_wall_block_vanishes_at_target(small):
    _, st = small
    Uc = H.cell_values(st.case, plug(st.case))
    coef, target, phi3, _, bad = st.wall_data(Uc)
    assert bad == 0 and (coef > 0).all() and (target > 0).all()
    for i in range(0, coef.size, 97):
        A = coef[i] * np.outer(phi3[i], phi3[i])
        b = coef[i] * target[i] * phi3[i]
        x = np.r_[np.full(4, target[i]), np.zeros(4)]
        assert np.abs(A @ x - b).max() <= 1e-12 * np.abs(b).max()


def _face_geometry(st, face):
    """Barycentric gradients, 3 volume / area and the face's vertex
    positions of a wall face's cell, from the vertex coordinates alone."""
    c, lf, n, area = face
    X = st.m.points[st.m.tets[c]]
    M = np.column_stack([np.ones(4), X])             # [1, x_a]
    grad = np.linalg.inv(M.T)[:, 1:]  # row a: grad lambda_a
    vol = abs(np.linalg.det(X[1:] - X[0])) / 6.0
    return grad, 3.0 * vol / area, X


def test_trace_block_closed_forms(small):
    """trace_log's symmetric Nitsche block on a wall face, summed over the
    face's twelve points, against the face integrals of the P1+BF3 basis:
    int lambda_a = A/3 on the face's vertices, int 27 l_j l_k l_l = 0.45 A
    for the face's own bubble, int lambda_j lambda_k = A/12.  A linear field
    equal to the target g on the face leaves mu s int phi_a (s its slope
    into the fluid); a constant field c leaves (c - g)(tau int phi_a - mu
    int dn phi_a), with tau = gamma mu / h and h = 3 volume / area."""
    _, st = small
    P = T.Params(omega_target="trace_log", trace_gamma=73.0)
    st2 = T.Step.__new__(T.Step)
    st2.__dict__.update(st.__dict__)
    st2.P = P
    mu, g, s, cval = 2.3e-3, 41.0, 5.5, 17.0
    for face in st.wall[:: max(1, len(st.wall) // 9)]:
        c, lf, n, area = face
        grad, h, X = _face_geometry(st, face)
        A = np.zeros((8, 8))
        b = np.zeros(8)
        for q in range(12):
            Aq, bq, ok = st2.trace_block(face, q, mu, g)
            assert ok
            A += Aq
            b += bq
        assert np.abs(A - A.T).max() <= 1e-13 * np.abs(A).max()
        fv = [j for j in range(4) if j != lf]
        int_phi = np.zeros(8)
        int_phi[fv] = area / 3.0
        int_phi[4 + lf] = 0.45 * area
        gn = grad @ n
        int_dn = np.zeros(8)
        int_dn[:4] = area * gn
        int_dn[4 + lf] = 27.0 * area / 12.0 * gn[fv].sum()
        for i in fv:
            int_dn[4 + i] = 27.0 * area / 12.0 * gn[lf]
        x0 = X[fv[0]]
        lin = np.r_[g - s * (X - x0) @ n, np.zeros(4)]
        assert np.allclose(A @ lin - b, mu * s * int_phi, rtol=0, atol=1e-11 * np.abs(b).max())
        tau = 73.0 * mu / h
        const = np.r_[np.full(4, cval), np.zeros(4)]
        want = (cval - g) * (tau * int_phi - mu * int_dn)
        assert np.allclose(A @ const - b, want, rtol=0, atol=1e-11 * np.abs(b).max())


def test_vector_fit_uses_the_tangential_speed(small):
    """The omega wall fit takes the speed tangential to the wall face.  On
    the straight pipe every wall face is parallel to the axis, so for an
    axial flow that speed is the axial speed |u_z| at each sample, and a
    radial component (along the wall normals, which are horizontal) leaves
    the speeds unchanged only where it is normal to the face."""
    _, st = small
    Uc = T.H.cell_values(st.case, plug(st.case))
    st.wall_data(Uc)
    cells = st.s_cell.reshape(-1)
    phi, _ = T.S.basis(st.s_l.reshape(-1, 4))
    uz = np.einsum("na,na->n", Uc[cells][:, :, 2], phi).reshape(-1, 5)
    assert max(abs(f[2][2]) for f in st.wall) < 1e-14
    assert np.allclose(st.last_speed, np.abs(uz), rtol=1e-14, atol=0)
    # the reference core's fit over 5 and over 1 active point both succeed
    out5, v5 = refbind.RefSst().fit_values(st.last_speed, np.broadcast_to(st.heights, uz.shape), st.nu, 5)
    out1, v1 = refbind.RefSst().fit_values(st.last_speed, np.broadcast_to(st.heights, uz.shape), st.nu, 1)
    assert v5.all() and v1.all() and (out5[:, 0] > 0).all() and (out1[:, 0] > 0).all()
    # with one point the fit is the pointwise Spalding root at the third sample
    root, ok = refbind.RefSst().spalding(st.last_speed[:, 2], np.full(uz.shape[0], st.heights[2]), st.nu)
    assert ok.all() and np.allclose(out1[:, 0], root[:, 0], rtol=1e-10)


def vertex_system(st, A, b):
    nv = st.m.nvert
    return A[:nv, :nv].tocsr(), b[:nv]


def test_relaxation_keeps_fixed_point(small):
    _, st = small
    U = plug(st.case)
    A, rhs = st.assemble(0, U)
    Av, bv = vertex_system(st, A, rhs)
    xs = spla.spsolve(Av.tocsc(), bv)
    x = np.zeros(st.m.ns)
    x[:st.m.nvert] = xs
    Ar, br, d = T.Step.relax(A, rhs, x, 0.3)
    Arv, brv = vertex_system(st, Ar, br)
    assert np.linalg.norm(Arv @ xs - brv) <= 1e-10 * np.linalg.norm(brv)
    # the delta added to the diagonal is (1/alpha - 1) times the full-row L1
    l1 = np.asarray(abs(A).sum(axis=1)).ravel()
    assert np.allclose(d - A.diagonal(), (1 / 0.3 - 1) * l1, rtol=1e-14, atol=0)
    # alpha = 1 leaves the system unchanged
    A1, b1, _ = T.Step.relax(A, rhs, x, 1.0)
    assert abs(A1 - A).max() == 0 and np.array_equal(b1, rhs)


def test_sweeps_converge_to_direct_solve(small):
    _, st = small
    U = plug(st.case)
    A, rhs = st.assemble(1, U)
    Ar, br, _ = T.Step.relax(A, rhs, st.w, 0.3)
    Arv, brv = vertex_system(st, Ar, br)
    xs = spla.spsolve(Arv.tocsc(), brv)
    x = st.sweeps(Ar, br, st.w, 300, 0.7)
    assert np.linalg.norm(x[:st.m.nvert] - xs) <= 1e-9 * np.linalg.norm(xs)
    assert np.all(x[st.m.nvert:] == 0.0)


def test_colouring_is_proper(small):
    _, st = small
    nv = st.m.nvert
    E = st.ent
    for c in range(E.shape[0]):
        v = E[c, :4]
        cols = st.colors[v]
        assert len(set(cols.tolist())) == 4  # the four vertices of a tet are mutually coupled


def test_product_topology_matches_oracle(small):
    h, st = small
    assert abs(h.turb["k_in"] - st.k_in) <= 1e-15 * st.k_in
    assert abs(h.turb["omega_in"] - st.omega_in) <= 1e-15 * st.omega_in
    assert h.turb["inlet_faces"] == len(st.inlet) and h.turb["wall_faces"] == len(st.wall)
    assert h.turb["nvert"] == st.m.nvert and h.turb["ns"] == st.m.ns
    assert np.allclose(h.get("turb_heights", 5), st.heights, rtol=1e-15, atol=0)
    n = st.s_cell.size
    assert np.array_equal(h.get("turb_sample_cell", n).astype(np.int64), st.s_cell.ravel())
    assert np.abs(h.get("turb_sample_l", 4 * n) - st.s_l.ravel()).max() < 1e-12
    assert np.array_equal(h.get("turb_color", st.m.nvert).astype(np.int64), st.colors)
    ent = h.get("turb_cell_ent", st.ent.size).astype(np.int64).reshape(-1, 8)
    assert np.array_equal(ent, st.ent)


def test_options_are_validated(mesh_dir):
    from nodals_amd._lib import H8, NodalsError

    h = H8(["--mesh", mesh_dir["small"]] + BASE, device=False)
    for bad in (["--sst-unknown", "1"], ["--sst-profile-fit-y1", "0.1"], ["--sst-gs-omega", "2.5"],
                ["--sst-omega-wall-mode", "log"], ["--sst-alpha-k", "0"]):
        with pytest.raises(NodalsError) as e:
            h.turb_init(bad)
        assert e.value.code == 1, bad
    h.turb_init(["--sst-profile-fit-y1", "0.01", "--sst-profile-fit-y2", "0.02", "--sst-profile-fit-y3", "0.03",
                 "--sst-profile-fit-y4", "0.04", "--sst-profile-fit-y5", "0.05", "--sst-omega-wall-mode", "none"])
    assert np.allclose(h.get("turb_heights", 5), [0.01, 0.02, 0.03, 0.04, 0.05])
    assert h.turb["wall_faces"] == 0
    h.close()
