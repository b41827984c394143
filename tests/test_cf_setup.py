"""CF/PMIS AMG setup: the oracle pinned by known answers and invariants, the
product's host hierarchy against the oracle (bitwise for integer data and
for arithmetic done in the same order).

Reference: gpu/serial_cuda/g8_cf_hierarchy_host.hpp (not compilable here: it
includes CUDA sources), tested by restatement in oracle/cf_oracle.py.
"""
import random

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import cf_oracle as CF


# ---------------------------------------------------------------- oracle pins
def test_splitmix64_published_vectors():
    assert CF.splitmix64(0) == 0xE220A8397B1DCDAF
    g = 0x9E3779B97F4A7C15
    seq = [CF.splitmix64((1234567 + k * g) & ((1 << 64) - 1)) for k in range(3)]
    assert seq == [6457827717110365317, 3203168211198807973, 9817491932198370423]


def _random_spd_graph(rng, n, p):
    A = sp.random(n, n, density=p, random_state=rng, data_rvs=lambda k: -rng.uniform(0.1, 1.0, k))
    A = A + A.T
    A = A.tolil()
    A.setdiag(0)
    A = A.tocsr()
    d = -np.asarray(A.sum(axis=1)).ravel() + 0.1
    return (A + sp.diags(d)).tocsr()


def test_pmis_equals_greedy_priority_mis_and_invariants():
    rng = np.random.RandomState(7)
    for trial in range(25):
        n = rng.randint(5, 80)
        A = _random_spd_graph(rng, n, rng.uniform(0.03, 0.3))
        A.sort_indices()
        theta = [0.1, 0.25, 0.5, 0.9][trial % 4]
        state, cid, rounds, strong, graph = CF.pmis_split(A, theta)
        assert np.array_equal(state, CF.greedy_mis_by_priority(graph))
        C = set(np.flatnonzero(state == 1))
        for i in range(n):
            if state[i] == 1:
                assert not any(j in C for j in graph[i])          # independent
            else:
                assert state[i] == 2 and any(j in C for j in graph[i])  # maximal
        assert sorted(cid[cid >= 0]) == list(range(len(C)))


def _lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()


def test_interpolation_hand_computed_rows():
    A = _lap1d(5)
    strong = CF.strength(A, 0.25)
    graph = CF.union_graph(strong)
    # split C F F C C  (coarse ids 0 . . 1 2)
    state = np.array([1, 2, 2, 1, 1], dtype=np.uint8)
    cid = np.array([0, -1, -1, 1, 2], dtype=np.int32)
    Pd = CF.cf_transfer(A, state, cid, strong, graph, 8, "direct").toarray()
    Pe = CF.cf_transfer(A, state, cid, strong, graph, 8, "exti").toarray()
    # direct: row 1 sees only C0, row 2 only C1
    assert np.array_equal(Pd[1], [1.0, 0.0, 0.0]) and np.array_equal(Pd[2], [0.0, 1.0, 0.0])
    # ext+i: row 1 adds path |a12| (-a23)/(-a23) = 1 to C1; row 2 adds 1 to C0
    assert np.array_equal(Pe[1], [0.5, 0.5, 0.0]) and np.array_equal(Pe[2], [0.5, 0.5, 0.0])
    for P in (Pd, Pe):
        assert np.array_equal(P[[0, 3, 4]], np.eye(3))
    # pmax 1 keeps the largest weight (ties to the smaller target)
    P1 = CF.cf_transfer(A, state, cid, strong, graph, 1, "exti").toarray()
    assert np.array_equal(P1[1], [1.0, 0.0, 0.0])


def test_interpolation_invariants_on_poisson():
    from nodals_amd import p1

    row, col, val = p1.kuhn_p1_laplacian(6)
    A = sp.csr_matrix((val, col, row))
    for interp in ("direct", "exti"):
        state, cid, rounds, strong, graph = CF.pmis_split(A, 0.25)
        P = CF.cf_transfer(A, state, cid, strong, graph, 4, interp)
        assert np.allclose(np.asarray(P.sum(axis=1)).ravel(), 1.0, rtol=0, atol=1e-14)
        assert np.diff(P.indptr).max() <= 4
        assert (P.data > 0).all()


# ---------------------------------------------------------------- product
def _h8_cf(mesh, interp="exti", pmax=8, aggressive=0, theta=0.25, terminal=1000):
    from nodals_amd._lib import H8

    return H8(["--mesh", mesh, "--max-outer", "10", "--amg-hierarchy", "cf", "--cf-interp", interp,
               "--cf-pmax", str(pmax), "--cf-aggressive-first", str(aggressive), "--cf-theta", str(theta),
               "--amg-terminal", str(terminal), "--amg-smoother", "jacobi", "--verbose", "0"], device=False)


def _product_fine(h):
    nc = h.info["cells"]
    fr = h.get("fine_row", nc + 1)
    fc = h.get("fine_col", int(fr[-1]))
    fv = h.get("exact_fine_val", int(fr[-1]))
    return sp.csr_matrix((fv, fc, fr.astype(np.int64)), shape=(nc, nc))


@pytest.mark.parametrize("interp,pmax", [("exti", 8), ("direct", 4)])
def test_product_cf_hierarchy_matches_oracle(mesh_dir, interp, pmax):
    from oracle import h8_oracle as O

    h = _h8_cf(mesh_dir["medium"], interp=interp, pmax=pmax, terminal=200)
    nc, nv = h.info["cells"], h.info["nv"]
    A = _product_fine(h)
    # exact fine CSR = B diag(rAU) B^T
    ref = CF.exact_fine_csr(h.get("bcoef", nc * 24), h.get("cell_ref", nc * 8), h.get("init_rau", nv), nv)
    assert abs(A - ref).max() <= 1e-13 * abs(ref).max()
    H = O.hierarchy_from_product(h)
    fine = A
    rounds = h.get("cf_rounds", len(H["P"]), np.int32)
    for l, (P, Ac) in enumerate(zip(H["P"], H["A"])):
        state, cid, rds, strong, graph = CF.pmis_split(fine, 0.25)
        assert np.array_equal(h.get(f"cf_state{l}", fine.shape[0], np.uint8), state)
        assert np.array_equal(h.get(f"cf_cid{l}", fine.shape[0], np.int32), cid)
        assert rounds[l] == rds
        Po = CF.cf_transfer(fine, state, cid, strong, graph, pmax, interp)
        assert np.array_equal(P.indptr, Po.indptr) and np.array_equal(P.indices, Po.indices)
        assert np.array_equal(P.data, Po.data)  # same operations in the same order
        G = (P.T @ fine @ P).tocsr()
        assert abs(Ac - G).max() <= 1e-12 * abs(G).max()
        fine = Ac
    assert H["A"][-1].shape[0] <= 200
    T = H["A"][-1].toarray()
    assert np.allclose(H["Tinv"] @ T, np.eye(T.shape[0]), atol=1e-9)
    h.close()


def test_product_cf_aggressive_first_composes_exactly(mesh_dir):
    from oracle import h8_oracle as O

    h = _h8_cf(mesh_dir["medium"], interp="exti", pmax=8, aggressive=1, terminal=100)
    nc = h.info["cells"]
    A = _product_fine(h)
    H = O.hierarchy_from_product(h)
    s0, c0, _, st0, g0 = CF.pmis_split(A, 0.25)
    assert np.array_equal(h.get("cf_state0", nc, np.uint8), s0)
    P01 = CF.cf_transfer(A, s0, c0, st0, g0, 8, "exti")
    A1 = (P01.T @ A @ P01).tocsr()
    if A1.shape[0] > 100:
        n1 = A1.shape[0]
        s1 = h.get("cf_state1", n1, np.uint8)
        # the product's A1 differs from the scipy product only in rounding;
        # the split of it must still be a valid PMIS split
        assert set(np.unique(s1)) <= {1, 2}
        P02 = H["P"][0]
        assert P02.shape == (nc, H["A"][0].shape[0])
        assert np.allclose(np.asarray(P02.sum(axis=1)).ravel(), 1.0, atol=1e-13)
    h.close()


def test_sys_cf_hierarchy_matches_oracle():
    from nodals_amd import p1
    from nodals_amd._lib import System

    row, col, val = p1.kuhn_p1_laplacian(8)
    A = sp.csr_matrix((val, col, row))
    S = System(row, col, val, "hierarchy=cf interp=exti pmax=4 smoother=jacobi terminal=64", device=False)
    H = S.hierarchy()
    fine = A
    for l, (P, Ac) in enumerate(zip(H["P"], H["A"])):
        state, cid, _, strong, graph = CF.pmis_split(fine, 0.25)
        assert np.array_equal(S.get(f"cf_state{l}", fine.shape[0], np.uint8), state)
        Po = CF.cf_transfer(fine, state, cid, strong, graph, 4, "exti")
        assert np.array_equal(P.indices, Po.indices) and np.array_equal(P.data, Po.data)
        G = (P.T @ fine @ P).tocsr()
        assert abs(Ac - G).max() <= 1e-12 * abs(G).max()
        fine = Ac
    S.close()
