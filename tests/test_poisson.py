"""Pressure-Poisson kernel bench operator (P1 Laplacian on a Kuhn cube) and
its SA hierarchy (CPU)."""
import numpy as np
import pytest
import scipy.sparse as sp

from nodals_amd import p1
from oracle import refbind


def test_kuhn_p1_is_seven_point_stencil():
    """On the Kuhn triangulation the P1 stiffness matrix is h times the 7-point
    stencil: interior rows have 15 stored entries (14 edge neighbours), 6h on
    the diagonal, -h on the 6 axis neighbours and 0 on the 8 diagonal edges."""
    n = 6
    row, col, val = p1.kuhn_p1_laplacian(n, pin=0)
    A = sp.csr_matrix((val, col, row))
    h = 1.0 / n
    m = n + 1
    i, j, k = 3, 2, 4
    v = i + m * (j + m * k)
    cols = col[row[v]:row[v + 1]]
    vals = val[row[v]:row[v + 1]]
    assert cols.size == 15
    for c, a in zip(cols, vals):
        d = abs(c - v)
        if c == v:
            assert abs(a - 6 * h) < 1e-14
        elif d in (1, m, m * m):
            assert abs(a + h) < 1e-14
        else:
            assert abs(a) < 1e-14
    # symmetric, pinned row/column
    assert abs(A - A.T).max() < 1e-15
    assert A[0, 0] == 1.0 and A[0].nnz == cols.size - 15 + A[0].nnz and np.all(A[0].data[A[0].indices != 0] == 0)


def test_config_sizes():
    row, col, val = p1.kuhn_p1_laplacian(64)
    assert row.size - 1 == 274625 and row[-1] == 4018753
    b = p1.poisson_rhs(row.size - 1)
    assert b[0] == 0.0 and abs(b.mean()) < 1e-15 * 10


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built")
def test_sa_hierarchy_matches_reference_functions():
    from nodals_amd._lib import System

    row, col, val = p1.kuhn_p1_laplacian(12)
    S = System(row, col, val, "smoother=mcgs sweeps=2 terminal=200", device=False)
    R = refbind.RefCase.blank()
    R.sa_from_csr(row, col, val, terminal=200)
    dims = R.get("csr_dims", np.int32)
    assert S.levels == dims[0] and S.terminal == dims[1]
    assert np.array_equal(R.get("csr_agg0", np.int32), S.get("agg0", row.size - 1, np.int32))
    nf = row.size - 1
    for l in range(S.levels):
        pr = R.get(f"csr_P{l}_row", np.int64)
        assert np.array_equal(pr, S.get(f"P{l}_row", nf + 1, np.int64))
        assert np.array_equal(R.get(f"csr_P{l}_col", np.int32), S.get(f"P{l}_col", int(pr[-1]), np.int32))
        assert np.array_equal(R.get(f"csr_P{l}_val", np.float64), S.get(f"P{l}_val", int(pr[-1])))
        ar = R.get(f"csr_A{l}_row", np.int64)
        nc = ar.size - 1
        assert np.array_equal(ar, S.get(f"A{l}_row", nc + 1, np.int64))
        assert np.array_equal(R.get(f"csr_A{l}_col", np.int32), S.get(f"A{l}_col", int(ar[-1]), np.int32))
        assert np.array_equal(R.get(f"csr_A{l}_val", np.float64), S.get(f"A{l}_val", int(ar[-1])))
        nf = nc
    t = R.get("csr_terminal_inv", np.float64)
    assert np.array_equal(t, S.get("terminal_inv", t.size))
