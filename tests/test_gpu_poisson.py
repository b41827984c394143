"""Pressure-Poisson kernel bench system on the GPU vs the oracle: CSR SpMV,
the MCGS-smoothed SA V-cycle in FP64 / FP32 / BF16-finest storage, and
FGMRES(30) to rtol 1e-8 with the FP64 outer residual.

Tolerances: FP64 kernels 1e-12 relative; the FP32 V-cycle differs from the
FP64 oracle by FP32 rounding (1e-5 relative); the BF16-finest V-cycle by
BF16 storage of the finest values (3e-2 relative).  Whatever the hierarchy
precision, the FGMRES solution must meet the FP64 true residual 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from nodals_amd import p1
from oracle import h8_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def poisson():
    row, col, val = p1.kuhn_p1_laplacian(10)
    A = sp.csr_matrix((val, col, row))
    b = p1.poisson_rhs(A.shape[0])
    return row, col, val, A, b


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16", 3e-2)])
def test_vcycle_and_spmv(poisson, precision, tol):
    from nodals_amd._lib import System

    row, col, val, A, b = poisson
    print(f"\n[gpu] poisson 10^3 precision={precision}", flush=True)
    S = System(row, col, val, f"precision={precision} smoother=mcgs sweeps=2 order=forward terminal=100")
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    assert rel(S.spmv(x), A @ x) < 1e-12
    vc = O.VCycle(S.hierarchy(), O.Smoother(kind="mcgs", mcgs_fine_sweeps=2, mcgs_coarse_sweeps=2,
                                            mcgs_order="forward"))
    vc.set_fine(A, A.diagonal())
    assert rel(S.precond(b), vc.apply(b)) < tol
    S.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_fgmres_true_residual(poisson, precision):
    from nodals_amd._lib import System

    row, col, val, A, b = poisson
    S = System(row, col, val, f"precision={precision} smoother=mcgs sweeps=2 order=forward terminal=100")
    x, info = S.solve(b, "fgmres", 1e-8, 30, 300)
    assert info["ok"], info
    assert np.linalg.norm(b - A @ x) <= 1e-8 * np.linalg.norm(b) * 1.0001
    if precision == "fp64":
        vc = O.VCycle(S.hierarchy(), O.Smoother(kind="mcgs", mcgs_fine_sweeps=2, mcgs_coarse_sweeps=2,
                                                mcgs_order="forward"))
        vc.set_fine(A, A.diagonal())
        xo, io = O.fgmres(A, vc.apply, b, 1e-8, 0.0, 300, 30)
        assert abs(io["its"] - info["its"]) <= 1
        assert rel(x, xo) < 1e-6
    ms, info2 = S.solve_timed("fgmres", 1e-8, 30, 300, 2)
    assert info2["ok"] and info2["its"] == info["its"] and ms > 0
    for lvl in range(-1, S.levels - 1):
        t, by = S.time_spmv(lvl, 5)
        assert t > 0 and by > 0
    S.close()


def test_pcg_symmetric_mcgs(poisson):
    from nodals_amd._lib import System

    row, col, val, A, b = poisson
    S = System(row, col, val, "precision=fp64 smoother=mcgs sweeps=1 order=symmetric terminal=100")
    x, info = S.solve(b, "pcg", 1e-8, 30, 300)
    assert info["ok"] and np.linalg.norm(b - A @ x) <= 1.0001e-8 * np.linalg.norm(b)
    S.close()
