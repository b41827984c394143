"""Assembled P1 Laplacian on a Kuhn cube: the pressure-Poisson kernel bench.

Operator: nodal P1 stiffness sum_T vol_T grad(lambda_a) . grad(lambda_b) over
every tetrahedron, CSR over the full edge graph (the Kuhn diagonal edges
carry exact or round-off zeros and stay in the pattern), with a symmetric
Dirichlet pin at one vertex (row and column zeroed, unit diagonal).  RHS:
seeded N(0,1) nodal values (MT19937, seed 20261202), mean removed, zero at
the pin.  64^3 hexes: 274,625 rows, 4,018,753 nonzeros.
"""
from __future__ import annotations

from typing import Tuple

import numpy as np
import scipy.sparse as sp

from . import meshgen

SEED = 20261202


def kuhn_p1_laplacian(n: int, pin: int = 0) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(row int64, col int32, val float64) of the pinned P1 Laplacian."""
    pts, tets = meshgen.kuhn_cube(n)
    X = pts[tets]                                        # (nt,4,3)
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)   # columns
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)                              # rows = grad lambda_1..3
    G = np.concatenate([-Jinv.sum(axis=1, keepdims=True), Jinv], axis=1)   # (nt,4,3)
    K = np.einsum("tad,tbd->tab", G, G) * (np.abs(det) / 6.0)[:, None, None]
    rows = np.repeat(tets, 4, axis=1).ravel()
    cols = np.tile(tets, (1, 4)).ravel()
    npt = pts.shape[0]
    A = sp.coo_matrix((K.ravel(), (rows, cols)), shape=(npt, npt)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    # symmetric Dirichlet pin, pattern kept
    r0, r1 = A.indptr[pin], A.indptr[pin + 1]
    A.data[r0:r1] = 0.0
    A.data[A.indices == pin] = 0.0
    A.data[r0 + np.flatnonzero(A.indices[r0:r1] == pin)[0]] = 1.0
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64)


def poisson_rhs(npts: int, pin: int = 0, seed: int = SEED) -> np.ndarray:
    rng = np.random.Generator(np.random.MT19937(seed))
    b = rng.standard_normal(npts)
    b[pin] = 0.0
    free = np.ones(npts, dtype=bool)
    free[pin] = False
    b[free] -= b[free].mean()
    return b
