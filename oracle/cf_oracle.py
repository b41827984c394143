"""CPU oracle of the CF/PMIS AMG setup (classical negative-coupling strength,
PMIS split, direct / extended+i interpolation, exact transfer composition).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

Restates gpu/serial_cuda/g8_cf_hierarchy_host.hpp:
  splitmix64            g8_hash64 (:90-95)
  priority_key          g8_priority_higher (:97-106)
  pmis_split            g8_build_pmis_split_reference (:169-293)
  cf_transfer           g8_prune_normalize + g8_build_transfer_reference (:625-785)
  compose               g11_compose_transfers_exact (:1100-1195)
  exact_fine_csr        g8_build_exact_fine_csr (:116-166)

How each is pinned (tests/test_cf_setup.py):
  splitmix64    published SplitMix64 test vectors (seed 0 -> 0xE220A8397B1DCDAF;
                seed 1234567 -> 6457827717110365317, 3203168211198807973, ...).
  pmis_split    equals the sequential greedy maximal independent set taken
                in descending priority order (a different algorithm that
                yields the same set for a fixed total order: the
                lexicographically-first MIS), plus independence and
                maximality on the symmetric strong graph.
  cf_transfer   hand-computed rows on a 1-D Laplacian (direct and ext+i
                paths), unit row sums, C-row injection, pmax.
  compose       scipy sparse product.
  exact_fine_csr  equals B diag(rAU) B^T assembled by scipy from the B table.
The reference header itself includes CUDA sources and cannot be compiled
here, so these are the pins.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np
import scipy.sparse as sp

_M64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """SplitMix64 output for state x (g8_cf_hierarchy_host.hpp:90-95)."""
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def strength(A: sp.csr_matrix, theta: float) -> List[List[int]]:
    """Strong rows: j != i, a_ij < 0, -a_ij >= theta * max_k(-a_ik)
    (g8_cf_hierarchy_host.hpp:174-205)."""
    A = A.tocsr()
    out: List[List[int]] = []
    for i in range(A.shape[0]):
        cols = A.indices[A.indptr[i]:A.indptr[i + 1]]
        vals = A.data[A.indptr[i]:A.indptr[i + 1]]
        off = cols != i
        neg = off & (vals < 0.0)
        mx = float(np.max(-vals[neg])) if np.any(neg) else 0.0
        thr = theta * mx
        if thr > 0.0:
            sel = neg & (-vals >= thr)
            out.append(sorted(int(c) for c in cols[sel]))
        else:
            out.append([])
    return out


def union_graph(strong: List[List[int]]) -> List[List[int]]:
    g: List[set] = [set() for _ in strong]
    for i, s in enumerate(strong):
        for j in s:
            g[i].add(j)
            g[j].add(i)
    return [sorted(x) for x in g]


def priority_key(i: int, graph: List[List[int]]) -> Tuple[int, int, int]:
    """Larger key = higher priority: (union degree, hash(i + 0x63021), i)."""
    return (len(graph[i]), splitmix64(i + 0x63021), i)


def pmis_split(A: sp.csr_matrix, theta: float):
    """PMIS rounds: every undecided row with no undecided higher-priority
    neighbour becomes C, then the undecided neighbours of the new C rows
    become F (g8_cf_hierarchy_host.hpp:207-250).  Returns (state, coarse_id,
    rounds, strong, graph) with state 1 = C, 2 = F."""
    strong = strength(A, theta)
    graph = union_graph(strong)
    n = len(graph)
    key = [priority_key(i, graph) for i in range(n)]
    state = np.zeros(n, dtype=np.uint8)
    rounds = 0
    undecided = n
    while undecided > 0:
        winners = [i for i in range(n) if state[i] == 0
                   and all(not (state[j] == 0 and key[j] > key[i]) for j in graph[i])]
        if not winners:
            raise RuntimeError("PMIS no progress")
        for i in winners:
            state[i] = 1
            undecided -= 1
        for i in winners:
            for j in graph[i]:
                if state[j] == 0:
                    state[j] = 2
                    undecided -= 1
        rounds += 1
    cid = np.full(n, -1, dtype=np.int32)
    c = np.flatnonzero(state == 1)
    cid[c] = np.arange(c.size, dtype=np.int32)
    return state, cid, rounds, strong, graph


def greedy_mis_by_priority(graph: List[List[int]]) -> np.ndarray:
    """Sequential greedy maximal independent set in descending priority
    order: 1 for members, 2 otherwise.  Independent check of pmis_split."""
    n = len(graph)
    order = sorted(range(n), key=lambda i: priority_key(i, graph), reverse=True)
    state = np.zeros(n, dtype=np.uint8)
    for i in order:
        if state[i] != 0:
            continue
        state[i] = 1
        for j in graph[i]:
            if state[j] == 0:
                state[j] = 2
    return state


def _entry(A: sp.csr_matrix, i: int, j: int) -> float:
    cols = A.indices[A.indptr[i]:A.indptr[i + 1]]
    k = np.searchsorted(cols, j)
    if k < cols.size and cols[k] == j:
        return float(A.data[A.indptr[i] + k])
    return 0.0


def cf_transfer(A: sp.csr_matrix, state, cid, strong, graph, pmax: int, interp: str) -> sp.csr_matrix:
    """Interpolation rows (g8_cf_hierarchy_host.hpp:647-785): C rows inject;
    F rows take |a_ij| of strong C neighbours, with ext+i also path weights
    |a_ik| (-a_kj) / sum_j(-a_kj) through strong F neighbours k, falling back
    to the union-graph C neighbours; candidates ordered by (weight desc,
    target asc), cut to pmax, normalised in that order, stored by target."""
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    nC = int((state == 1).sum())
    rows, cols, vals = [], [], []
    for i in range(n):
        if state[i] == 1:
            rows.append(i)
            cols.append(int(cid[i]))
            vals.append(1.0)
            continue
        acc: Dict[int, float] = {}

        def add(c: int, w: float):
            acc[c] = acc.get(c, 0.0) + w

        for j in strong[i]:
            if state[j] != 1:
                continue
            a = _entry(A, i, j)
            w = -a if a < 0.0 else abs(a)
            if w > 0.0:
                add(int(cid[j]), w)
        if interp == "exti":
            for k in strong[i]:
                if state[k] != 2:
                    continue
                aik = _entry(A, i, k)
                path = -aik if aik < 0.0 else 0.0
                if not path > 0.0:
                    continue
                denom = 0.0
                for j in strong[k]:
                    if state[j] != 1:
                        continue
                    akj = _entry(A, k, j)
                    if akj < 0.0:
                        denom += -akj
                if not denom > 0.0:
                    continue
                for j in strong[k]:
                    if state[j] != 1:
                        continue
                    akj = _entry(A, k, j)
                    if akj < 0.0:
                        add(int(cid[j]), path * ((-akj) / denom))
        if not acc:
            for j in graph[i]:
                if state[j] != 1:
                    continue
                w = abs(_entry(A, i, j))
                if w > 0.0:
                    add(int(cid[j]), w)
        if not acc:
            raise RuntimeError("interpolation row has no C target")
        cand = [(c, w) for c, w in acc.items() if w > 0.0]
        cand.sort(key=lambda e: (-e[1], e[0]))
        if pmax > 0 and len(cand) > pmax:
            cand = cand[:pmax]
        s = 0.0
        for _, w in cand:
            s += w
        cand = sorted(((c, w / s) for c, w in cand), key=lambda e: e[0])
        for c, w in cand:
            rows.append(i)
            cols.append(c)
            vals.append(w)
    return sp.csr_matrix((np.array(vals), (np.array(rows), np.array(cols))), shape=(n, nC))


def compose(P01: sp.csr_matrix, P12: sp.csr_matrix) -> sp.csr_matrix:
    """Exact P01 P12 (g8_cf_hierarchy_host.hpp:1100-1195)."""
    return (P01 @ P12).tocsr()


def exact_fine_csr(bcoef: np.ndarray, cell_ref: np.ndarray, rau: np.ndarray, nv: int) -> sp.csr_matrix:
    """B diag(rAU) B^T with B[c, 3 g + d] = bcoef[c, a, d] for free DOF g of
    cell c (g8_cf_hierarchy_host.hpp:116-166 as a matrix identity)."""
    ref = cell_ref.reshape(-1, 8)
    nc = ref.shape[0]
    b = bcoef.reshape(nc, 8, 3)
    rows, cols, vals = [], [], []
    for a in range(8):
        g = ref[:, a]
        m = g >= 0
        for d in range(3):
            rows.append(np.flatnonzero(m))
            cols.append(3 * g[m] + d)
            vals.append(b[m, a, d])
    B = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nc, 3 * nv))
    R = sp.diags(np.repeat(rau, 3))
    return (B @ R @ B.T).tocsr()
