"""CPU oracle of the learned pressure-correction network (forward, loss and
gradients), with the device's BF16 rounding points.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker; never by
the product package.

The network has no counterpart in the reference (its only data-driven
pressure aid is the snapshot POD of
turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/pod_balanced.inc),
so there is nothing of the reference to follow: this restates the network
the product defines,
    h0 = relu(bf(x) bf(W0) + b0)
    h_l = h_{l-1} + relu(bf([h_{l-1} | mean_N h_{l-1}]) bf(W_l) + b_l)
    y = bf(h4) bf(w) + b + x[:, 2],   loss = mean((y - t)^2)
with bf() rounding to BF16 where the device feeds its matrix cores, and its
backward pass (bf() applied to both GEMM operands there as well).  Pinned by
central finite differences of the loss on a small graph with rounding
switched off (tests/test_gnn_oracle.py): a transposed weight, a wrong
aggregation transpose or a missing ReLU mask fails that check.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

IN, HID, CAT, LAYERS = 16, 64, 128, 4
SKIP = 2  # input column added to the output


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout() -> Dict[str, Tuple[int, Tuple[int, ...]]]:
    """Offsets and shapes of the parameters in the flat FP32 vector."""
    o, out = 0, {}

    def put(name, shape):
        nonlocal o
        out[name] = (o, shape)
        o += int(np.prod(shape))

    put("W0", (IN, HID))
    put("b0", (HID,))
    for l in range(LAYERS):
        put(f"W{l + 1}", (CAT, HID))
        put(f"b{l + 1}", (HID,))
    put("w", (HID,))
    put("b", (1,))
    out["_size"] = (o, ())
    return out


def unpack(p: np.ndarray) -> Dict[str, np.ndarray]:
    L = layout()
    return {k: np.asarray(p[o:o + int(np.prod(s))], dtype=np.float64).reshape(s)
            for k, (o, s) in L.items() if k != "_size"}


def pack(d: Dict[str, np.ndarray]) -> np.ndarray:
    L = layout()
    p = np.zeros(L["_size"][0])
    for k, (o, s) in L.items():
        if k != "_size":
            p[o:o + int(np.prod(s))] = np.asarray(d[k]).ravel()
    return p


class Graph:
    def __init__(self, off: np.ndarray, nbr: np.ndarray):
        self.off, self.nbr = np.asarray(off), np.asarray(nbr)
        self.n = self.off.size - 1
        self.deg = np.diff(self.off)
        rows = np.repeat(np.arange(self.n), self.deg)
        import scipy.sparse as sp
        w = 1.0 / np.maximum(self.deg[rows], 1)
        self.mean = sp.csr_matrix((w, (rows, self.nbr)), shape=(self.n, self.n))  # row i: mean of neighbours


def forward(p: np.ndarray, g: Graph, x: np.ndarray, rounding: bool = True):
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    cache = {"X": x}
    Z0 = r(x) @ r(P["W0"]) + P["b0"]
    H = np.maximum(Z0, 0.0)
    cache["Z0"], cache["H0"] = Z0, H
    for l in range(LAYERS):
        Xl = np.concatenate([H, g.mean @ H], axis=1)
        Z = r(Xl) @ r(P[f"W{l + 1}"]) + P[f"b{l + 1}"]
        H = H + np.maximum(Z, 0.0)
        cache[f"X{l + 1}"], cache[f"Z{l + 1}"], cache[f"H{l + 1}"] = Xl, Z, H
    y = r(H) @ r(P["w"]) + P["b"][0] + np.asarray(x, dtype=np.float64)[:, SKIP]
    return y, cache


def loss_and_grad(p: np.ndarray, g: Graph, x: np.ndarray, t: np.ndarray, rounding: bool = True):
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    y, c = forward(p, g, x, rounding)
    n = g.n
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / n
    G: Dict[str, np.ndarray] = {}
    H4 = c[f"H{LAYERS}"]
    G["w"] = r(H4).T @ dy
    G["b"] = np.array([dy.sum()])
    dH = np.outer(dy, r(P["w"]))
    for l in range(LAYERS, 0, -1):
        dZ = dH * (c[f"Z{l}"] > 0)
        G[f"W{l}"] = r(c[f"X{l}"]).T @ r(dZ)
        G[f"b{l}"] = dZ.sum(axis=0)
        dX = r(dZ) @ r(P[f"W{l}"]).T
        dH = dH + dX[:, :HID] + g.mean.T @ dX[:, HID:]
    dZ0 = dH * (c["Z0"] > 0)
    G["W0"] = r(x).T @ r(dZ0)
    G["b0"] = dZ0.sum(axis=0)
    return loss, pack(G), y
