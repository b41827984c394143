"""Pins of the learned-warm-start oracle: its analytic gradient equals central
finite differences of its loss (rounding off, float64), and its BF16
rounding matches the definition on known values."""
import numpy as np

from oracle import gnn_oracle as G


def _ring_graph(n):
    off = np.arange(0, 2 * n + 1, 2)
    nbr = np.array([[(i - 1) % n, (i + 1) % n] for i in range(n)]).ravel()
    return G.Graph(off, nbr)


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159265, 0.0], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    # 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    assert list(G.bf16(v)) == [1.0, 1.0, 1.0078125, 1.0078125, -3.140625, 0.0]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    g = _ring_graph(9)
    p = rng.standard_normal(G.layout()["_size"][0]) * 0.3
    x = rng.standard_normal((g.n, G.IN))
    t = rng.standard_normal(g.n)
    loss, grad, _ = G.loss_and_grad(p, g, x, t, rounding=False)
    idx = rng.choice(p.size, 60, replace=False)
    # include every block of the layout
    L = G.layout()
    idx = np.unique(np.concatenate([idx, [o for k, (o, s) in L.items() if k != "_size"]]))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = G.loss_and_grad(q, g, x, t, rounding=False)[0]
        q[i] -= 2 * h
        lm = G.loss_and_grad(q, g, x, t, rounding=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])
