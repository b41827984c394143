"""Seeded input sets for the SST-2003 local state and the wall-law kernels
(shared by the CPU and GPU tests)."""
import numpy as np


def sst_inputs(n=4000, seed=20261202):
    rng = np.random.default_rng(seed)
    R = 0.5
    r = R * np.sqrt(rng.uniform(0.0, 1.0, n))
    r[:8] = [0.0, R, R * (1 - 1e-12), R * 0.999, 0.25, 0.4999, 0.1, 0.49]
    th = rng.uniform(0, 2 * np.pi, n)
    xy = np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
    gu = rng.standard_normal((n, 9)) * 10.0 ** rng.uniform(-3, 3, (n, 1))
    gu[:20] = 0.0  # zero strain
    k = 10.0 ** rng.uniform(-10, 1, n)
    om = 10.0 ** rng.uniform(-5, 5, n)
    k[20:30] = 0.0  # below the floor
    om[30:40] = -1.0
    gk = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-3, 2, (n, 1))
    gw = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-2, 4, (n, 1))
    dist = R - np.hypot(xy[:, 0], xy[:, 1])
    return dict(gu=gu, k=k, omega=om, grad_k=gk, grad_omega=gw, nu=1e-5, radius=R, xy=xy, dist=dist)


def spalding_inputs(n=4000, seed=7):
    rng = np.random.default_rng(seed)
    slip = np.concatenate([[0.0, -1e-20, 1e-14, -3.0, 1e30], 10.0 ** rng.uniform(-8, 3, n - 5)])
    slip *= np.where(rng.uniform(size=n) < 0.3, -1.0, 1.0)
    y = 10.0 ** rng.uniform(-7, 0, n)
    return slip, y, 1.5e-5


def profile_inputs(n=2000, seed=11):
    rng = np.random.default_rng(seed)
    base = 10.0 ** rng.uniform(-6, -2, (n, 1))
    y = base * np.array([0.6, 0.8, 1.0, 1.2, 1.4])
    U = 10.0 ** rng.uniform(-4, 2, (n, 1)) * (np.log1p(y / base * 20.0) + rng.uniform(0, 0.1, (n, 5)))
    U[:5] = 0.0
    Us = U * np.where(rng.uniform(size=(n, 1)) < 0.5, -1.0, 1.0)
    y[5, 3] = y[5, 2]  # non-increasing heights: invalid
    return U, Us, y, 1.5e-5


def omega_inputs(n=3000, seed=5):
    rng = np.random.default_rng(seed)
    a = np.empty((n, 15))
    a[:, 0] = 10.0 ** rng.uniform(-6, -1, n)     # y
    a[:, 1] = 10.0 ** rng.uniform(-6, -4, n)     # nu
    a[:, 2] = 10.0 ** rng.uniform(-3, 2, n)      # Ut
    a[:, 3] = 10.0 ** rng.uniform(-4, 1, n)      # uTau
    a[:, 4] = 10.0 ** rng.uniform(-10, 1, n)     # k
    a[:, 5] = 10.0 ** rng.uniform(-3, 6, n)      # omega
    a[:, 6] = 1e-8
    a[:, 7] = 1e-3
    a[:, 8] = 0.41
    a[:, 9] = 9.8
    a[:, 10] = 0.075
    a[:, 11] = 10.0 ** rng.uniform(0, 2, n)      # reBlend
    a[:, 12] = 10.0
    a[:, 13] = 10.0 ** rng.uniform(-6, 3, n)     # pk
    a[:, 14] = 10.0 ** rng.uniform(-7, -2, n)    # nut
    shared = (rng.uniform(size=n) < 0.5).astype(np.int32)
    return a, shared
