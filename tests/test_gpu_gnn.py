"""Learned pressure-correction warm start on the GPU.

Forward pass and parameter gradients of the BF16-MFMA network against the
CPU oracle (same BF16 rounding points; the device accumulates in FP32 in its
own order, so a rare rounding flip of an intermediate activation is allowed:
1e-3 relative in norm); training on samples recorded from the solver's own
SIMPLE runs lowers the loss; the trained network, as the initial guess of
the FGMRES pressure solve at a Reynolds number not trained on, lowers the
pressure iterations while the converged fields stay those of the plain
method (1e-6 relative L2).
"""
import numpy as np
import pytest

from oracle import gnn_oracle as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, dtype=np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small(mesh_dir):
    from nodals_amd._lib import H8

    args = ["--mesh", mesh_dir["small"], "--run-mode", "converge", "--max-outer", "3000", "--simple-tol", "1e-6",
            "--amg-smoother", "jacobi", "--pressure-solver", "fgmres", "--p-rtol", "0.1", "--p-max", "200",
            "--verbose", "0"]
    return args


def test_forward_and_gradient_match_oracle(small):
    from nodals_amd._lib import Gnn, H8

    print("\n[gpu] GNN forward/gradient vs oracle, small pipe", flush=True)
    h = H8(small + ["--re", "20"])
    g = Gnn(h, seed=3)
    gr = h.gnn_graph()
    graph = G.Graph(gr[0], gr[1])
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((g.n, G.IN)) * 0.7).astype(np.float32)
    t = rng.standard_normal(g.n).astype(np.float32)
    p = g.params().astype(np.float64)
    # make every block active: perturb the zero-initialised parts
    p = (p + rng.standard_normal(p.size) * 0.05).astype(np.float32)
    g.set_params(p)
    y = g.forward(x)
    yo, _ = G.forward(p.astype(np.float64), graph, x.astype(np.float64))
    assert rel(y, yo) < 1e-3
    grad, loss = g.gradient(x, t)
    lo, go, _ = G.loss_and_grad(p.astype(np.float64), graph, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lo) <= 1e-3 * lo
    L = G.layout()
    for k, (o, s) in L.items():
        if k == "_size":
            continue
        sl = slice(o, o + int(np.prod(s)))
        assert rel(grad[sl], go[sl]) < 2e-3, k
    g.close()
    h.close()


def test_training_and_warm_start_gain(small):
    from nodals_amd._lib import Gnn, H8

    train_re = [10.0, 25.0, 45.0, 80.0, 120.0]
    h0 = H8(small + ["--re", "60"])
    g = Gnn(h0, seed=1)
    g.reserve(1500)
    for re in train_re:
        print(f"\n[gpu] record SIMPLE samples small pipe Re={re}", flush=True)
        h = H8(small + ["--re", str(re)])
        h.attach_gnn(g, warm_start=False, record_every=2)
        r = h.run()
        h.attach_gnn(None)
        assert r["converged"] == 1 and r["recorded"] > 0
        h.close()
    assert g.samples > 100
    l0 = g.evaluate()
    print(f"[gpu] train on {g.samples} samples (initial loss {l0:.4e})", flush=True)
    last = g.train(epochs=30, lr=1e-3, seed=2)
    assert last < 0.5 * l0, (l0, last)
    # held-out Reynolds number: plain FGMRES from zero vs learned warm start
    print("[gpu] held-out Re=60: plain vs learned warm start", flush=True)
    tight = ["--simple-tol", "1e-9"]
    hp = H8(small + ["--re", "60"] + tight)
    rp = hp.run()
    fp = [hp.get(f"u{d}", hp.info["nv"]) for d in range(3)] + [hp.get("p", hp.info["cells"])]
    hw = H8(small + ["--re", "60"] + tight)
    hw.attach_gnn(g, warm_start=True)
    rw = hw.run()
    hw.attach_gnn(None)
    fw = [hw.get(f"u{d}", hw.info["nv"]) for d in range(3)] + [hw.get("p", hw.info["cells"])]
    print(f"[gpu] pressure its plain={rp['pressure_its']} warm={rw['pressure_its']} "
          f"outer {rp['outer']}/{rw['outer']} fallbacks={rw['fallbacks']}/{rw['warm_starts']}", flush=True)
    assert rp["converged"] == 1 and rw["converged"] == 1
    assert rw["warm_starts"] == rw["outer"] - 1  # from the second iteration on
    assert rw["pressure_its"] < rp["pressure_its"]
    for a, b in zip(fw, fp):
        assert rel(a, b) < 1e-6
    for x in (hp, hw):
        x.close()
    g.close()
    h0.close()
