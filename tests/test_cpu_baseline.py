"""The C++/OpenMP CPU baseline (oracle/cpu_h8.cpp) against the numpy
oracle's SIMPLE loop: same fields to rounding and the same pressure
iteration counts, on two pipes, at one thread and at several.  The numpy
oracle is itself pinned against the reference's host code and golden
runs in test_h8_oracle.py; this file only checks that the timed baseline
computes the same iterations."""
import numpy as np
import pytest

from oracle import cpubind
from oracle import h8_oracle as O

pytestmark = pytest.mark.skipif(not cpubind.available(), reason="oracle/_cpu not built (make -C oracle cpu)")


def _case(mesh):
    from nodals_amd._lib import H8

    h = H8(["--mesh", mesh, "--max-outer", "10", "--verbose", "0"], device=False)
    case = O.case_from_product(h)
    h.close()
    return case


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("which", ["small", "medium"])
def test_cpu_port_matches_numpy_oracle(mesh_dir, which):
    case = _case(mesh_dir[which])
    ref = O.simple(case, p_solver="fgmres", p_rtol=0.5, p_max=200, smoother=O.Smoother(kind="jacobi"), max_outer=10)
    cs = cpubind.CpuSimple(case)
    r = cs.simple(10)
    assert [h["p_its"] for h in ref["hist"]] == r["p_its"].tolist()
    assert max(_rel(r["U"][d], ref["U"][d]) for d in range(3)) < 1e-10
    assert _rel(r["p"], ref["p"]) < 1e-10
    # one thread computes the same iterations
    r1 = cs.simple(10, threads=1)
    assert r1["p_its"].tolist() == r["p_its"].tolist()
    assert max(_rel(r1["U"][d], r["U"][d]) for d in range(3)) < 1e-10
    cs.close()


def test_cpu_port_restarts_from_state(mesh_dir):
    """Two runs of 4 and 6 iterations, the second from the first's state,
    equal one run of 10 (the bench times the port in such chunks)."""
    case = _case(mesh_dir["small"])
    cs = cpubind.CpuSimple(case)
    whole = cs.simple(10, threads=1)
    a = cs.simple(4, threads=1)
    b = cs.simple(6, threads=1, U0=a["U"], p0=a["p"])
    assert np.array_equal(b["U"], whole["U"]) and np.array_equal(b["p"], whole["p"])
    cs.close()


def test_cpu_port_planted_relaxation_mistake_is_seen(mesh_dir):
    """A planted change of the momentum under-relaxation in the oracle run
    (0.5 -> 0.55) moves the fields far outside the tolerance used above."""
    case = _case(mesh_dir["small"])
    bad = O.simple(case, alpha_u=0.55, p_solver="fgmres", p_rtol=0.5, p_max=200, smoother=O.Smoother(kind="jacobi"),
                   max_outer=10)
    r = cpubind.CpuSimple(case).simple(10)
    assert max(_rel(r["U"][d], bad["U"][d]) for d in range(3)) > 1e-4
