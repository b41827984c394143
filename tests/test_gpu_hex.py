"""Q1+BF2 hexahedral element blocks on the GPU (nodals_hex_element_blocks)
against the oracle (oracle/hex_oracle.py, pinned in tests/test_hex_oracle.py)
on curved hexahedra with random velocities: diffusion and convection blocks
to 1e-12 relative per element; an inverted element is counted, not used."""
import numpy as np
import pytest

from oracle import hex_oracle as X

pytestmark = pytest.mark.gpu


def test_hex_blocks_match_oracle():
    from nodals_amd import _lib

    n = 203                                   # not a multiple of the 4 elements per block
    print(f"\n[gpu] hex element blocks, n={n}", flush=True)
    coords = X.random_hexes(n, seed=11)
    rng = np.random.default_rng(12)
    u = rng.standard_normal((n, 14, 3))
    D, C, bad = _lib.hex_element_blocks(coords, u, 0.013)
    assert bad == 0
    Do, Co = X.batch(coords, u, 0.013)
    ed = max(np.abs(D[c] - Do[c]).max() / np.abs(Do[c]).max() for c in range(n))
    ec = max(np.abs(C[c] - Co[c]).max() / np.abs(Co[c]).max() for c in range(n))
    print(f"[gpu]   diffusion {ed:.2e} convection {ec:.2e}", flush=True)
    assert ed < 1e-12 and ec < 1e-12


def test_hex_blocks_count_inverted_elements():
    from nodals_amd import _lib

    coords = X.random_hexes(3, seed=13)
    coords[1] = coords[1][[4, 5, 6, 7, 0, 1, 2, 3]]   # top and bottom swapped: negative Jacobian
    _, _, bad = _lib.hex_element_blocks(coords, np.zeros((3, 14, 3)), 1.0)
    assert bad == 27
