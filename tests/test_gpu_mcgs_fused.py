"""Coarse MCGS levels smoothed in one workgroup (k_mcgs_level_wg, the
iterate in LDS, a barrier between colours) give bit-identical results to
one launch per colour: fixed-10 SIMPLE with an MCGS-smoothed SA hierarchy
(pipe, FGMRES) and with the CF hierarchy of the wall-function path (helix,
PCG), with the fused levels on (the default) and off
(--amg-mcgs-fuse-nnz 0)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(h, nv, nc):
    return [h.get(f"u{d}", nv) for d in range(3)] + [h.get("p", nc)]


def test_fused_mcgs_levels_are_bit_identical_sa(mesh_dir):
    from nodals_amd._lib import H8

    print("\n[gpu] fused vs per-colour MCGS levels, SA, medium pipe", flush=True)
    base = ["--mesh", mesh_dir["medium"], "--max-outer", "10", "--amg-smoother", "mcgs", "--pressure-solver", "fgmres",
            "--p-max", "200", "--verbose", "0"]
    out = []
    for fuse in ("250000", "0"):
        h = H8(base + ["--amg-mcgs-fuse-nnz", fuse])
        r = h.run()
        assert r["pass_"] == 1
        out.append((_fields(h, h.info["nv"], h.info["cells"]), r["avg_pressure_its"]))
        h.close()
    assert out[0][1] == out[1][1]
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))


def test_fused_mcgs_levels_are_bit_identical_wall_path(mesh_dir):
    from nodals_amd._lib import H8

    print("[gpu] fused vs per-colour MCGS levels, CF + PCG wall path, small helix", flush=True)
    wall = ["--duct", "1", "--wall-model", "spalding", "--re", "5000", "--alpha-u", "0.7", "--alpha-p", "0.3",
            "--amg-hierarchy", "cf", "--cf-interp", "exti", "--amg-smoother", "mcgs", "--pressure-solver", "pcg",
            "--p-rtol", "1e-3", "--p-max", "600", "--max-outer", "10", "--verbose", "0"]
    out = []
    for fuse in ("250000", "0"):
        h = H8(["--mesh", mesh_dir["helix"], *wall, "--amg-mcgs-fuse-nnz", fuse])
        h.turb_init([])
        h.wall_init([])
        r = h.run()
        out.append((_fields(h, h.info["nv"], h.info["cells"]), r["avg_pressure_its"]))
        h.close()
    assert out[0][1] == out[1][1]
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
