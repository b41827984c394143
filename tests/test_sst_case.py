"""The SST case launcher (nodals_amd/sst_case.py) and the Gmsh -> HXT1
conversion (nodals_amd/hxt1.py), restating the assertions of the
reference's own launcher tests (turbulence/sst_gpu_v2/tests/
test_software.py:21-171) on this build's launcher: dry-run argv, case
overrides, validation errors, the execution path with a stand-in binary
(argv, environment, run directory, compact console), explicit sample heights,
restart iteration limits, the generic single-tetrahedron conversion and the
reporting cadences.  CPU only.

The case is tests/golden/sst_ubend_reference.case: the reference's
U-bend reference case with the values those tests assert on; its mesh is a
small generated helix in Gmsh ASCII 2.2 (the U-bend mesh is not in the
reference's tree).
"""
import json
import os
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
CASE = ROOT / "tests" / "golden" / "sst_ubend_reference.case"


@pytest.fixture(scope="module")
def casedir(tmp_path_factory):
    from nodals_amd import meshgen

    d = tmp_path_factory.mktemp("sstcase")
    shutil.copy(CASE, d / "case.case")
    meshgen.write_helix_msh2(d / "mesh.msh", core=2, ring=1, nslab=4)
    return d


def invoke(casedir, *rest, env=None):
    return subprocess.run([sys.executable, "-m", "nodals_amd.sst_case", "run", "--case", str(casedir / "case.case"),
                           *map(str, rest)], capture_output=True, text=True, env=env, cwd=ROOT)


def fake_binary(path: Path, body: str) -> Path:
    path.write_text("#!" + sys.executable + "\n" + body)
    path.chmod(0o755)
    return path


def test_baseline_dry_run(casedir):
    p = invoke(casedir, "--dry-run", "--steps", "10")
    assert p.returncode == 0, p.stderr
    for s in ("NODALS_SOFTWARE_MODE=m3-fgs scalar=q1 maxOuter=10", "--simple-coupling corrected",
              "--sst-g2-start plug", "--sst-k-fgs-sweeps 8", "--sst-omega-fgs-sweeps 8",
              "--sst-profile-average-closure 1", "--sst-profile-average-wall-law spalding_profile5",
              "--sst-profile-fit-y1 0.001632945926233649", "--sst-profile-fit-y5 0.008164729631168245",
              "--sst-shared-spalding-wall-state 1", "--sst-of-production 0"):
        assert s in p.stdout, s


def test_m4_selection_is_opt_in(casedir):
    p = invoke(casedir, "--dry-run", "--momentum-solver", "m4-hybrid")
    assert p.returncode == 0, p.stderr
    assert "NODALS_SOFTWARE_MODE=m4-hybrid" in p.stdout
    assert "M4 is experimental" in p.stderr


def test_bad_density(casedir):
    p = invoke(casedir, "--dry-run", "--set", "physics.density=1.3")
    assert p.returncode != 0
    assert "density must be 1" in p.stderr


def test_case_overrides(casedir):
    p = invoke(casedir, "--dry-run", "--set", "sst.k_sweeps=9")
    assert p.returncode == 0, p.stderr
    assert "--sst-k-fgs-sweeps 9" in p.stdout


def test_execution_path_with_a_stand_in_binary(casedir, tmp_path):
    mesh = tmp_path / "mesh.bin"
    mesh.write_bytes(b"TEST_HXT1")
    fake = fake_binary(tmp_path / "mock_native",
                       "import json,os,sys\nprint('FAKE_NATIVE_EXECUTION_OK')\n"
                       "with open('mock.json','w') as f:\n"
                       " json.dump({'argv':sys.argv[1:],'env':{k:v for k,v in os.environ.items() "
                       "if k.startswith(('SST','STAGE','NODALS'))}},f)\n")
    out = tmp_path / "run"
    env = os.environ.copy()
    env["STAGE26_INSTALL_HASH"] = "previous_stale_hash"
    env["NODALS_M3_MODE"] = "off"
    p = invoke(casedir, "--mesh", mesh, "--binary", fake, "--output", out, "--steps", "3", env=env)
    assert p.returncode == 0, p.stdout + "\n" + p.stderr
    assert "FAKE_NATIVE_EXECUTION_OK" in (out / "terminal.log").read_text()
    assert "FAKE_NATIVE_EXECUTION_OK" not in p.stdout
    rec = json.loads((out / "mock.json").read_text())
    e, argv = rec["env"], rec["argv"]
    assert e["NODALS_M3_MODE"] == "fgs"
    assert e["SST22_TRACTION_MODEL"] == "vector_projected"
    assert e["SST22D4D_MOM_MODE"] == "coupled"
    assert e["STAGE27_INITIAL_VELOCITY"] == "potential"
    assert e["NODALS_F1B_FROZEN_WALL"] == "1"
    assert e["NODALS_F2_FACE_SPALDING"] == "1"
    assert e["SST_WALL_SAMPLING_MODE"] == "ray_trace"
    assert e["SST_WALL_RAY_MAX_CROSSINGS"] == "4096"
    assert e["STAGE27_S02B_PACK"] == "1"
    assert argv[argv.index("--sst-profile-average-closure") + 1] == "1"
    assert "STAGE26_INSTALL_HASH" not in e
    assert argv[argv.index("--sst-g2-max") + 1] == "3"
    assert (out / "invocation.json").exists() and (out / "terminal.log").exists()


def test_explicit_sample_heights(casedir):
    p = invoke(casedir, "--dry-run", "--set", "wall_function.distances=1e-4, 2e-4, 3e-4, 4e-4, 5e-4")
    assert p.returncode == 0, p.stderr
    assert "--sst-profile-fit-y1 0.0001" in p.stdout
    assert "--sst-profile-fit-y5 0.00050000000000000001" in p.stdout
    p2 = invoke(casedir, "--dry-run", "--set", "wall_function.distances=0.4, 0.3, 0.2, 0.1, 0.05")
    assert p2.returncode != 0
    assert "strictly increasing" in p2.stderr


def test_restart_steps_are_absolute_native_limit(casedir, tmp_path):
    ckpt = tmp_path / "NODALS_STAGE31_STATE_it200.ckpt"
    ckpt.write_bytes(b"placeholder")
    p = invoke(casedir, "--restart", ckpt, "--steps", "10", "--dry-run")
    assert p.returncode == 0, p.stderr
    assert "--sst-g2-max 210" in p.stdout


def test_generic_mesh_orientation(tmp_path):
    """One tetrahedron: the inlet normal is -x and the outlet normal is not."""
    from nodals_amd import hxt1

    mesh = tmp_path / "single_tet.msh"
    mesh.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n"
                    "$EndNodes\n$Elements\n5\n1 2 2 1 1 1 3 4\n2 2 2 2 2 2 3 4\n3 2 2 3 3 1 2 3\n"
                    "4 2 2 3 3 1 2 4\n5 4 2 4 4 1 2 3 4\n$EndElements\n")
    d = hxt1.convert(mesh, tmp_path / "hxt")
    assert d["tetrahedra"] == 1
    assert (tmp_path / "hxt" / "HXT1_mesh.bin").is_file()
    assert d["inlet_faces"] == 1 and d["outlet_faces"] == 1
    r = hxt1.read_hxt1(tmp_path / "hxt" / "HXT1_mesh.bin")
    assert r["counts"]["ntetface"] == 4 and r["counts"]["nvel"] == 8
    # positive orientation, and the records name the tagged faces
    assert r["geom"][0, 2] > 0 and abs(r["geom"][0, 0] - 1.0 / 6.0) < 1e-15
    for rec, tri in ((r["inlet"][0], (0, 2, 3)), (r["outlet"][0], (1, 2, 3))):
        c, lf, bub = rec
        face = sorted(set(range(4)) - {int(r["tets"][c, lf])})
        assert tuple(sorted(int(r["tets"][c, j]) for j in range(4) if j != lf)) == tri == tuple(face)
        assert bub == 4 + sorted([(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]).index(tri)


def test_hxt1_round_trip_through_the_mesh_reader(casedir, tmp_path):
    """The converted helix read back by the product's HXT1 reader gives the
    same setup as the polyMesh of the same helix (cells, DOFs, patterns)."""
    from nodals_amd import hxt1, meshgen
    from nodals_amd._lib import H8

    d = hxt1.convert(casedir / "mesh.msh", tmp_path / "hxt")
    poly = meshgen.write_helix_case(tmp_path / "poly", core=2, ring=1, nslab=4)
    args = ["--duct", "1", "--wall-model", "spalding", "--re", "5000", "--max-outer", "10", "--verbose", "0"]
    a = H8(["--mesh", str(poly), *args], device=False)
    b = H8(["--mesh", d["hxt1_path"], *args], device=False)
    assert a.info == b.info
    for name, n, t in (("mesh_sizes", 4, np.int64), ("g2free", None, np.int32)):
        if n is None:
            n = int(a.get("mesh_sizes", 4, np.int64)[0] + a.get("mesh_sizes", 4, np.int64)[1])
        assert np.array_equal(a.get(name, n, t), b.get(name, n, t))
    assert np.allclose(a.get("pipe_geometry", 12), b.get("pipe_geometry", 12), rtol=1e-12, atol=1e-15)


def test_lean_reporting_case_and_cadences(casedir):
    from nodals_amd import sst_case

    p = invoke(casedir, "--dry-run")
    assert p.returncode == 0, p.stderr
    for s in ("--sst-g2-print-every 10", "--sst-g2-metrics-every 10", "--sst-fgs-check-every 10"):
        assert s in p.stdout
    assert sst_case.compact_line("NODALS_SST_RESIDUALS it=10 UxAbs=1\n")
    assert not sst_case.compact_line("NODALS_SST_G1_DIAG equation=k nonlinearIt=2\n")
    assert not sst_case.compact_line("STAGE27_AUDIT_SCHEDULE it=2\n")
    bad = invoke(casedir, "--dry-run", "--set", "output.metrics_every=7")
    assert bad.returncode != 0
    assert "multiple" in bad.stderr


def test_case_restart_path_and_native_output_environment(casedir, tmp_path):
    ckpt = tmp_path / "NODALS_STAGE31_STATE_it200.ckpt"
    ckpt.write_bytes(b"placeholder")
    mesh = tmp_path / "mesh.hxt1"
    mesh.write_bytes(b"HXT1")
    p = invoke(casedir, "--dry-run", "--mesh", mesh, "--set", f"restart.path={ckpt}",
               "--set", "restart.input_fingerprint=" + "a" * 64, "--set", "simple.max_iterations=300")
    assert p.returncode == 0, p.stderr
    assert "--sst-g2-max 300" in p.stdout
    mock = fake_binary(tmp_path / "mock_native",
                       "import os,sys,json\nprint('NODALS_SST_G1_DIAG equation=k nonlinearIt=1')\n"
                       "print('NODALS_SST_RESIDUALS it=10 UxAbs=1')\n"
                       "print('NODALS_SOFTWARE_COMPLETION=COMPLETE iterations=10 converged=0')\n"
                       "keys=['STAGE31_RESTART_FROM','STAGE31_CHECKPOINT_DIR','STAGE31_CHECKPOINT_EVERY',"
                       "'SST_G2_VTU_DIR','SST_G2_VTU_EVERY','NODALS_SST_DIAGNOSTIC_EVERY',"
                       "'NODALS_LEGACY_FINAL_DIAGNOSTICS']\n"
                       "with open('props.json','w') as f: json.dump({k:v for k,v in os.environ.items() "
                       "if k in keys},f)\n")
    out = tmp_path / "run"
    q = invoke(casedir, "--mesh", mesh, "--binary", mock, "--output", out, "--set", f"restart.path={ckpt}",
               "--set", "restart.input_fingerprint=" + "a" * 64)
    assert q.returncode == 0, q.stderr
    assert "NODALS_SST_RESIDUALS it=10" in q.stdout
    assert "NODALS_SST_G1_DIAG" not in q.stdout
    assert "NODALS_SST_G1_DIAG" in (out / "terminal.log").read_text()
    props = json.loads((out / "props.json").read_text())
    assert props["STAGE31_RESTART_FROM"] == str(ckpt)
    assert props["STAGE31_CHECKPOINT_DIR"] == str(out)
    assert props["STAGE31_CHECKPOINT_EVERY"] == "100"
    assert props["SST_G2_VTU_EVERY"] == "100"
    assert props["SST_G2_VTU_DIR"] == str(out)
    assert props["NODALS_SST_DIAGNOSTIC_EVERY"] == "0"
    assert props["NODALS_LEGACY_FINAL_DIAGNOSTICS"] == "0"


def test_unsupported_inlet_profile_is_an_error(casedir):
    p = invoke(casedir, "--dry-run", "--set", "inlet.main.profile_csv=/nonexistent.csv")
    assert p.returncode != 0
    assert "profile" in p.stderr
