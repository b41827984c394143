"""`.case` front-end: the assertions of the reference's own translation tests
(tests/test_gpu_case_translation.py:21-91, tests/test_gpu_supg_case_translation.py:10-46)
restated against nodals_amd.case, plus launcher behaviour with a stand-in
executable (the technique of turbulence/sst_gpu_v2/tests/test_software.py:53-85).
"""
import json
import os
import stat
import subprocess
import sys
from pathlib import Path

import pytest

from nodals_amd import case as m

ROOT = Path(__file__).resolve().parents[1]
CASE = ROOT / "tests" / "golden" / "gpu-h8-annotated.case"
REF_CASE = Path("/root/reference/cases/gpu-h8-annotated.case")


def pairs(opts):
    assert len(opts) % 2 == 0, opts
    return {opts[i]: opts[i + 1] for i in range(0, len(opts), 2)}


@pytest.fixture(params=["golden", "reference"])
def case_file(request):
    if request.param == "reference":
        if not REF_CASE.exists():
            pytest.skip("reference case template not present")
        return REF_CASE
    return CASE


def test_annotated_case_selects_cf_pmis_family(case_file):
    cp = m.read_case(case_file)
    o = pairs(m.build_gpu_options(cp))
    assert o["--amg-hierarchy"] == "cf"
    assert o["--cf-coarsening"] == "pmis"
    assert o["--cf-strength"] == "classical-negative"
    assert o["--cf-theta"] == "0.25"
    assert o["--cf-interp"] == "exti"
    assert o["--cf-pmax"] == "8"
    assert o["--cf-aggressive-first"] == "0"
    assert o["--amg-smoother"] == "jacobi"
    assert o["--pressure-solver"] == "richardson"
    assert o["--pressure-richardson-omega"] == "1.0"
    assert o["--fine-csr-refresh-every"] == "1"
    assert o["--momentum-work"] == "fgs1"
    assert o["--run-mode"] == "converge"
    assert m.precision(cp) == "fp32"


def test_sa_family_one_override_away(case_file):
    cp = m.read_case(case_file)
    cp.set("pressure_amg", "hierarchy", "sa")
    cp.set("pressure_amg", "smoother", "cheb2")
    cp.set("pressure", "solver", "pcg")
    o = pairs(m.build_gpu_options(cp))
    assert o["--amg-hierarchy"] == "sa"
    assert o["--amg-smoother"] == "cheb2"
    assert o["--amg-spectrum-policy"] == "auto"
    assert o["--amg-power-its"] == "16"
    assert o["--amg-cheb-degree"] == "2"
    assert o["--pressure-solver"] == "pcg"


def test_l1jacobi_and_outer_chebyshev(case_file):
    cp = m.read_case(case_file)
    cp.set("pressure_amg", "smoother", "l1jacobi")
    cp.set("pressure", "solver", "cheb")
    o = pairs(m.build_gpu_options(cp))
    assert o["--amg-smoother"] == "l1jacobi"
    assert o["--pressure-solver"] == "cheb"
    assert o["--pressure-cheb-degree"] == "3"
    assert o["--pressure-power-its"] == "6"


def test_fixed10_consistency(case_file):
    cp = m.read_case(case_file)
    cp.set("run", "run_mode", "fixed10")
    cp.set("simple", "max_iterations", "10")
    o = pairs(m.build_gpu_options(cp))
    assert o["--run-mode"] == "fixed10" and o["--max-outer"] == "10"
    cp.set("simple", "max_iterations", "11")
    with pytest.raises(ValueError, match="fixed10 requires"):
        m.build_gpu_options(cp)


def test_cheb2_without_spectrum_rejected(case_file):
    cp = m.read_case(case_file)
    cp.set("pressure_amg", "smoother", "cheb2")
    cp.set("pressure_amg", "spectrum_policy", "off")
    with pytest.raises(ValueError, match="requires spectrum_policy"):
        m.build_gpu_options(cp)


def test_unknown_key_rejected(case_file, tmp_path):
    cp = m.read_case(case_file)
    cp.set("pressure_amg", "smother", "jacobi")
    with pytest.raises(ValueError, match="unknown GPU case key"):
        m.build_gpu_options(cp)
    bad = tmp_path / "bad_key.case"
    bad.write_text("[solver]\nmesh = x\nreynolds = 20\n")
    with pytest.raises(ValueError, match=r"unknown GPU case key \[solver\] reynolds"):
        m.read_case(bad)
    missing = tmp_path / "no_mesh.case"
    missing.write_text("[solver]\nre = 20\n")
    with pytest.raises(ValueError, match=r"missing required \[solver\] mesh"):
        m.build_gpu_options(m.read_case(missing))


def test_value_types_checked(case_file):
    cp = m.read_case(case_file)
    cp.set("simple", "max_iterations", "many")
    with pytest.raises(ValueError, match="expected an integer"):
        m.build_gpu_options(cp)
    cp = m.read_case(case_file)
    cp.set("pressure_amg_cf", "aggressive_first", "maybe")
    with pytest.raises(ValueError, match="expected boolean"):
        m.build_gpu_options(cp)
    cp.set("pressure_amg_cf", "aggressive_first", "on")
    assert pairs(m.build_gpu_options(cp))["--cf-aggressive-first"] == "1"


def test_supg_translation(case_file):
    cp = m.read_case(case_file)
    for k, v in (("enabled", "true"), ("tau_scale", "0.05"), ("magic", "9.0"), ("form", "implicit"),
                 ("quad_points", "64")):
        cp.set("supg", k, v)
    o = pairs(m.build_gpu_options(cp))
    assert o["--supg"] == "1" and o["--supg-tau-scale"] == "0.05" and o["--supg-magic"] == "9.0"
    assert o["--supg-form"] == "implicit" and o["--supg-quad-points"] == "64"
    cp.set("supg", "enabled", "false")
    assert pairs(m.build_gpu_options(cp))["--supg"] == "0"
    cp.set("supg", "enabled", "true")
    cp.set("supg", "form", "explicit")
    with pytest.raises(ValueError, match="form=implicit"):
        m.build_gpu_options(cp)
    cp.set("supg", "form", "implicit")
    cp.set("supg", "quad_points", "27")
    with pytest.raises(ValueError, match="quad_points=64"):
        m.build_gpu_options(cp)


def test_set_override_and_fgmres_extension():
    cp = m.read_case(CASE)
    m.apply_overrides(cp, ["pressure.solver=fgmres", "pressure.restart=20", "pressure_amg.hierarchy=sa"])
    o = pairs(m.build_gpu_options(cp))
    assert o["--pressure-solver"] == "fgmres" and o["--pressure-restart"] == "20"
    with pytest.raises(ValueError, match="section.key=value"):
        m.apply_overrides(cp, ["novalue"])
    with pytest.raises(ValueError, match="unknown GPU case key"):
        m.apply_overrides(cp, ["pressure.nope=1"])


def test_translated_argv_accepted_by_solver_parser(tmp_path):
    """Every option the translator emits parses in the native solver (host
    setup only, small generated mesh)."""
    from nodals_amd import meshgen
    from nodals_amd._lib import H8

    mesh = meshgen.write_pipe_case(tmp_path / "pipe", core=4, ring=2, nslab=4)
    for hierarchy in ("cf", "sa"):  # cf (PMIS, ext+i, pmax 8) is the translator default
        cp = m.read_case(CASE)
        m.apply_overrides(cp, [f"solver.mesh={mesh}", f"pressure_amg.hierarchy={hierarchy}"])
        h = H8(m.build_gpu_options(cp), device=False)
        assert h.info["cells"] == (16 + 32) * 4 * 6
        h.close()


def _fake_solver(tmp_path: Path, body: str) -> Path:
    exe = tmp_path / "fake_h8"
    exe.write_text("#!" + sys.executable + "\nimport json, sys, os\n" + body)
    exe.chmod(exe.stat().st_mode | stat.S_IEXEC)
    return exe


def _run_launcher(tmp_path, exe, extra=()):
    out = tmp_path / "out"
    cmd = [str(ROOT / "bin" / "nodals-gpu"), str(CASE), "--exe", str(exe), "--set", f"run.output_root={out}",
           "--set", "solver.mesh=/nonexistent", *extra]
    return subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120), out


def test_launcher_pass_detection_and_argv(tmp_path):
    exe = _fake_solver(tmp_path, "json.dump(sys.argv[1:], open(os.environ['FAKE_ARGV'], 'w'))\n"
                                 "print('NODALS_GPU_RESULT gate=H8 tag=x status=PASS')\n")
    os.environ["FAKE_ARGV"] = str(tmp_path / "argv.json")
    try:
        r, out = _run_launcher(tmp_path, exe, ["--", "--verbose", "0"])
    finally:
        del os.environ["FAKE_ARGV"]
    assert r.returncode == 0, r.stdout
    assert "NODALS_GPU_CASE_RESULT status=PASS" in r.stdout
    argv = json.loads((tmp_path / "argv.json").read_text())
    assert argv[-2:] == ["--verbose", "0"] and pairs(argv[:-2])["--mesh"] == "/nonexistent"
    assert (out / "gpu_h8_cf_pmis.log").exists()


def test_launcher_not_converged_and_failure_codes(tmp_path):
    exe = _fake_solver(tmp_path, "print('NODALS_GPU_RESULT gate=H8 tag=x status=FAIL')\n")
    r, _ = _run_launcher(tmp_path, exe)
    assert r.returncode == 21 and "NOT_CONVERGED" in r.stdout
    exe = _fake_solver(tmp_path, "sys.exit(35)\n")
    r, _ = _run_launcher(tmp_path, exe)
    assert r.returncode == 35
    # a missing --exe falls through to the in-tree solver (reference lookup
    # order); the real binary fails on the missing mesh before touching the
    # GPU and reports the exception exit code
    r, _ = _run_launcher(tmp_path, tmp_path / "missing_exe")
    assert r.returncode == 90 and "NODALS_GPU_H8_EXCEPTION" in r.stdout
    bad = tmp_path / "bad.case"
    bad.write_text("[nope]\nx = 1\n")
    rr = subprocess.run([str(ROOT / "bin" / "nodals-gpu"), str(bad), "--dry-run"], stdout=subprocess.PIPE,
                        stderr=subprocess.STDOUT, text=True, timeout=60)
    assert rr.returncode == 4 and "unknown GPU case section" in rr.stdout


def test_dry_run_prints_command():
    r = subprocess.run([str(ROOT / "bin" / "nodals-gpu"), str(CASE), "--dry-run", "--dump-options-json"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0
    assert "NODALS_GPU_CASE_RESULT status=DRY_RUN" in r.stdout
    line = [l for l in r.stdout.splitlines() if l.startswith("NODALS_GPU_OPTIONS_JSON=")][0]
    opts = json.loads(line.split("=", 1)[1])
    assert pairs(opts)["--amg-hierarchy"] == "cf"


def test_launcher_time_limit_while_printing(tmp_path):
    """The time limit applies while the solver is still printing: the child
    group is stopped and the launcher reports 124."""
    exe = _fake_solver(tmp_path, "import time\nprint('started', flush=True)\ntime.sleep(60)\n")
    r, out = _run_launcher(tmp_path, exe, ["--set", "run.timeout_seconds=1"])
    assert r.returncode == 124, r.stdout
    assert "started" in (out / "gpu_h8_cf_pmis.log").read_text()
