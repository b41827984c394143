"""Adaptive y+ segment runner (nodals_amd/adaptive_case.py) on the CPU: the
refresh schedule, the plan and dry-run commands, the prerequisites, and the
whole stage loop with a stand-in native solver (the technique of
turbulence/sst_gpu_v2/tests/test_software.py:53-85), including --resume and
the checks that stop a stage whose outputs or markers are missing.
Behaviour restated from examples/spiral_swept_207k/run_adaptive_case.py."""
import json
import os
import stat
import subprocess
import sys
from pathlib import Path

import pytest

from nodals_amd import adaptive_case as A

ROOT = Path(__file__).resolve().parents[1]
CASE = ROOT / "tests" / "golden" / "sst_ubend_reference.case"


def test_refresh_schedule():
    assert A.refresh_schedule(1000, True, 100, 100, -1)[0] == list(range(100, 1000, 100))
    r, s = A.refresh_schedule(650, True, 200, 250, -1)
    assert r == [200, 450] and s == [(0, 200), (200, 450), (450, 650)]
    assert A.refresh_schedule(650, True, 200, 250, 1)[0] == [200]
    assert A.refresh_schedule(650, False, 200, 250, -1) == ([], [(0, 650)])
    assert A.refresh_schedule(650, True, 200, 250, 0) == ([], [(0, 650)])
    assert A.refresh_schedule(200, True, 200, 10, -1) == ([], [(0, 200)])  # a refresh at the end is no refresh
    for bad in ((0, True, 1, 1, -1), (10, True, 0, 1, -1), (10, True, 1, 0, -1), (10, True, 1, 1, -2)):
        with pytest.raises(ValueError):
            A.refresh_schedule(*bad)


@pytest.fixture()
def case_dir(tmp_path):
    from nodals_amd import meshgen

    d = tmp_path / "c"
    d.mkdir()
    (d / "case.case").write_text(CASE.read_text())
    meshgen.write_helix_msh2(d / "mesh.msh", core=4, ring=2, nslab=12)
    sets = ["mesh.path=mesh.msh", "simple.max_iterations=30", "adaptive_sampling.first_refresh=10",
            "adaptive_sampling.refresh_every=10", "output.checkpoint_every=7", "output.vtu_every=0"]
    return d, sets


def _runner(d, sets, *extra, binary=None):
    cmd = [sys.executable, "-m", "nodals_amd.adaptive_case", "--case", str(d / "case.case"), "--output",
           str(d / "out"), "--cache", str(d / "cache")]
    if binary:
        cmd += ["--binary", str(binary)]
    for s in sets:
        cmd += ["--set", s]
    env = dict(os.environ, PYTHONPATH=str(ROOT))
    return subprocess.run(cmd + list(extra), capture_output=True, text=True, cwd=d, timeout=300, env=env)


def test_dry_run_prints_plan_and_segment_commands(case_dir):
    d, sets = case_dir
    r = _runner(d, sets, "--dry-run")
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "NODALS_ADAPTIVE_PLAN final=30 refreshAt=[10, 20]" in out
    for i, (b, e) in enumerate([(0, 10), (10, 20), (20, 30)]):
        assert f"NODALS_ADAPTIVE_STAGE index={i} from={b} to={e} " in out
    cmds = [l for l in out.splitlines() if l.startswith("NODALS_ADAPTIVE_DRY_CMD")]
    assert len(cmds) == 3
    # checkpoint cadence 7 does not divide 10, so each segment saves at its end
    assert "output.checkpoint_every=10" in cmds[0] and "output.checkpoint_every=30" in cmds[2]
    assert "environment.NODALS_FACE_YPLUS_UTAU_CSV=" in cmds[0] and "--restart" not in cmds[0]
    assert "stage_0000_0010/NODALS_STAGE31_STATE_it10.ckpt" in cmds[1]
    assert "stage_0010_0020/NODALS_FACE_UTAU_it20.csv" in cmds[2]
    assert "--steps 10" in cmds[2] and "environment.NODALS_FACE_YPLUS_EXPORT=1" in cmds[1]
    assert "NODALS_ADAPTIVE_DRY_RUN iteration=30" in out
    assert not (d / "out").exists()


def test_prerequisites_and_errors(case_dir):
    d, sets = case_dir
    r = _runner(d, sets + ["wall_function.samples=5"], "--dry-run")
    assert r.returncode == 2 and "NODALS_ADAPTIVE_ERROR: Current face-y+ patch requires" in r.stderr
    r = _runner(d, sets + ["adaptive_sampling.min_scale=0.5", "adaptive_sampling.max_scale=0.4"], "--dry-run")
    assert r.returncode == 2 and "Invalid target y+" in r.stderr
    r = _runner(d, sets + ["mesh.path=missing.msh"], "--dry-run")
    assert r.returncode == 2 and "Mesh missing" in r.stderr
    # without refreshes the prerequisites do not apply
    r = _runner(d, sets + ["wall_function.samples=5", "adaptive_sampling.enabled=false"], "--dry-run")
    assert r.returncode == 0 and "refreshAt=[]" in r.stdout


FAKE = r'''
import os, sys
argv = sys.argv[1:]
last = int(argv[argv.index("--max-outer") + 1])
ck = os.environ["STAGE31_CHECKPOINT_DIR"]
log = open(os.environ["FAKE_LOG"], "a")
log.write(f"{os.environ.get('STAGE31_RESTART_FROM', '')}|{os.environ.get('NODALS_FACE_YPLUS_UTAU_CSV', '')}|{last}\n")
if os.environ.get("STAGE31_RESTART_FROM"):
    print("STAGE31_RESTART_LOAD=PASS path=%s iteration=0" % os.environ["STAGE31_RESTART_FROM"])
if os.environ.get("NODALS_FACE_YPLUS_UTAU_CSV"):
    assert os.path.isfile(os.environ["NODALS_FACE_YPLUS_UTAU_CSV"])
    print("NODALS_FACE_YPLUS_LOADED faces=3 status=PASS")
skip = os.environ.get("FAKE_SKIP_EXPORT_AT")
open(os.path.join(ck, "NODALS_STAGE31_STATE_it%d.ckpt" % last), "wb").write(b"NODALSCK")
if os.environ.get("NODALS_FACE_YPLUS_EXPORT") == "1" and str(last) != skip:
    open(os.path.join(ck, "NODALS_FACE_UTAU_it%d.csv" % last), "w").write("0,0,0,0.05,30\n")
print("NODALS_SOFTWARE_COMPLETION=COMPLETE iterations=%d converged=0" % last)
'''


def _fake(d):
    exe = d / "fake_sst"
    exe.write_text("#!" + sys.executable + "\n" + FAKE)
    exe.chmod(exe.stat().st_mode | stat.S_IEXEC)
    return exe


def test_stage_loop_with_a_stand_in_solver(case_dir, monkeypatch):
    d, sets = case_dir
    exe = _fake(d)
    log = d / "fake.log"
    monkeypatch.setenv("FAKE_LOG", str(log))
    r = _runner(d, sets, binary=exe)
    assert r.returncode == 0, r.stdout + r.stderr
    runs = [l.split("|") for l in log.read_text().splitlines()]
    assert [int(x[2]) for x in runs] == [10, 20, 30]  # absolute limits from the restart names
    assert runs[0][0] == "" and runs[0][1] == ""
    assert runs[1][0].endswith("stage_0000_0010/NODALS_STAGE31_STATE_it10.ckpt")
    assert runs[2][1].endswith("stage_0010_0020/NODALS_FACE_UTAU_it20.csv")
    m = json.loads((d / "out" / "adaptive_schedule.json").read_text())
    assert m["completed"] == [10, 20, 30] and m["refreshes"] == [10, 20]
    assert r.stdout.count("NODALS_ADAPTIVE_PASS") == 3 and "NODALS_ADAPTIVE_COMPLETE iteration=30" in r.stdout
    # --resume skips every complete stage
    log.write_text("")
    r = _runner(d, sets, "--resume", binary=exe)
    assert r.returncode == 0 and r.stdout.count("NODALS_ADAPTIVE_SKIP_COMPLETE") == 3 and log.read_text() == ""
    # a missing face export stops the loop at that stage
    monkeypatch.setenv("FAKE_SKIP_EXPORT_AT", "20")
    r = _runner(d, sets, binary=exe)
    assert r.returncode == 2 and "did not export F2 wall u_tau" in r.stderr
    assert r.stdout.count("NODALS_ADAPTIVE_PASS") == 1
