"""Adaptive y+ segment runner for SST cases.

``python -m nodals_amd.adaptive_case --case X.case [--resume] [--dry-run]``
runs an SST ``.case`` in segments of SIMPLE iterations.  Each segment ends
with a checkpoint and an export of the wall faces' fitted friction
velocities (``NODALS_FACE_UTAU_it<N>.csv``).  The next segment restarts from
that checkpoint and moves every wall face's Spalding sample to the height
that puts it at the target y+ (``NODALS_FACE_YPLUS_*`` environment of the
native solver).

Restates the behaviour of the reference's runner
(examples/spiral_swept_207k/run_adaptive_case.py:56-204): the
``[adaptive_sampling]`` keys and defaults (enabled true, first_refresh 100,
refresh_every 100, max_refreshes -1, targets 30, min_scale 0.02, max_scale
0.35), the refresh schedule (first, first + every, ... below the final
iteration, at most max_refreshes), the one-point/ray-trace prerequisites,
the stage directories ``stage_BBBB_EEEE``, ``adaptive_schedule.json``, the
``--resume`` skip rule, the per-stage checks (checkpoint, face export,
restart and face-load markers) and the ``NODALS_ADAPTIVE_*`` marker lines;
errors print ``NODALS_ADAPTIVE_ERROR:`` and exit 2.  Each segment runs
through this package's SST launcher (``nodals_amd.sst_case run``) instead of
the reference's ``nodals.py``, so there is no ``--root``.
"""
from __future__ import annotations

import argparse
import configparser
import json
import os
import re
import shlex
import shutil
import subprocess
import sys
from dataclasses import dataclass, field
from datetime import datetime
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

from .sst_case import Case

STAGE_DIR = re.compile(r"stage_\d{4,}_\d{4,}")
# exit codes of a segment that ran its iterations (3: the reference solver's
# "steps ended without formal convergence")
SEGMENT_OK = (0, 3)


class AdaptiveError(ValueError):
    pass


def _flag(text: str) -> bool:
    v = str(text).strip().lower()
    if v in ("yes", "true", "on", "1"):
        return True
    if v in ("no", "false", "off", "0"):
        return False
    raise AdaptiveError(f"Expected true/false; got {text!r}")


def _path(text, base: Path) -> Path:
    p = Path(os.path.expandvars(os.path.expanduser(str(text))))
    return (p if p.is_absolute() else base / p).resolve()


def refresh_schedule(final: int, enabled: bool, first: int, every: int, limit: int) -> Tuple[List[int], List[Tuple[int, int]]]:
    """Refresh iterations and the (begin, end) segments they cut [0, final]
    into."""
    if final < 1 or first < 1 or every < 1 or limit < -1:
        raise AdaptiveError("final iterations, first_refresh and refresh_every must be positive; max_refreshes >= -1")
    at: List[int] = []
    if enabled and limit != 0:
        it = first
        while it < final and (limit < 0 or len(at) < limit):
            at.append(it)
            it += every
    cuts = [0] + at + [final]
    return at, list(zip(cuts[:-1], cuts[1:]))


@dataclass
class Plan:
    case: Path
    mesh: Path
    mesh_override: Optional[Path]
    output: Path
    final: int
    refreshes: List[int]
    stages: List[Tuple[int, int]]
    target_sp: float
    target_omega: float
    min_scale: float
    max_scale: float
    checkpoint_every: int
    vtu_every: int
    overrides: List[str] = field(default_factory=list)
    binary: Optional[Path] = None  # native solver of each segment (the launcher's default when None)
    cache: Optional[Path] = None   # HXT1 mesh cache of the launcher (its default when None)

    @staticmethod
    def stage_dir(out: Path, begin: int, end: int) -> Path:
        return out / f"stage_{begin:04d}_{end:04d}"


def make_plan(case: Path, overrides: Sequence[str], final: Optional[int], mesh: Optional[Path],
              output: Optional[Path]) -> Plan:
    cp = Case(case, list(overrides)).cp

    def get(section: str, key: str, default) -> str:
        return cp.get(section, key, fallback=str(default))

    last = final if final is not None else int(get("simple", "max_iterations", 1000))
    on = _flag(get("adaptive_sampling", "enabled", "true"))
    first = int(get("adaptive_sampling", "first_refresh", 100))
    every = int(get("adaptive_sampling", "refresh_every", 100))
    limit = int(get("adaptive_sampling", "max_refreshes", -1))
    sp = float(get("adaptive_sampling", "target_spalding_yplus", 30))
    om = float(get("adaptive_sampling", "target_omega_yplus", 30))
    lo = float(get("adaptive_sampling", "min_scale", 0.02))
    hi = float(get("adaptive_sampling", "max_scale", 0.35))
    if not (0 < sp <= 500 and 0 < om <= 500 and 0 < lo <= hi <= 1):
        raise AdaptiveError("Invalid target y+ or sampling-scale bounds")
    refreshes, stages = refresh_schedule(last, on, first, every, limit)
    out = _path(output, Path.cwd()) if output else _path(get("output", "directory", "runs/ubend_adaptive"), case.parent)
    source = _path(mesh, Path.cwd()) if mesh else _path(get("mesh", "path", ""), case.parent)
    if not source.is_file():
        raise FileNotFoundError("Mesh missing: " + str(source))
    if refreshes:
        one_point = int(get("wall_function", "samples", 1)) == 1 and int(get("omega_wall", "sample_index", 3)) == 3
        rays = get("wall_sampling", "mode", "ray_trace").lower() == "ray_trace"
        if not (one_point and rays and _flag(get("wall_function", "face_spalding", "true"))):
            raise AdaptiveError("Current face-y+ patch requires wall_function.samples=1, omega_wall.sample_index=3, "
                                "wall_sampling.mode=ray_trace, wall_function.face_spalding=true")
    if out in {Path("/"), Path.home().resolve(), Path(__file__).resolve().parents[1]}:
        raise AdaptiveError("Output directory must be a dedicated run directory, not home/root/source")
    return Plan(case=case, mesh=source, mesh_override=_path(mesh, Path.cwd()) if mesh else None, output=out,
                final=last, refreshes=refreshes, stages=stages, target_sp=sp, target_omega=om, min_scale=lo,
                max_scale=hi, checkpoint_every=int(get("output", "checkpoint_every", 100)),
                vtu_every=int(get("output", "vtu_every", 100)), overrides=list(overrides))


def _cadence(requested: int, end: int) -> int:
    """A cadence that saves at `end`: the requested one if it divides end."""
    return requested if requested > 0 and end % requested == 0 else end


def segment_command(plan: Plan, index: int, prev: Optional[Path]) -> List[str]:
    begin, end = plan.stages[index]
    cmd = [sys.executable, "-u", "-m", "nodals_amd.sst_case", "run", "--case", str(plan.case)]
    if plan.binary:
        cmd += ["--binary", str(plan.binary)]
    if plan.cache:
        cmd += ["--cache", str(plan.cache)]
    if plan.mesh_override:
        cmd += ["--mesh", str(plan.mesh_override)]
    for o in plan.overrides:
        cmd += ["--set", o]
    export = "1" if plan.refreshes else "0"
    sets = [f"output.checkpoint_every={_cadence(plan.checkpoint_every, end)}",
            f"output.vtu_every={_cadence(plan.vtu_every, end)}",
            "output.checkpoint_dir=.", "output.vtu_dir=.",
            f"environment.NODALS_FACE_YPLUS_EXPORT={export}",
            f"environment.NODALS_FACE_YPLUS_SP={plan.target_sp:.17g}",
            f"environment.NODALS_FACE_YPLUS_OMEGA={plan.target_omega:.17g}",
            f"environment.NODALS_FACE_YPLUS_MIN_SCALE={plan.min_scale:.17g}",
            f"environment.NODALS_FACE_YPLUS_MAX_SCALE={plan.max_scale:.17g}",
            "restart.path="]
    for s in sets:
        cmd += ["--set", s]
    if index == 0:
        cmd += ["--set", "environment.NODALS_FACE_YPLUS_UTAU_CSV="]  # no stale heights on a fresh start
    else:
        assert prev is not None
        cmd += ["--restart", str(prev / f"NODALS_STAGE31_STATE_it{begin}.ckpt"),
                "--set", f"environment.NODALS_FACE_YPLUS_UTAU_CSV={prev / f'NODALS_FACE_UTAU_it{begin}.csv'}"]
    cmd += ["--steps", str(end - begin), "--output", str(Plan.stage_dir(plan.output, begin, end))]
    return cmd


def _nonempty(p: Path) -> bool:
    return p.is_file() and p.stat().st_size > 0


class Runner:
    def __init__(self, plan: Plan, resume: bool, dry_run: bool):
        self.plan, self.resume, self.dry = plan, resume, dry_run
        self.manifest = {"case": str(plan.case), "mesh": str(plan.mesh), "refreshes": plan.refreshes,
                         "stages": plan.stages, "final_iterations": plan.final,
                         "target_sp_yplus": plan.target_sp, "target_omega_yplus": plan.target_omega,
                         "started_utc": datetime.now().isoformat(), "overrides": plan.overrides, "completed": []}

    def _say(self, line: str) -> None:
        print(line, flush=True)

    def _save_manifest(self) -> None:
        if not self.dry:
            (self.plan.output / "adaptive_schedule.json").write_text(json.dumps(self.manifest, indent=2) + "\n")

    def _outputs(self, seg: Path, end: int) -> Tuple[Path, Path]:
        return seg / f"NODALS_STAGE31_STATE_it{end}.ckpt", seg / f"NODALS_FACE_UTAU_it{end}.csv"

    def _complete(self, seg: Path, end: int) -> bool:
        ck, face = self._outputs(seg, end)
        return _nonempty(ck) and (not self.plan.refreshes or _nonempty(face))

    def announce(self) -> None:
        p = self.plan
        self._say(f"NODALS_ADAPTIVE_CASE={p.case}")
        self._say(f"NODALS_ADAPTIVE_MESH={p.mesh}")
        self._say(f"NODALS_ADAPTIVE_OUTPUT={p.output}")
        self._say(f"NODALS_ADAPTIVE_PLAN final={p.final} refreshAt={p.refreshes} spYplus={p.target_sp:g} "
                  f"omegaYplus={p.target_omega:g}")
        for i, (b, e) in enumerate(p.stages):
            self._say(f"NODALS_ADAPTIVE_STAGE index={i} from={b} to={e} "
                      f"sampling={'fixed' if i == 0 else 'target_yplus'} refreshCount={i}")

    def run(self) -> int:
        p = self.plan
        self.announce()
        if not self.dry:
            p.output.mkdir(parents=True, exist_ok=True)
            if not self.resume:  # only the runner's own stage directories are replaced
                for child in p.output.iterdir():
                    if child.is_dir() and STAGE_DIR.fullmatch(child.name):
                        shutil.rmtree(child)
        self._save_manifest()
        prev: Optional[Path] = None
        for i, (b, e) in enumerate(p.stages):
            seg = Plan.stage_dir(p.output, b, e)
            if self.resume and not self.dry and self._complete(seg, e):
                self._say(f"NODALS_ADAPTIVE_SKIP_COMPLETE from={b} to={e}")
                prev = seg
                self.manifest["completed"].append(e)
                continue
            cmd = segment_command(p, i, prev)
            if self.dry:
                self._say(f"NODALS_ADAPTIVE_DRY_CMD from={b} to={e}: {shlex.join(cmd)}")
                prev = seg
                continue
            if i > 0:
                ck, face = self._outputs(prev, b)
                if not (ck.is_file() and face.is_file()):
                    raise RuntimeError(f"Missing stage inputs: {ck} / {face}")
            if seg.exists():
                shutil.rmtree(seg)
            self._say(f"NODALS_ADAPTIVE_START from={b} to={e} refresh={i} cmd={shlex.join(cmd)}")
            env = dict(os.environ)
            root = str(Path(__file__).resolve().parents[1])
            env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
            rc = subprocess.call(cmd, env=env)
            if rc not in SEGMENT_OK:
                raise RuntimeError(f"NodalS segment {b}->{e} exited {rc}; see {seg}/terminal.log")
            ck, face = self._outputs(seg, e)
            if not _nonempty(ck):
                raise RuntimeError(f"Segment {b}->{e} did not save checkpoint {ck}")
            if p.refreshes and not _nonempty(face):
                raise RuntimeError(f"Segment {b}->{e} did not export F2 wall u_tau {face}")
            if i > 0:
                log = (seg / "terminal.log").read_text(errors="replace")
                if "STAGE31_RESTART_LOAD=PASS" not in log or not re.search(r"NODALS_FACE_YPLUS_LOADED[^\n]*status=PASS",
                                                                           log):
                    raise RuntimeError(f"Segment {b}->{e} did not verify both checkpoint restart and adaptive ray "
                                       "loading")
            prev = seg
            self.manifest["completed"].append(e)
            self._save_manifest()
            self._say(f"NODALS_ADAPTIVE_PASS from={b} to={e} checkpoint={ck} faceUTau={face}")
        self._say(f"NODALS_ADAPTIVE_{'DRY_RUN' if self.dry else 'COMPLETE'} iteration={p.final} output={p.output}")
        return 0


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(description="Run an SST .case in segments, moving the wall samples to a target y+")
    ap.add_argument("--case", type=Path, required=True, help="SST .case with an [adaptive_sampling] section")
    ap.add_argument("--mesh", type=Path, help="Override [mesh] path")
    ap.add_argument("--output", type=Path, help="Override [output] directory")
    ap.add_argument("--final-iterations", type=int, help="Override [simple] max_iterations")
    ap.add_argument("--set", action="append", default=[], metavar="SECTION.KEY=VALUE")
    ap.add_argument("--binary", type=Path, help="Native SST solver of every segment (default: the launcher's)")
    ap.add_argument("--cache", type=Path, help="HXT1 mesh cache of every segment (default: the launcher's)")
    ap.add_argument("--resume", action="store_true", help="Skip complete stages; rerun incomplete ones")
    ap.add_argument("--dry-run", action="store_true", help="Print the plan and the segment commands only")
    return ap


def main(argv: Optional[Sequence[str]] = None) -> int:
    a = parser().parse_args(argv)
    try:
        case = _path(a.case, Path.cwd())
        if not case.is_file():
            raise FileNotFoundError(case)
        plan = make_plan(case, a.set, a.final_iterations, a.mesh, a.output)
        plan.binary = _path(a.binary, Path.cwd()) if a.binary else None
        plan.cache = _path(a.cache, Path.cwd()) if a.cache else None
        return Runner(plan, a.resume, a.dry_run).run()
    except (OSError, ValueError, RuntimeError, configparser.Error) as exc:
        print("NODALS_ADAPTIVE_ERROR: " + str(exc), file=sys.stderr, flush=True)
        return 2


if __name__ == "__main__":
    raise SystemExit(main())
