"""SST .case front end: a case file to the native SST solver's argv and
environment, and the run driver around it.

The interface restated here is the reference's case launcher,
turbulence/sst_gpu_v2/nodals.py (``nodals.py run``): the case sections and
keys and their defaults (:41-53 reading, :120-420 validation and the
environment table, :432-509 the native argv), the dry run and its report
lines (:660-668), the run directory with ``invocation.json`` and
``terminal.log`` (:676-704), the compact console filter (:97-113) and the
``NODALS_NATIVE_EXIT_CODE`` line.  The argv flag names and the environment
variable names and values are the native solver's interface, so they are
kept exactly; the native solver this build runs by default is its own
``nodals_amd/bin/nodals_sst``.

Differences, each an error rather than a silent change: developed inlet
profiles (``[inlet.main] profile_csv``) and hexahedral meshes are not
supported.
"""
from __future__ import annotations

import argparse
import configparser
import datetime
import hashlib
import json
import math
import os
import re
import shlex
import subprocess
import sys
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence

PACKAGE = Path(__file__).resolve().parent
DEFAULT_BINARY = PACKAGE / "bin" / "nodals_sst"
DEFAULT_CACHE = PACKAGE.parent / ".cache" / "meshes"
REQUIRED_SECTIONS = ("mesh", "physics", "inlet.main", "simple", "sst", "solver.u")
MOMENTUM_MODES = {"m3-fgs": "fgs", "m3-sgs": "sgs", "m4-hybrid": "m4", "coupled-fgmres": "off"}
# Physical wall-sample heights of the original case runner for a 0.004 m
# pipe; scaled by diameter / 0.004 and the sample fraction (nodals.py:290-293).
BASE_HEIGHTS = (1.4694676501540146e-4, 2.9389353003080293e-4, 4.4084029504620439e-4,
                5.8778706006160585e-4, 7.3473382507700737e-4)
CONSOLE_PREFIXES = ("NODALS_SST_RESIDUALS ", "NODALS_SST_PLOT_METRICS ", "NODALS_SST_G2_VTU_CHECKPOINT ",
                    "STAGE31_RESTART_SAVE=", "STAGE31_RESTART_LOAD=", "NODALS_SST_G2_FIXED10_TIMING ",
                    "NODALS_SOFTWARE_COMPLETION=", "HXT5B_ERROR ", "NODALS_SOFTWARE_ERROR ", "CUDA_ERROR",
                    "NODALS_SST_G2_FINAL ", "NODALS_METRIC_GATE ")
STALE_ENV = re.compile(r"^(SST|STAGE|NODALS|P_AMG|HXT)")


class CaseError(ValueError):
    pass


def fmt(x) -> str:
    """Shortest round-trip decimal of a double (17 significant digits)."""
    return format(float(x), ".17g")


def compact_line(line: str) -> bool:
    """Lines the compact console shows (terminal.log keeps everything)."""
    return line.startswith(CONSOLE_PREFIXES)


class Case:
    """A parsed .case file with typed accessors; --set overrides applied."""

    def __init__(self, path: Path, overrides: Sequence[str] = ()):
        cp = configparser.ConfigParser(interpolation=None, inline_comment_prefixes=("#", ";"), strict=True)
        if not cp.read(path):
            raise FileNotFoundError(f"Case file not found: {path}")
        missing = [s for s in REQUIRED_SECTIONS if not cp.has_section(s)]
        if missing:
            raise CaseError(f"Missing required case section [{missing[0]}]")
        for item in overrides:
            if "=" not in item or "." not in item.split("=", 1)[0]:
                raise CaseError(f"--set expects SECTION.KEY=VALUE: {item}")
            lhs, value = item.split("=", 1)
            section, key = lhs.rsplit(".", 1)
            if not cp.has_section(section):
                cp.add_section(section)
            cp.set(section, key, value)
        self.cp = cp
        self.path = Path(path)

    def text(self, section: str, key: str, default: Optional[str] = None) -> Optional[str]:
        return self.cp.get(section, key, fallback=default)

    def word(self, section: str, key: str, default: str) -> str:
        return str(self.text(section, key, default)).strip().lower()

    def num(self, section: str, key: str, default=None) -> float:
        raw = self.text(section, key, None if default is None else str(default))
        if raw is None:
            raise CaseError(f"Required numeric key missing: {section}.{key}")
        value = float(raw)
        if not math.isfinite(value):
            raise CaseError(f"Nonfinite {section}.{key}")
        return value

    def count(self, section: str, key: str, default: int) -> int:
        value = int(self.text(section, key, str(default)))
        if value < 0:
            raise CaseError(f"Negative value for {section}.{key}")
        return value

    def flag(self, section: str, key: str, default: bool = False) -> bool:
        raw = str(self.text(section, key, str(default))).strip().lower()
        if raw in ("1", "yes", "true", "on"):
            return True
        if raw in ("0", "no", "false", "off"):
            return False
        raise CaseError(f"Expected boolean true/false for {section}.{key}; got {raw!r}")

    def vec3(self, section: str, key: str, default: Optional[str] = None) -> List[float]:
        raw = str(self.text(section, key, default)).strip().strip("()[]")
        vals = [float(v.strip()) for v in raw.split(",")]
        if len(vals) != 3 or not all(math.isfinite(v) for v in vals):
            raise CaseError(f"{section}.{key} must be a finite 3-component vector")
        return vals


def _resolve(raw, parent: Path) -> Path:
    p = Path(os.path.expandvars(os.path.expanduser(str(raw))))
    return (p if p.is_absolute() else parent / p).resolve()


def _child_dir(raw: Optional[str], output: Path) -> Path:
    raw = (raw or "").strip()
    return output if raw in ("", ".") else _resolve(raw, output)


@dataclass
class Invocation:
    argv: List[str]
    env: Dict[str, str]
    settings: Dict[str, object] = field(default_factory=dict)


def _inlet(case: Case):
    mode = case.word("inlet.main", "mode", "velocity")
    if mode == "velocity":
        vel = case.vec3("inlet.main", "velocity")
        bulk = math.sqrt(sum(v * v for v in vel))
    elif mode == "speed":
        bulk = case.num("inlet.main", "speed")
        d = case.vec3("inlet.main", "direction", "(1,0,0)")
        n = math.sqrt(sum(v * v for v in d))
        if not n:
            raise CaseError("Cannot use zero inlet direction")
        vel = [bulk * v / n for v in d]
    else:
        raise CaseError("Supported inlet modes: velocity, speed (with direction)")
    if bulk <= 0:
        raise CaseError("Inlet velocity must be nonzero")
    return vel, bulk


def _sample_heights(case: Case, diameter: float, scale: float) -> List[float]:
    raw = case.text("wall_function", "distances")
    if raw:
        try:
            heights = [float(v.strip()) for v in raw.split(",")]
        except ValueError as e:
            raise CaseError("wall_function.distances must contain five numbers") from e
    else:
        heights = [d * diameter / 0.004 * scale for d in BASE_HEIGHTS]
    ok = len(heights) == 5 and all(math.isfinite(y) for y in heights) and 0 < heights[0] and all(
        heights[i] < heights[i + 1] for i in range(4))
    if not ok:
        raise CaseError("wall_function.distances must be five strictly increasing positive values")
    return heights


def build_invocation(case: Case, opts, mesh: Path, mesh_sha: Optional[str], output: Path,
                     warn=lambda m: print(m, file=sys.stderr)) -> Invocation:
    """argv, environment and a summary of the settings for one native run."""
    if case.text("physics", "density", "1") and abs(case.num("physics", "density", 1) - 1) > 1e-12:
        raise CaseError("Current native solver uses kinematic pressure; physics.density must be 1")
    diameter = case.num("mesh", "diameter")
    if diameter <= 0:
        raise CaseError("mesh.diameter must be positive")
    reynolds = case.num("physics", "reynolds")
    if reynolds <= 0:
        raise CaseError("physics.reynolds must be positive")
    vel, bulk = _inlet(case)
    nu = bulk * diameter / reynolds
    if case.text("physics", "kinematic_viscosity") is not None:
        given = case.num("physics", "kinematic_viscosity")
        if given <= 0 or abs(given / nu - 1) > 0.001:
            raise CaseError("inlet speed, diameter, Reynolds and kinematic_viscosity disagree (>0.1%)")
        nu = given

    start = case.word("initialization", "mode", "potential_flow")
    if start not in ("potential_flow", "plug"):
        raise CaseError("initialization.mode must be potential_flow or plug")
    state = case.word("sst", "state_space", "q1")
    if state not in ("q1", "bernstein_q2", "q1bf2_unlimited"):
        raise CaseError("sst.state_space must be q1, bernstein_q2 or q1bf2_unlimited")
    if state != "q1":
        warn("WARNING: non-reference scalar discretization selected; numerical parity not established.")
    mode = (getattr(opts, "momentum_solver", None) or case.word("solver.u", "method", "m3-fgs")).lower()
    if mode not in MOMENTUM_MODES:
        raise CaseError("momentum mode must be " + ", ".join(MOMENTUM_MODES))
    convection = case.word("convection", "form", "advective")
    if convection not in ("advective", "conservative", "skew_symmetric", "emac"):
        raise CaseError("convection.form must be advective, conservative, skew_symmetric, or emac")
    if convection in ("conservative", "skew_symmetric"):
        if state != "q1":
            raise CaseError("conservative momentum convection requires sst.state_space=q1")
        if mode == "m4-hybrid":
            raise CaseError("conservative momentum convection not supported with experimental M4")
    if convection == "emac":
        if state != "q1":
            raise CaseError("EMAC momentum requires sst.state_space=q1")
        if mode not in ("m3-fgs", "m3-sgs"):
            raise CaseError("EMAC momentum currently requires M3 FGS or SGS")
        if case.flag("supg", "momentum", False):
            raise CaseError("EMAC momentum does not support momentum SUPG yet")
    stress = case.word("stress", "form", "vector_laplacian")
    if stress not in ("vector_laplacian", "deviatoric"):
        raise CaseError("stress.form must be vector_laplacian or deviatoric")
    if stress == "deviatoric":
        if state != "q1":
            raise CaseError("deviatoric stress currently requires sst.state_space=q1")
        if mode == "m4-hybrid":
            raise CaseError("deviatoric stress not supported for experimental M4")
    if case.flag("supg", "k", False) or case.flag("supg", "omega", False):
        raise CaseError("Full strong-residual scalar SUPG is not verified for k or omega; both must be false")
    if case.flag("supg", "momentum", False):
        if stress != "deviatoric" or convection not in ("advective", "skew_symmetric", "conservative"):
            raise CaseError("Full momentum SUPG is only available with deviatoric stress and advective, "
                            "skew_symmetric or conservative convection")
        if state != "q1":
            raise CaseError("Full momentum SUPG requires sst.state_space=q1")
        if mode not in ("m3-fgs", "m3-sgs"):
            raise CaseError("Full momentum SUPG requires M3 FGS or SGS")
        if not 0 < case.num("supg", "nu_fd_rel", 0.001) <= 0.01:
            raise CaseError("supg.nu_fd_rel must be in (0,0.01]")
        if not case.num("supg", "tau_scale", 0.05) > 0:
            raise CaseError("supg.tau_scale must be positive")
    sweeps = case.count("solver.u", "sweeps", 5)
    if not 1 <= sweeps <= 60:
        raise CaseError("solver.u.sweeps outside 1..60")

    # absolute native iteration limit (a restart's _itNNN counts)
    limit = opts.outer_iterations if getattr(opts, "outer_iterations", None) is not None else \
        case.count("simple", "max_iterations", 10)
    if getattr(opts, "steps", None) is not None:
        if getattr(opts, "restart", None):
            m = re.search(r"_it(\d+)\.ckpt$", Path(opts.restart).name)
            if not m:
                raise CaseError("For --steps and --restart, checkpoint name must contain _itNNN.ckpt; "
                                "use --outer-iterations otherwise")
            limit = int(m.group(1)) + opts.steps
        else:
            limit = opts.steps
    if limit < 1:
        raise CaseError("Iteration limit must be positive")
    iterations = int(limit)

    print_every = case.count("output", "print_every", 10)
    metrics_every = case.count("output", "metrics_every", 10)
    check_every = case.count("sst", "check_every", 10)
    metric = case.word("convergence", "metric", "legacy")
    if metric not in ("legacy", "pressure_drop", "wall_traction", "both"):
        raise CaseError("convergence.metric must be legacy, pressure_drop, wall_traction or both")
    conv_tol = case.num("convergence", "relative_tolerance", 0.01)
    conv_window = case.count("convergence", "window_iterations", 25)
    conv_residuals = case.flag("convergence", "require_residuals", True)
    if not (0 < conv_tol < 1) or not (2 <= conv_window <= 10000):
        raise CaseError("convergence tolerance must be (0,1), window_iterations 2..10000")
    if metric != "legacy" and metrics_every != 1:
        raise CaseError("convergence.metric requires output.metrics_every=1 for consecutive SIMPLE iterations")
    if print_every < 1 or metrics_every < 1 or print_every % metrics_every:
        raise CaseError("output.print_every must be a positive multiple of output.metrics_every")
    if check_every > 0 and print_every % check_every:
        raise CaseError("sst.check_every must divide output.print_every (or be 0 to disable scalar true residuals)")

    samples = case.count("wall_function", "samples", 1)
    omega_sample = case.count("omega_wall", "sample_index", 3)
    if not (1 <= samples <= 5 and 1 <= omega_sample <= 5):
        raise CaseError("wall sample indices must be between 1 and 5")
    initial = case.word("wall_sampling", "initial_distance", "legacy_scale")
    if initial not in ("legacy_scale", "first_cell_centroid"):
        raise CaseError("wall_sampling.initial_distance must be legacy_scale or first_cell_centroid")
    auto = initial == "first_cell_centroid"

    def sampling(section, key, fallback):
        raw = case.word(section, key, fallback)
        if raw == "auto":
            if not auto:
                raise CaseError(f"{section}.{key}=auto requires [wall_sampling] initial_distance=first_cell_centroid")
            return 0.5, True
        return float(raw), False

    scale, auto_sp = sampling("wall_function", "spalding_sample_scale", "auto" if auto else "0.5")
    omega_scale, auto_om = sampling("omega_wall", "sample_scale", "auto" if auto else str(scale))
    if not (0.001 <= scale <= 1 and 0.001 <= omega_scale <= 1):
        raise CaseError("Sampling scales must be between 0.001 and 1")
    if case.text("wall_function", "law", "spalding") != "spalding":
        raise CaseError("Frozen reference supports wall_function.law=spalding")
    if case.text("omega_wall", "method", "of_auto") != "of_auto":
        raise CaseError("Frozen reference supports omega_wall.method=of_auto")
    omega_mode = case.word("omega_wall", "stage41_mode", "sample_log")
    if omega_mode not in ("sample_log", "of_auto", "trace_log", "couette_point3"):
        raise CaseError("omega_wall.stage41_mode must be sample_log, of_auto or couette_point3")
    if omega_mode == "trace_log":
        if omega_sample != 3:
            raise CaseError("trace_log requires omega sample index 3")
        if case.flag("omega_wall", "production", False):
            raise CaseError("trace_log requires production=false")
        if state != "q1":
            raise CaseError("trace_log currently supports masked-P1/Q1 SST")
        if not 10 <= case.num("omega_wall", "trace_gamma", 50.0) <= 100000:
            raise CaseError("omega_wall.trace_gamma must be between 10 and 100000")
    if omega_mode == "couette_point3":
        if samples != 5 or omega_sample != 3:
            raise CaseError("couette_point3 needs wall_function.samples=5 and omega_wall.sample_index=3")
        if case.flag("omega_wall", "production", False):
            raise CaseError("couette_point3 is omega-only: set omega_wall.production=false")
        if state != "q1":
            raise CaseError("couette_point3 is supported only for masked-P1/Q1 SST")
    heights = _sample_heights(case, diameter, scale)
    ray = case.word("wall_sampling", "mode", "ray_trace")
    if ray not in ("ray_trace", "legacy_endpoint"):
        raise CaseError("wall_sampling.mode must be ray_trace or legacy_endpoint")
    if ray == "legacy_endpoint":
        warn("WARNING: legacy_endpoint uses different wall sampling than the M3 reference ray_trace")
    crossings = case.count("wall_sampling", "max_crossings", 4096)
    if crossings < 1:
        raise CaseError("wall_sampling.max_crossings must be positive")
    bary = case.num("wall_sampling", "bary_tolerance", 1e-9)
    if bary <= 0:
        raise CaseError("wall_sampling.bary_tolerance must be positive")

    gamma_n = fmt(case.num("wall_function", "nitsche_gamma", 1e-6))
    k_sweeps = str(case.count("sst", "k_sweeps", 8))
    omega_sweeps = str(case.count("sst", "omega_sweeps", 8))
    checkpoint_every = str(case.count("output", "checkpoint_every", 100))
    env: Dict[str, str] = {
        "SST_STATE_SPACE": state,
        "SST_COMPACT_P1": str(int(case.flag("sst", "compact_p1", False))),
        "SST_COMPACT_QUAD": case.text("sst", "compact_quad", "24"),
        "NODALS_P2_STAGE5_RUN": "0",
        "NODALS_P2_STAGE4_TWO_STEP": "0",
        "NODALS_P2_NATIVE_PILOT": "0",
        "SST26_SPIRAL_DIAMETER": fmt(diameter),
        "SST23_INLET_VELOCITY": ",".join(fmt(v) for v in vel),
        "SST_WALL_DISTANCE_METHOD": case.text("wall_distance", "method", "geometric"),
        "SST_WALL_SAMPLING_MODE": ray,
        "NODALS_AUTO_CENTROID_MODE": str(int(auto)),
        "NODALS_AUTO_CENTROID_SP": str(int(auto_sp)),
        "NODALS_AUTO_CENTROID_OM": str(int(auto_om)),
        "NODALS_AUTO_CORNER_DEGREES": case.text("wall_sampling", "corner_angle_degrees", "45"),
        "NODALS_AUTO_THIN_POLICY": case.word("wall_sampling", "thin_channel_policy", "error"),
        "SST_WALL_RAY_MAX_CROSSINGS": str(crossings),
        "SST_WALL_RAY_BARY_TOLERANCE": fmt(bary),
        "STAGE27_S02B_PACK": "1",
        "STAGE27_INITIAL_VELOCITY": "potential" if start == "potential_flow" else "plug",
        "STAGE27_SPALDING_ZERO_LIMIT": str(int(case.flag("wall_function", "zero_limit", True))),
        "STAGE27_LOCAL_RAY_ADAPT": "1" if samples == 1 and omega_sample == 3 else "0",
        "SST24B_ACTIVE_POINTS": str(samples),
        "SST24B_OMEGA_SAMPLE_INDEX": str(omega_sample),
        "STAGE32_SPALDING_SAMPLE_SCALE": fmt(scale),
        "STAGE32_OMEGA_SAMPLE_SCALE": fmt(omega_scale),
        "SST22_TRACTION_MODEL": "vector_projected",
        "SST22D4_NITSCHE_ENABLE": "1",
        "SST22D4_GAMMA_N": gamma_n,
        "SST22D4I_BF3": "1",
        "SST22D4I3A_SCHUR": "1",
        "SST22D4D_MOM_MODE": "coupled",
        "SST22D4C_AUDIT_EVERY": str(iterations),
        "NODALS_MOMENTUM_STRESS_FORM": stress,
        "NODALS_MOMENTUM_CONVECTION_FORM": convection,
        "NODALS_M3_MODE": MOMENTUM_MODES[mode],
        "NODALS_M3_SWEEPS": str(sweeps),
        "NODALS_M3_OMEGA": fmt(case.num("solver.u", "omega", 0.7)),
        "NODALS_F1B_FROZEN_WALL": str(int(case.flag("wall_function", "frozen_wall_cache", True))),
        "NODALS_F2_FACE_SPALDING": str(int(case.flag("wall_function", "face_spalding", True))),
        "NODALS_F1_FACE_LOCAL": "0",
        "NODALS_MOM_WALL_OPERATOR": "original",
        "NODALS_P_REFRESH_EVERY": "1",
        "NODALS_F1B_GPU_GATE": "0",
        "STAGE27_WALL_AUDIT_MODE": "off",
        "STAGE27_S05A_K_MODE": case.text("sst", "k_solver", "fgs-fixed"),
        "STAGE27_S05A_OMEGA_MODE": case.text("sst", "omega_solver", "fgs-fixed"),
        "STAGE27_S05A_K_SWEEPS": k_sweeps,
        "STAGE27_S05A_OMEGA_SWEEPS": omega_sweeps,
        "STAGE27_S05A_CHECK_EVERY": str(check_every),
        "STAGE27_S05B2_FAST_SCALAR_ACTION": "1",
        "STAGE27_S05B1_MOM_MODE": "inherit",
        "STAGE27_S05B3_MOM_MODE": "inherit",
        "STAGE33A_MOMENTUM_SUPG": "0",
        "NODALS_MOMENTUM_FULL_SUPG": str(int(case.flag("supg", "momentum", False))),
        "NODALS_MOMENTUM_SUPG_NU_FD_REL": fmt(case.num("supg", "nu_fd_rel", 0.001)),
        "STAGE33BC_K_SUPG": str(int(case.flag("supg", "k", False))),
        "STAGE33BC_OMEGA_SUPG": str(int(case.flag("supg", "omega", False))),
        "STAGE33A_TAU_SCALE": fmt(case.num("supg", "tau_scale", 0.05)),
        "STAGE33BC_K_TAU_SCALE": fmt(case.num("supg", "k_tau_scale", 0.05)),
        "STAGE33BC_OMEGA_TAU_SCALE": fmt(case.num("supg", "omega_tau_scale", 0.05)),
        "NODALS_G2_CHILD_PROFILE": str(int(case.flag("output", "profile", False))),
        "NODALS_PREDICTOR_AUDIT": "0",
        "STAGE31_CHECKPOINT_EVERY": checkpoint_every,
        "STAGE31_RESTART_EVERY": checkpoint_every,
        "SST_G2_VTU_EVERY": str(case.count("output", "vtu_every", 100)),
        "NODALS_SST_DIAGNOSTIC_EVERY": str(case.count("output", "diagnostics_every", 0)),
        "NODALS_LEGACY_FINAL_DIAGNOSTICS": str(int(case.flag("output", "legacy_final_diagnostics", False))),
        "P_AMG_SMOOTHER": case.text("amg", "smoother", "jacobi"),
        "P_AMG_FINE_SWEEPS": str(case.count("amg", "fine_sweeps", 1)),
        "P_AMG_COARSE_SWEEPS": str(case.count("amg", "coarse_sweeps", 1)),
        "STAGE41_OMEGA_MODE": omega_mode,
        "STAGE41_TRACE_GAMMA": fmt(case.num("omega_wall", "trace_gamma", 50.0)),
        "STAGE27_S05B1_BLOCK_MAX": "3",
        "STAGE27_S05B1_COMPONENT_MAX": "6",
        "STAGE27_S05B1_PRECOND_FGS": "1",
        "SST22D4D_RESTART": str(case.count("solver.u", "restart", 30)),
        "SST22D4D_MAXITS": str(case.count("solver.u", "max_iterations", 800)),
    }
    env["SST_G2_VTU_DIR"] = str(_child_dir(case.text("output", "vtu_dir", ""), output))
    env["STAGE31_CHECKPOINT_DIR"] = str(_child_dir(case.text("output", "checkpoint_dir", ""), output))
    if mode == "m4-hybrid":
        warn("WARNING: M4 is experimental and has not passed GPU parity.")
    restart = getattr(opts, "restart", None)
    if restart:
        if not Path(restart).is_file():
            raise FileNotFoundError(f"Restart checkpoint missing: {restart}")
        env["STAGE31_RESTART_FROM"] = str(restart)
    if mesh_sha:
        fixed = {"mesh_sha256": mesh_sha, "profile_sha256": None, "re": reynolds, "nu": nu, "diameter": diameter,
                 "density": 1.0}
        env["STAGE31_INPUT_FINGERPRINT"] = hashlib.sha256(
            json.dumps(fixed, sort_keys=True, separators=(",", ":")).encode()).hexdigest()
    if restart and not mesh_sha and not case.text("restart", "input_fingerprint"):
        raise CaseError("Restart from preconverted HXT1 requires restart.input_fingerprint (or original .msh)")
    if case.text("restart", "input_fingerprint"):
        env["STAGE31_INPUT_FINGERPRINT"] = case.text("restart", "input_fingerprint")
    if case.text("inlet.main", "profile_csv", "").strip():
        raise CaseError("inlet.main.profile_csv: developed inlet profiles are not supported by this solver")
    if case.text("mesh", "patch_map"):
        patch = _resolve(case.text("mesh", "patch_map"), case.path.parent)
        if not patch.is_file():
            raise FileNotFoundError(f"Boundary patch map missing: {patch}")
        env["SST_PATCH_MAP_FILE"] = str(patch)
    if case.cp.has_section("environment"):
        blocked = {"NODALS_SKEW_STREAMLINE_DIFFUSION", "NODALS_SKEW_SD_TAU_SCALE", "NODALS_SKEW_FULL_SUPG",
                   "NODALS_SKEW_FULL_NU_FD_REL", "NODALS_CONSERVATIVE_STREAMLINE_DIFFUSION",
                   "NODALS_CONSERVATIVE_SD_TAU_SCALE", "NODALS_CONSERVATIVE_FULL_SUPG",
                   "NODALS_CONSERVATIVE_FULL_NU_FD_REL", "STAGE33A_MOMENTUM_SUPG", "NODALS_MOMENTUM_FULL_SUPG",
                   "NODALS_MOMENTUM_SUPG_NU_FD_REL"}
        for k, v in case.cp.items("environment"):
            name = k.upper()
            if not re.fullmatch(r"[A-Z_][A-Z_0-9]*", name):
                raise CaseError(f"Invalid environment key {k}")
            if name in blocked:
                raise CaseError(f"Deprecated/restricted momentum SUPG environment key {name}; "
                                "use [supg] momentum, tau_scale, nu_fd_rel")
            env[name] = v
    env["NODALS_CONV_METRIC"] = metric
    env["NODALS_CONV_RTOL"] = fmt(conv_tol)
    env["NODALS_CONV_WINDOW"] = str(conv_window)
    env["NODALS_CONV_REQUIRE_RESIDUALS"] = str(int(conv_residuals))

    binary = getattr(opts, "binary", None) or DEFAULT_BINARY
    argv = [str(binary), "--mesh", str(mesh), "--vtu", str(output / "flow_sst.vtu"),
            "--re", fmt(reynolds), "--bulk", fmt(bulk),
            "--sst-g2", "1", "--sst-g2-start", "plug", "--simple-coupling", "corrected",
            "--max-outer", str(iterations), "--sst-g2-max", str(iterations),
            "--sst-g2-physical-gate", str(int(case.flag("simple", "physical_gate", False))),
            "--sst-g2-print-every", str(print_every), "--sst-g2-metrics-every", str(metrics_every),
            "--sst-g2-cont-atol", fmt(case.num("simple", "continuity_atol", 1e-6)),
            "--sst-g2-plateau-tol", fmt(case.num("simple", "plateau_tol", 1e-2)),
            "--sst-g2-plateau-samples", str(case.count("simple", "plateau_samples", 5)),
            "--sst-shared-spalding-wall-state", str(int(case.flag("wall_function", "shared_wall_state", True))),
            "--sst-profile-integral-utau", str(int(case.flag("wall_function", "integral_utau", True))),
            "--sst-profile-integral-nq", str(case.count("wall_function", "integral_nq", 12)),
            "--sst-profile-integral-closure", "0", "--sst-profile-average-closure", "1",
            "--sst-profile-average-wall-law", "spalding_profile5", "--sst-profile-match-y", "0"]
    for i, y in enumerate(heights):
        argv += [f"--sst-profile-fit-y{i + 1}", fmt(y)]
    argv += ["--sst-of-production", str(int(case.flag("omega_wall", "production", False))),
             "--sst-of-kappa", fmt(case.num("sst", "of_kappa", 0.4)),
             "--sst-of-E", fmt(case.num("sst", "of_e", 9.025013499434122)),
             "--sst-of-beta1", fmt(case.num("sst", "of_beta1", 0.075)),
             "--sst-of-re-blend", fmt(case.num("sst", "of_re_blend", 11.0)),
             "--sst-intensity", fmt(case.num("sst", "intensity", 0.05)),
             "--sst-length-ratio", fmt(case.num("sst", "length_ratio", 0.07)),
             "--sst-alpha-k", fmt(case.num("sst", "ur_k", 0.7)),
             "--sst-alpha-omega", fmt(case.num("sst", "ur_omega", 0.7)),
             "--sst-k-solver", case.text("sst", "k_solver", "fgs-fixed"),
             "--sst-omega-solver", case.text("sst", "omega_solver", "fgs-fixed"),
             "--sst-k-fgs-sweeps", k_sweeps, "--sst-omega-fgs-sweeps", omega_sweeps,
             "--sst-fgs-check-every", str(check_every),
             "--sst-omega-wall-mode", case.text("omega_wall", "method", "of_auto"),
             "--sst-wall-penalty-gamma", fmt(case.num("omega_wall", "gamma", 1e6)),
             "--alpha-u", fmt(case.num("simple", "ur_u", 0.7)),
             "--alpha-p", fmt(case.num("simple", "ur_p", 0.7)),
             "--simple-tol", fmt(case.num("simple", "simple_tol", 1e-12)),
             "--gamma", gamma_n, "--momentum-work", "fgmres-fgs",
             "--mom-max", str(case.count("solver.u", "max_iterations", 800)),
             "--mom-rtol", fmt(case.num("solver.u", "rtol", 1e-6)),
             "--mom-atol", fmt(case.num("solver.u", "atol", 0.0)),
             "--mom-fgmres-restart", str(case.count("solver.u", "restart", 30)),
             "--pressure-solver", case.text("solver.p", "method", "pcg_amg"),
             "--p-rtol", fmt(case.num("solver.p", "rtol", 0.5)),
             "--p-atol", fmt(case.num("solver.p", "atol", 1e-12)),
             "--p-max", str(case.count("solver.p", "max_iterations", 600)),
             "--amg-hierarchy", case.text("amg", "hierarchy", "cf"),
             "--cf-coarsening", case.text("amg", "coarsening", "pmis"),
             "--cf-strength", case.text("amg", "strength", "classical-negative"),
             "--cf-interp", case.text("amg", "interpolation", "exti"),
             "--cf-theta", fmt(case.num("amg", "theta", 0.25)),
             "--amg-terminal", str(case.count("amg", "terminal", 1000)),
             "--amg-smoother", case.text("amg", "smoother", "jacobi"),
             "--amg-fine-mode", case.text("amg", "fine_mode", "explicit"),
             "--amg-jacobi-fine-sweeps", str(case.count("amg", "fine_sweeps", 1)),
             "--amg-jacobi-coarse-sweeps", str(case.count("amg", "coarse_sweeps", 1)),
             "--profile-audit", str(int(case.flag("output", "profile_audit", False)))]
    if case.cp.has_section("native"):
        for k, v in case.cp.items("native"):
            if not re.fullmatch(r"[a-z][a-z0-9-]*", k):
                raise CaseError(f"Invalid native CLI key {k}")
            argv += ["--" + k, v]
    for opt in getattr(opts, "native_option", None) or []:
        if "=" not in opt:
            raise CaseError("--native-option expects flag=value")
        k, v = opt.split("=", 1)
        if not re.fullmatch(r"[a-z][a-z0-9-]*", k):
            raise CaseError(f"Invalid native option {k}")
        argv += ["--" + k, v]
    settings = dict(reynolds=reynolds, bulk=bulk, diameter=diameter, nu=nu, scalar_space=state, momentum=mode,
                    outer_iteration_limit=iterations, wall_sampling=ray, profile_fit_heights=heights)
    return Invocation(argv, env, settings)


def prepare_mesh(source: Path, cache: Path, convert: bool):
    """(native mesh path, sha256 of a .msh source or None).  A .msh is
    converted once into cache/<sha256[:20]>/HXT1_mesh.bin."""
    suffix = source.suffix.lower()
    if suffix == ".msh":
        from nodals_amd import hxt1

        digest = hxt1.sha256(source)
        folder = cache / digest[:20]
        target = folder / "HXT1_mesh.bin"
        if convert and not target.is_file():
            print("NODALS_MESH_CONVERT=" + str(source), flush=True)
            hxt1.convert(source, folder)
        if convert and not target.is_file():
            raise RuntimeError(f"HXT1 conversion did not produce {target}")
        return target, digest
    if suffix in (".bin", ".hxt1bin", ".hxt1"):
        return source, None
    raise CaseError("Unsupported mesh format: expected .msh (Gmsh ASCII v2.2) or HXT1 binary")


def run(opts) -> int:
    case = Case(opts.case, opts.set)
    if opts.restart is None:
        path = (case.text("restart", "path", "") or "").strip()
        if path:
            opts.restart = _resolve(path, opts.case.parent)
    source = _resolve(opts.mesh or case.text("mesh", "path"), opts.case.parent)
    if not source.is_file():
        raise FileNotFoundError(f"Mesh not found: {source}")
    if case.word("mesh", "format", "msh2") not in ("msh2", "hxt1"):
        raise CaseError("mesh.format must be msh2 or HXT1")
    cache = _resolve(opts.cache, Path.cwd())
    if opts.output:
        out = _resolve(opts.output, Path.cwd())
    elif (case.text("output", "directory", "") or "").strip():
        out = _resolve(case.text("output", "directory"), opts.case.parent)
    else:
        stamp = datetime.datetime.now().strftime("%Y%m%d_%H%M%S_%f")
        out = PACKAGE.parent / "runs" / f"{opts.case.stem}_{stamp}"
    mesh, digest = prepare_mesh(source, cache, not opts.dry_run)
    inv = build_invocation(case, opts, mesh, digest, out)
    print(f"NODALS_SOFTWARE_RUN case={opts.case} mesh={mesh}")
    print(f"NODALS_SOFTWARE_MODE={inv.settings['momentum']} scalar={inv.settings['scalar_space']} "
          f"maxOuter={inv.settings['outer_iteration_limit']}")
    print("NODALS_SOFTWARE_STRESS_FORM=" + inv.env["NODALS_MOMENTUM_STRESS_FORM"])
    print("NODALS_SOFTWARE_MOMENTUM_CONVECTION=" + inv.env["NODALS_MOMENTUM_CONVECTION_FORM"])
    print("NODALS_NATIVE_CMD=" + shlex.join(inv.argv))
    if opts.dry_run:
        print("NODALS_DRY_RUN=YES; no mesh conversion, GPU invocation or run directory created")
        return 0
    if opts.prepare_only:
        print(f"NODALS_MESH_READY={mesh}; GPU solver not invoked")
        return 0
    binary = Path(inv.argv[0])
    if not binary.is_file():
        raise FileNotFoundError(f"Executable not present: {binary}")
    if not os.access(binary, os.X_OK):
        raise PermissionError(f"Executable is not executable: {binary}")
    out.mkdir(parents=True, exist_ok=True)
    for d in (inv.env["SST_G2_VTU_DIR"], inv.env["STAGE31_CHECKPOINT_DIR"]):
        Path(d).mkdir(parents=True, exist_ok=True)
    (out / "invocation.json").write_text(json.dumps(
        {"case": str(opts.case), "mesh_source": str(source), "compiled_mesh": str(mesh), "command": inv.argv,
         "environment_overrides": inv.env, "settings": inv.settings, "source_mesh_sha256": digest},
        indent=2) + "\n")
    console = case.word("output", "console", "compact")
    if console not in ("compact", "full"):
        raise CaseError("output.console must be compact or full")
    base = {k: v for k, v in os.environ.items() if not STALE_ENV.match(k)}
    with (out / "terminal.log").open("w") as log:
        proc = subprocess.Popen(inv.argv, env={**base, **inv.env}, cwd=out, stdout=subprocess.PIPE,
                                stderr=subprocess.STDOUT, text=True, errors="replace", bufsize=1)
        assert proc.stdout is not None
        for line in proc.stdout:
            log.write(line)
            if console == "full" or compact_line(line):
                print(line, end="", flush=True)
        rc = proc.wait()
    print(f"NODALS_NATIVE_EXIT_CODE={rc} output={out}")
    return rc


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(description="Run an SST .case with the native solver")
    sub = ap.add_subparsers(dest="command", required=True)
    p = sub.add_parser("run", help="Convert the mesh if needed and run the native solver")
    p.add_argument("--case", type=Path, required=True)
    p.add_argument("--mesh", type=Path, help="Override the case mesh")
    p.add_argument("--output", type=Path, help="Output directory (default: runs/<case>_<timestamp>)")
    p.add_argument("--binary", type=Path, default=DEFAULT_BINARY)
    p.add_argument("--cache", type=Path, default=DEFAULT_CACHE)
    p.add_argument("--outer-iterations", type=int, help="Absolute native iteration limit, restart count included")
    p.add_argument("--steps", type=int, help="Iterations from a fresh start or from an _itNNN.ckpt restart")
    p.add_argument("--restart", type=Path, help="Checkpoint to restart from")
    p.add_argument("--momentum-solver", choices=list(MOMENTUM_MODES))
    p.add_argument("--set", action="append", default=[], help="Override SECTION.KEY=VALUE (repeatable)")
    p.add_argument("--native-option", action="append", default=[], help="Pass --FLAG VALUE to the solver")
    p.add_argument("--dry-run", action="store_true", help="Check the input and print the invocation only")
    p.add_argument("--prepare-only", action="store_true", help="Convert the mesh, do not run")
    return ap


def main(argv: Optional[Sequence[str]] = None) -> int:
    opts = parser().parse_args(argv)
    try:
        return run(opts)
    except (ValueError, RuntimeError, FileNotFoundError, PermissionError, subprocess.CalledProcessError) as e:
        print("NODALS_SOFTWARE_ERROR: " + str(e), file=sys.stderr)
        return 2


if __name__ == "__main__":
    sys.exit(main())
