"""`.case` front-end of the H8 GPU solver (``bin/nodals-gpu CASE``).

Interface kept from the reference launcher (scripts/nodals_gpu_case.py): the
INI sections and keys with their defaults, strict name checking, the
``--set section.key=value`` overrides, the H8 flags each key becomes
(parser: gpu/serial_cuda/h8_main_base.cu:1787-1911), the executable lookup
order (``--exe``, ``NODALS_GPU_<PRECISION>_EXE``, the package's ``bin/``,
``<prefix>/libexec/nodals``), the ``NODALS_GPU_*`` launcher lines, PASS
detection on ``NODALS_GPU_RESULT gate=H8 ... status=PASS`` and the exit
codes (0 pass, 21 not converged, 2 solver missing, 4 bad case, 124 time
limit, otherwise the solver's own status).

The translation itself is the key table `H8_SCHEMA` (one row per case key)
plus the cross-key rules below it, evaluated by :mod:`nodals_amd.casefile`.

Extensions of this build, off by default: ``[pressure] solver = fgmres`` with
``[pressure] restart``, and ``[run] precision = bf16`` (BF16 finest AMG
level).
"""
from __future__ import annotations

import argparse
import datetime as _dt
import json
import os
import shlex
import sys
from pathlib import Path
from typing import List, Mapping, Optional, Sequence

from .casefile import (CaseDoc, CaseError, Key, Schema, boolean, integer, number, one_of, only, path_text,
                       run_logged, text)

EXIT_PASS, EXIT_MISSING_SOLVER, EXIT_BAD_CASE, EXIT_NOT_CONVERGED, EXIT_TIMEOUT = 0, 2, 4, 21, 124


def _fixed10_needs_ten(v: Mapping[str, str]) -> None:
    if v["run.run_mode"] == "fixed10" and v["simple.max_iterations"] != "10":
        raise CaseError("[run] run_mode=fixed10 requires [simple] max_iterations=10")


def _supg_forms(v: Mapping[str, str]) -> None:
    if v["supg.enabled"] != "1":
        return
    if v["supg.form"] != "implicit":
        raise CaseError("current H8 GPU SUPG supports only [supg] form=implicit")
    if v["supg.quad_points"] != "64":
        raise CaseError("current H8 GPU SUPG supports only [supg] quad_points=64")


def _cheb2_needs_spectrum(v: Mapping[str, str]) -> None:
    if v["pressure_amg.smoother"] == "cheb2" and v["pressure_amg.spectrum_policy"] == "off":
        raise CaseError("[pressure_amg] smoother=cheb2 requires spectrum_policy=auto or always")


def _is_fgmres(v: Mapping[str, str]) -> bool:
    return v["pressure.solver"] == "fgmres"


H8_SCHEMA = Schema(
    label="GPU case",
    required=("solver.mesh",),
    keys=(
        # [run]: name is the solver tag; the rest steer the launcher itself.
        Key("run.name", "nodals_gpu", "--tag"),
        Key("run.precision", "fp32", None, one_of("fp32", "fp64", "bf16",
                                                  message="[run] precision must be fp32 or fp64 (or bf16)")),
        Key("run.run_mode", "converge", "--run-mode", one_of("fixed10", "converge",
                                                             message="[run] run_mode must be fixed10 or converge")),
        Key("run.output_root", "auto"),
        Key("run.timeout_seconds", "0", None, integer),
        # [solver], [boundary]
        Key("solver.mesh", None, "--mesh", path_text),
        Key("solver.re", "20", "--re", number),
        Key("solver.bulk_velocity", "1.0", "--bulk", number),
        Key("boundary.wall_patch", "patch_0_0", "--wall"),
        Key("boundary.inlet_patch", "patch_2_0", "--inlet"),
        Key("boundary.outlet_patch", "patch_1_0", "--outlet"),
        # [simple]
        Key("simple.rtol", "1e-3", "--simple-tol", number),
        Key("simple.max_iterations", "2500", "--max-outer", integer),
        Key("simple.alpha_u", "0.5", "--alpha-u", number),
        Key("simple.alpha_p", "0.5", "--alpha-p", number),
        # [momentum]
        Key("momentum.work", "fgs1", "--momentum-work", only("fgs1", "current H8 supports only [momentum] work=fgs1")),
        Key("momentum.rtol", "1e-6", "--mom-rtol", number),
        Key("momentum.atol", "1e-12", "--mom-atol", number),
        Key("momentum.relative_drop", "0.1", "--mom-drop", number),
        Key("momentum.omega", "1.0", "--mom-omega", number),
        Key("momentum.max_iterations", "20000", "--mom-max", integer),
        # [supg]: the implicit 64-point SUPG of the native solver.
        Key("supg.enabled", "false", "--supg", boolean),
        Key("supg.tau_scale", "0.05", "--supg-tau-scale", number),
        Key("supg.magic", "9.0", "--supg-magic", number),
        Key("supg.form", "implicit", "--supg-form", lambda r, w: r.strip().lower()),
        Key("supg.quad_points", "64", "--supg-quad-points", integer),
        # [pressure]
        Key("pressure.solver", "richardson", "--pressure-solver",
            one_of("pcg", "richardson", "cheb", "fgmres",
                   message="[pressure] solver must be pcg, richardson, cheb or fgmres")),
        Key("pressure.rtol", "0.5", "--p-rtol", number),
        Key("pressure.atol", "1e-12", "--p-atol", number),
        Key("pressure.max_iterations", "20", "--p-max", integer),
        Key("pressure.snapshot_tolerance", "5e-6", "--snapshot-tol", number),
        Key("pressure.fine_csr_refresh_every", "1", "--fine-csr-refresh-every",
            only("1", "current H8 requires [pressure] fine_csr_refresh_every=1")),
        Key("pressure.richardson_omega", "1.0", "--pressure-richardson-omega", number),
        Key("pressure.chebyshev_degree", "3", "--pressure-cheb-degree", integer),
        Key("pressure.power_iterations", "6", "--pressure-power-its", integer),
        Key("pressure.chebyshev_low_fraction", "0.05", "--pressure-cheb-low-fraction", number),
        Key("pressure.power_safety", "1.15", "--pressure-power-safety", number),
        Key("pressure.restart", "30", "--pressure-restart", integer, when=_is_fgmres),
        # [pressure_amg]
        Key("pressure_amg.hierarchy", "cf", "--amg-hierarchy",
            one_of("sa", "cf", message="[pressure_amg] hierarchy must be sa or cf")),
        Key("pressure_amg.smoother", "jacobi", "--amg-smoother",
            one_of("cheb2", "jacobi", "l1jacobi", message="[pressure_amg] smoother must be cheb2, jacobi, or l1jacobi")),
        Key("pressure_amg.terminal_rows", "1000", "--amg-terminal", integer),
        Key("pressure_amg.spectrum_policy", "auto", "--amg-spectrum-policy",
            one_of("auto", "always", "off", message="[pressure_amg] spectrum_policy must be auto, always, or off")),
        Key("pressure_amg.power_iterations", "16", "--amg-power-its", integer),
        Key("pressure_amg.chebyshev_degree", "2", "--amg-cheb-degree", integer),
        Key("pressure_amg.lambda_safety", "1.5", "--amg-lambda-safety", number),
        Key("pressure_amg.lambda_low_fraction", "0.05", "--amg-lambda-low-fraction", number),
        Key("pressure_amg.jacobi_omega", "0.7", "--amg-jacobi-omega", number),
        # [pressure_amg_cf]
        Key("pressure_amg_cf.coarsening", "pmis", "--cf-coarsening",
            only("pmis", "current H8 CF hierarchy supports only [pressure_amg_cf] coarsening=pmis")),
        Key("pressure_amg_cf.strength", "classical-negative", "--cf-strength",
            only("classical-negative",
                 "current H8 CF hierarchy supports only [pressure_amg_cf] strength=classical-negative")),
        Key("pressure_amg_cf.theta", "0.25", "--cf-theta", number),
        Key("pressure_amg_cf.interpolation", "exti", "--cf-interp",
            one_of("direct", "exti", message="[pressure_amg_cf] interpolation must be direct or exti")),
        Key("pressure_amg_cf.pmax", "8", "--cf-pmax", integer),
        Key("pressure_amg_cf.aggressive_first", "false", "--cf-aggressive-first", boolean),
    ),
    rules=(_fixed10_needs_ten, _supg_forms, _cheb2_needs_spectrum),
)


# ---------------------------------------------------------------- public API


def read_case(path: Path) -> CaseDoc:
    """Read a case file and check every section and key name."""
    doc = CaseDoc.read(path)
    H8_SCHEMA.check_names(doc)
    return doc


def apply_overrides(doc: CaseDoc, overrides: Sequence[str]) -> None:
    doc.apply(H8_SCHEMA, overrides)


def build_gpu_options(doc: CaseDoc) -> List[str]:
    """The H8 command line for a case (flag, value pairs)."""
    return H8_SCHEMA.argv(doc)


def precision(doc: CaseDoc) -> str:
    return H8_SCHEMA.resolve(doc)["run.precision"]


def find_solver(precision_name: str, explicit: Optional[str] = None) -> Path:
    """First executable among: `explicit`, $NODALS_GPU_<P>_EXE, the
    package's bin/, <prefix>/libexec/nodals.  With none present, the first
    candidate is returned (and reported missing by the caller)."""
    binary = f"nodals_gpu_h8_{precision_name}"
    pkg = Path(__file__).resolve().parent
    env_value = os.environ.get(f"NODALS_GPU_{precision_name.upper()}_EXE")
    options = [path_text(explicit, "") if explicit else None, path_text(env_value, "") if env_value else None,
               str(pkg / "bin" / binary), str(pkg.parent / "libexec" / "nodals" / binary)]
    found = [Path(o) for o in options if o]
    for cand in found:
        if cand.is_file() and os.access(cand, os.X_OK):
            return cand
    return found[0]


def _log_passed(log_path: Path) -> bool:
    with log_path.open("r", encoding="utf-8", errors="replace") as f:
        return any(line.startswith("NODALS_GPU_RESULT gate=H8 ") and "status=PASS" in line for line in f)


def _output_root(doc_value: str, name: str) -> Path:
    if doc_value.strip().lower() != "auto":
        return Path(path_text(doc_value, ""))
    stamp = _dt.datetime.now().strftime("%Y%m%d_%H%M%S")
    return Path.home() / "Downloads" / f"NodalS_GPU_{name}_{stamp}"


def _parse_cli(argv: Sequence[str]):
    own, extra = list(argv), []
    if "--" in own:
        cut = own.index("--")
        own, extra = own[:cut], own[cut + 1:]
    ap = argparse.ArgumentParser(prog="nodals-gpu", description="Run the NodalS H8 GPU solver from a .case file")
    ap.add_argument("case", type=Path)
    ap.add_argument("--dry-run", action="store_true", help="print the solver command and stop")
    ap.add_argument("--dump-options-json", action="store_true", help="also print the options as JSON")
    ap.add_argument("--exe", help="solver executable to use")
    ap.add_argument("--set", action="append", default=[], metavar="SECTION.KEY=VALUE")
    return ap.parse_args(own), extra


def main(argv: Sequence[str]) -> int:
    ns, extra = _parse_cli(argv)
    case_path = ns.case.expanduser().resolve()
    doc = read_case(case_path)
    apply_overrides(doc, ns.set)
    vals = H8_SCHEMA.resolve(doc)
    options = H8_SCHEMA.argv(doc)
    prec = vals["run.precision"]
    exe = find_solver(prec, ns.exe)
    name = doc.get("run", "name") or case_path.stem
    outroot = _output_root(vals["run.output_root"], name)
    cmd = [str(exe), *options, *extra]

    print(f"NODALS_GPU_CASE name={name} file={case_path} precision={prec} output={outroot}")
    print("NODALS_GPU_COMMAND " + shlex.join(cmd))
    if ns.dump_options_json:
        print("NODALS_GPU_OPTIONS_JSON=" + json.dumps(options))
    if ns.dry_run:
        print("NODALS_GPU_CASE_RESULT status=DRY_RUN")
        return EXIT_PASS
    if not (exe.is_file() and os.access(exe, os.X_OK)):
        print(f"NODALS_GPU_CASE_RESULT status=FAIL reason=missing_solver executable={exe}", file=sys.stderr)
        return EXIT_MISSING_SOLVER
    sys.stdout.flush()
    log_path = outroot / f"{name}.log"
    rc = run_logged(cmd, log_path, timeout_s=float(vals["run.timeout_seconds"]))
    print(f"NODALS_GPU_ARTIFACTS log={log_path} outroot={outroot}")
    if rc != 0:
        print(f"NODALS_GPU_CASE_RESULT status=FAIL rc={rc}")
        return rc
    if _log_passed(log_path):
        print("NODALS_GPU_CASE_RESULT status=PASS")
        return EXIT_PASS
    print("NODALS_GPU_CASE_RESULT status=NOT_CONVERGED reason=no_NODALS_GPU_RESULT_PASS")
    return EXIT_NOT_CONVERGED


def cli() -> None:
    try:
        code = main(sys.argv[1:])
    except (CaseError, OSError) as e:
        print(f"NODALS_GPU_CASE_ERROR {e}", file=sys.stderr)
        code = EXIT_BAD_CASE
    raise SystemExit(code)


if __name__ == "__main__":
    cli()
