#!/usr/bin/env python3
"""Benchmark: seconds-to-converge of the H8 SIMPLE pipeline on one MI355X.

Workload (one "step" = one SIMPLE solve from rest to the convergence test,
setup excluded): laminar straight circular pipe, D=1, L=10D, Re=20, O-grid
section of 224 quads x 30 slabs, 6 tets per hex (40,320 tets), Poiseuille
Dirichlet inlet, zero-pressure outlet, no-slip wall; FGS momentum sweep,
SA-AMG-preconditioned FGMRES pressure solve, to continuity and momentum
residual 1e-6.  `value` is the mean wall-clock seconds of the FP32-AMG form
(AMG hierarchy and smoother in FP32 inside the FP64 FGMRES); the plain FP64
form is timed beside it.  Inputs (mesh, setup, operators) are resident in
HBM before the timed region.

Also reported: the fine pressure SpMV roofline (algorithmic bytes per launch
over the mean launch time from HIP events on the solver stream) and the CPU
oracle (FP64 numpy/scipy restatement) timed on a bounded sample of the same
case on the host.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "s-to-converge 207k SST (res 1e-5): learned+FP32 AMG vs FP64; SpMV GB/s"
HBM_PEAK_GBS = 8000.0


# The pressure tolerance is the reference binary's default p-rtol 0.5
# (gpu/serial_cuda/h8_main_base.cu:1753); the SIMPLE iteration count
# to 1e-6 is the same from 0.05 to 0.5 (1373-1377 iterations, same Delta p to
# 1e-8), so a tighter inner tolerance only costs time.
def solver_args(mesh: str, precision: str) -> list:
    return ["--mesh", mesh, "--run-mode", "converge", "--max-outer", "20000", "--simple-tol", "1e-6",
            "--alpha-u", "0.5", "--alpha-p", "0.5", "--amg-smoother", "jacobi", "--amg-jacobi-omega", "0.7",
            "--pressure-solver", "fgmres", "--p-rtol", "0.5", "--p-max", "200", "--pressure-restart", "30",
            "--amg-precision", precision, "--verbose", "0", "--tag", f"bench_{precision}"]


def cpu_baseline(mesh: str, gpu_outer: int, budget_s: float, dp_gpu: float) -> dict:
    """The FP64 C++/OpenMP port of the same SIMPLE loop (oracle/cpu_h8.cpp)
    on the host's cores, from rest, for the GPU's iteration count: chunks
    of 50 iterations until the count or the time budget is reached; any
    remainder is projected at the measured per-iteration rate."""
    import os

    from nodals_amd._lib import H8
    from oracle import cpubind
    from oracle import h8_oracle as O

    h = H8(solver_args(mesh, "fp64"), device=False)
    case = O.case_from_product(h)
    h.close()
    cs = cpubind.CpuSimple(case)  # gather lists, P^T: setup, not timed
    usable = cpubind.usable_cores()
    threads = min(usable, int(os.environ.get("OMP_NUM_THREADS") or usable))
    U, p = None, None
    done, dt = 0, 0.0
    while done < gpu_outer and dt < budget_s:
        n = min(50, gpu_outer - done)
        t0 = time.perf_counter()
        r = cs.simple(n, threads=threads, U0=U, p0=p)
        dt += time.perf_counter() - t0
        U, p, done = r["U"], r["p"], done + n
    cs.close()
    per_it = dt / done
    total = per_it * gpu_outer
    dp_cpu = O.pressure_drop_fit(case, p)
    full = done == gpu_outer
    tail = (f"all {gpu_outer} iterations timed; its pressure drop {dp_cpu:.6f} vs the GPU FP64 run's {dp_gpu:.6f}"
            if full else f"{done} of the GPU's {gpu_outer} iterations timed, the rest projected at that rate")
    print(f"[bench] CPU port: {threads} threads, {done} its in {dt:.1f} s ({per_it * 1e3:.1f} ms/it)",
          file=sys.stderr, flush=True)
    return {"value": total, "unit": "s", "cores": threads, "kind": "port",
            "cpu_model": cpubind.cpu_model(), "usable_cores": usable, "ms_per_iteration": per_it * 1e3,
            "iterations_timed": done, "projected": not full,
            "pressure_drop": dp_cpu if full else None,
            "sample": (f"FP64 C++/OpenMP port of the same SIMPLE loop (FGS momentum, FGMRES(30) + SA V-cycle with "
                       f"damped-Jacobi smoothing, p-rtol 0.5) on the same 40,320-tet pipe from rest, {threads} "
                       f"threads of {cpubind.cpu_model()}; {tail}")}


def poisson_leg(n: int, reps: int) -> dict:
    """Pressure-Poisson kernel bench: 64^3 Kuhn P1 Laplacian, SA V-cycle with 2
    forward MCGS sweeps in FGMRES(30) to rtol 1e-8, hierarchy in FP64 / FP32 /
    BF16-finest storage, FP64 outer residual; SpMV GB/s per level."""
    from nodals_amd import p1
    from nodals_amd._lib import System

    row, col, val = p1.kuhn_p1_laplacian(n)
    b = p1.poisson_rhs(row.size - 1)
    out = {"rows": int(row.size - 1), "nnz": int(row[-1])}
    for prec in ("fp64", "fp32", "bf16"):
        S = System(row, col, val, f"precision={prec} smoother=mcgs sweeps=2 order=forward")
        S.set_rhs(b)
        ms, info = S.solve_timed("fgmres", 1e-8, 30, 500, reps)
        levels = []
        for lvl in range(-1, S.levels - 1):
            t, by = S.time_spmv(lvl, 100)
            levels.append({"level": lvl, "rows": S.n if lvl < 0 else S.level_rows[lvl],
                           "ms": t, "GB_s": by / (t * 1e-3) / 1e9})
        out[prec] = {"solve_ms": ms, "fgmres_its": info["its"], "true_rel_residual": info["rel"], "spmv": levels}
        print(f"[bench] poisson {prec}: {ms:.2f} ms, its={info['its']} rel={info['rel']:.2e} "
              f"fine SpMV {levels[0]['GB_s']:.0f} GB/s", file=sys.stderr, flush=True)
        S.close()
    return out


def reset(h) -> None:
    """Start state of a step: fluid at rest, zero pressure (set before the timer)."""
    import numpy as np
    h.set("p", np.zeros(h.info["cells"]))
    for d in range(3):
        h.set(f"u{d}", np.zeros(h.info["nv"]))


def traffic_per_launch() -> dict | None:
    """HBM bytes per momentum Gauss-Seidel colour launch from the committed
    rocprofv3 counter summary: FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
    each scaled by what the same counter reads for a streaming copy of known
    size."""
    f = ROOT / "profiles" / "mom_gs_pipe40k_pmc.txt"
    if not f.exists():
        return None
    vals = {}
    for line in f.read_text().splitlines():
        if line.startswith("traffic_"):
            k, v = line.split("=", 1)
            vals[k.strip()] = float(v)
    if "traffic_bytes_per_launch" not in vals:
        return None
    return vals


# Published H8 numbers of the reference (RTX 3060, gpu/serial_cuda/README.md):
# fixed-10 FP64 at 768,530 cells 100.373 ms per SIMPLE = 7.657 MIUPS
# (:156-158); Re=20 convergence gate at 111,183 cells to simpleTol 1e-3:
# 407 SIMPLE iterations, 13.006 ms per SIMPLE = 8.549 MIUPS (:144-152).
REF_FIXED10_FP64_MIUPS = 7.657
REF_GATE_MS_PER_SIMPLE, REF_GATE_ITS = 13.006, 407


def reference_legs() -> dict:
    """The reference's own H8 benchmarks with its binary defaults (PCG, p-rtol
    0.5, p-max 20, Chebyshev(2) smoother with power-iteration spectrum, SA
    hierarchy, fgs1), on generated O-grid pipes of the same cell counts:
    fixed-10 FP64 at 768,768 cells (16x16 core, 10 rings, 143 slabs) and the
    Re = 20 convergence gate to simpleTol 1e-3 at 112,200 cells (10x10 core,
    6 rings, 55 slabs).  MIUPS = cells / (ms per SIMPLE) / 1000."""
    from nodals_amd import meshgen
    from nodals_amd._lib import H8

    out = {}
    d = ROOT / "_meshes" / "pipe_896x143"
    if not (d / "faces").exists():
        meshgen.write_pipe_case(d, core=16, ring=10, nslab=143)
    for prec in ("fp64", "fp32"):
        h = H8(["--mesh", str(d), "--run-mode", "fixed10", "--max-outer", "10", "--amg-precision", prec,
                "--verbose", "0"])
        h.run()  # warm-up
        reset(h)
        r = h.run()
        ms = r["loop_ms"] / 10.0
        out[f"fixed10_768k_{prec}"] = {"cells": h.info["cells"], "ms_per_simple": ms,
                                       "miups": h.info["cells"] / ms / 1000.0,
                                       "avg_pressure_its": r["avg_pressure_its"],
                                       "stage_ms": {k: r[f"{k}_ms"] / 10.0 for k in
                                                    ("assembly", "momentum", "continuity", "pressure", "pupdate")}}
        print(f"[bench] fixed-10 {h.info['cells']} cells {prec}: {ms:.3f} ms/SIMPLE "
              f"({out[f'fixed10_768k_{prec}']['miups']:.2f} MIUPS)", file=sys.stderr, flush=True)
        h.close()
    out["fixed10_768k_fp64"]["miups_vs_reference_rtx3060"] = out["fixed10_768k_fp64"]["miups"] / REF_FIXED10_FP64_MIUPS
    d = ROOT / "_meshes" / "pipe_340x55"
    if not (d / "faces").exists():
        meshgen.write_pipe_case(d, core=10, ring=6, nslab=55)
    h = H8(["--mesh", str(d), "--run-mode", "converge", "--max-outer", "5000", "--simple-tol", "1e-3",
            "--verbose", "0"])
    r = h.run()
    ms = r["loop_ms"] / max(r["outer"], 1)
    out["gate_112k_fp64"] = {"cells": h.info["cells"], "outer": r["outer"], "converged": r["converged"],
                             "ms_per_simple": ms, "loop_s": r["loop_ms"] / 1e3,
                             "pressure_drop": r["pressure_drop"], "exact_pressure_drop": r["exact_pressure_drop"],
                             "reference_outer": REF_GATE_ITS, "reference_ms_per_simple": REF_GATE_MS_PER_SIMPLE,
                             "speedup_per_simple_vs_reference_rtx3060": REF_GATE_MS_PER_SIMPLE / ms}
    print(f"[bench] gate {h.info['cells']} cells fp64: {r['outer']} SIMPLE, {ms:.3f} ms/SIMPLE", file=sys.stderr,
          flush=True)
    h.close()
    return out


def cavity_leg(steps: int) -> dict:
    """Lid-driven cavity (unit cube, 16^3 hexes x 6 Kuhn tets = 24,576 tets,
    lid u = (1, 0, 0), Re = 100), plain FP64, SIMPLE to 1e-6."""
    from nodals_amd import meshgen
    from nodals_amd._lib import H8

    d = ROOT / "_meshes" / "cavity_16"
    if not (d / "faces").exists():
        meshgen.write_cavity_case(d, n=16)
    h = H8(["--mesh", str(d), "--lid", "lid", "--wall", "walls", "--re", "100", "--run-mode", "converge",
            "--max-outer", "20000", "--simple-tol", "1e-6", "--alpha-u", "0.5", "--alpha-p", "0.5",
            "--amg-smoother", "jacobi", "--amg-jacobi-omega", "0.7", "--pressure-solver", "fgmres", "--p-rtol", "0.1",
            "--p-max", "200", "--amg-precision", "fp64", "--verbose", "0"])
    times, r = [], None
    for _ in range(max(1, steps)):
        reset(h)
        t0 = time.perf_counter()
        r = h.run()
        times.append(time.perf_counter() - t0)
    print(f"[bench] cavity 16^3 fp64: {np.mean(times):.3f} s, outer={r['outer']} converged={r['converged']}",
          file=sys.stderr, flush=True)
    out = {"s": float(np.mean(times)), "outer": r["outer"], "converged": r["converged"], "cells": h.info["cells"],
           "avg_pressure_its": r["avg_pressure_its"]}
    h.close()
    return out


def turb_leg(h, steps: int) -> dict:
    """k-omega SST transport step (k then omega: convection, volume terms,
    inlet flux, omega wall penalty at the third of five wall-normal samples,
    row-L1 relaxation, 20 + 20 coloured Gauss-Seidel sweeps, floors) on the
    converged pipe velocity of context h, with the reference's spiral-case
    settings (alpha 0.3, sweep omega 0.7, penalty 1e6)."""
    def run(args):
        t0 = time.perf_counter()
        h.turb_init(args)
        setup_s = time.perf_counter() - t0
        for _ in range(2):
            h.turb_step()
        dev, wall = [], []
        for _ in range(max(1, steps)):
            ts = time.perf_counter()
            r = h.turb_step()
            wall.append(time.perf_counter() - ts)
            dev.append(r["ms"])
        ns = h.turb["ns"]
        return setup_s, dev, wall, h.get("k", ns), h.get("omega", ns)

    setup_s, dev, wall, k64, w64 = run([])
    out = {"ms_per_step_device": float(np.mean(dev)), "ms_per_step_wall": 1e3 * float(np.mean(wall)),
           "steps": len(dev), "setup_s": setup_s, "scalar_dofs": h.turb["ns"], "vertex_rows": h.turb["nvert"],
           "colors": h.turb["colors"], "wall_faces": h.turb["wall_faces"], "inlet_faces": h.turb["inlet_faces"],
           "sweeps_per_equation": 20}
    # the same steps with the sweep operator's values stored in FP32
    _, dev32, _, k32, w32 = run(["--sst-sweep-precision", "fp32"])
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))  # noqa: E731
    out["fp32_sweep_values"] = {"ms_per_step_device": float(np.mean(dev32)), "k_rel_l2_vs_fp64": rel(k32, k64),
                                "omega_rel_l2_vs_fp64": rel(w32, w64), "steps_compared": 2 + len(dev32)}
    print(f"[bench] k-omega step: {out['ms_per_step_device']:.3f} ms device, {out['ms_per_step_wall']:.3f} ms wall "
          f"({h.turb['colors']} colours, {h.turb['wall_faces']} wall faces); FP32 sweep values "
          f"{np.mean(dev32):.3f} ms, k/omega vs FP64 {rel(k32, k64):.1e}/{rel(w32, w64):.1e}",
          file=sys.stderr, flush=True)
    return out


# The SST case's solver settings (examples/spiral_swept_207k/
# SPIRAL_SWEPT_ADAPTIVE.case): 20 momentum sweeps at omega 0.7 ([solver.u]
# m3-fgs), 20 + 20 k-omega sweeps ([sst]), pressure to rtol 1e-3
# ([solver.p]), alpha_u 0.7 / alpha_p 0.3 ([simple]).
SST_SOLVER = ["--mom-sweeps", "20", "--mom-omega", "0.7", "--p-rtol", "1e-3", "--alpha-u", "0.7", "--alpha-p", "0.3"]


def sst_coupled_leg(mesh: str) -> dict:
    """Coupled SST SIMPLE on the same 40,320-tet pipe at Re = 200: eddy
    viscosity in the momentum and one k-omega step after every pressure
    update, from rest (k, omega at their inlet values) to residual 1e-5,
    with the SST case's solver settings, FP32 AMG in FGMRES."""
    from nodals_amd._lib import H8

    args = ["--mesh", mesh, "--run-mode", "converge", "--max-outer", "5000", "--simple-tol", "1e-5",
            "--amg-smoother", "jacobi", "--amg-jacobi-omega", "0.7", "--pressure-solver", "fgmres", "--p-max", "200",
            "--amg-precision", "fp32", "--verbose", "0", "--re", "200"] + SST_SOLVER
    h = H8(args)
    h.turb_init([])
    h.turb_couple(True)
    t0 = time.perf_counter()
    r = h.run()
    dt = time.perf_counter() - t0
    out = {"s": dt, "outer": r["outer"], "converged": r["converged"], "pressure_its": r["pressure_its"],
           "turb_ms": r["turb_ms"], "pressure_drop": r["pressure_drop"], "re": 200}
    print(f"[bench] coupled SST pipe Re=200: {dt:.3f} s, outer={r['outer']} converged={r['converged']} "
          f"(k-omega {r['turb_ms'] / 1e3:.3f} s)", file=sys.stderr, flush=True)
    h.close()
    return out


def helix_leg() -> dict:
    """Config 5's geometry: the 207,360-tet helix (576-quad section x 60 slabs,
    coil radius 5D, pitch 2D, two turns) in duct mode, FP32 AMG with
    multicolour Gauss-Seidel smoothing (the spiral case's smoother; the
    section cells are about 20 times longer than wide, which Jacobi-smoothed
    aggregation does not handle).  Laminar Re = 20 to 1e-6.  (The SST runs on
    this mesh are config5_leg's.)"""
    from nodals_amd import meshgen
    from nodals_amd._lib import H8

    d = ROOT / "_meshes" / "helix_207k"
    if not (d / "faces").exists():
        meshgen.write_helix_case(d)
    base = ["--mesh", str(d), "--duct", "1", "--run-mode", "converge", "--max-outer", "6000", "--amg-smoother", "mcgs",
            "--pressure-solver", "fgmres", "--p-max", "400", "--amg-precision", "fp32", "--verbose", "0"]
    out = {}
    for key, extra, sst in (("laminar_re20", ["--re", "20", "--simple-tol", "1e-6", "--alpha-u", "0.5", "--alpha-p", "0.5",
                                              "--p-rtol", "0.5"], False),):
        h = H8(base + extra)
        if sst:
            h.turb_init([])
            h.turb_couple(True)
        t0 = time.perf_counter()
        r = h.run()
        dt = time.perf_counter() - t0
        out[key] = {"s": dt, "outer": r["outer"], "converged": r["converged"], "pressure_its": r["pressure_its"],
                    "ms_per_simple": 1e3 * dt / max(r["outer"], 1), "turb_s": r["turb_ms"] / 1e3,
                    "pressure_drop": r["pressure_drop"], "cells": h.info["cells"]}
        print(f"[bench] helix 207k {key}: {dt:.2f} s, outer={r['outer']} converged={r['converged']} "
              f"pressure its={r['pressure_its']}", file=sys.stderr, flush=True)
        h.close()
    return out


# Config 5 with the spiral case's wall closure and solver settings
# (examples/spiral_swept_207k/SPIRAL_SWEPT_ADAPTIVE.case): Spalding wall
# functions with a one-point face-averaged fit and the Nitsche normal
# penalty gamma 1e-6 ([wall_function]), deviatoric stress ([stress]), 20 M3
# block sweeps at omega 0.7 ([solver.u]), 20 + 20 k-omega sweeps at alpha
# 0.3 ([sst]), PCG-AMG pressure to rtol 1e-3 on a CF/PMIS extended+i
# hierarchy with symmetric MCGS ([solver.p], [amg], [native]), alpha_p 0.3
# ([simple]), rAU scale 2 (the SST application's default), potential-flow
# start ([initialization]).  Two settings differ from that case, both of
# them case keys of the reference: advective convection ([convection] form)
# and alpha_u 0.5 ([simple] ur_u).  At Re 50,000 on this O-grid the case's
# skew-symmetric form drives wall-bubble rows to non-positive relaxed
# diagonals (which the reference's s22c_finalize_scalar rejects) or wall
# fits to failure within tens of iterations, at alpha_u 0.7 and 0.5 alike;
# the advective form runs, and alpha_u 0.5 damps its plateau the most.
def config5_args(mesh: str, precision: str, max_outer: int, tol: float) -> list:
    return ["--mesh", mesh, "--duct", "1", "--wall-model", "spalding", "--re", "50000", "--run-mode", "converge",
            "--max-outer", str(max_outer), "--simple-tol", str(tol), "--alpha-u", "0.5", "--alpha-p", "0.3",
            "--amg-hierarchy", "cf", "--cf-interp", "exti", "--cf-theta", "0.25", "--amg-terminal", "1000",
            "--amg-smoother", "mcgs", "--amg-mcgs-order", "symmetric", "--pressure-solver", "pcg",
            "--p-rtol", "1e-3", "--p-atol", "1e-12", "--p-max", "600", "--amg-precision", precision,
            "--verbose", "0", "--tag", f"config5_{precision}"]


CONFIG5_WALL = ["--convection-form", "advective"]
# the spiral case's wall sampling: one Spalding point (the third), ray traced
# with local shortening ([wall_sampling] mode = ray_trace, samples = 1)
CONFIG5_SAMPLING = ["--sst-wall-sampling", "ray_trace", "--sst-wall-active-points", "1",
                    "--sst-wall-ray-local-adapt", "1",
                    # [omega_wall] stage41_mode = trace_log: the log-layer omega
                    # target imposed on the wall faces by symmetric Nitsche
                    "--sst-omega-wall-target", "trace_log"]


def yplus30_sample_scale(re: float) -> float:
    """Sample-height scale that puts the Spalding sample (the third of the
    five reference heights d_i D / 0.004 * scale) at y+ = 30 for the Blasius
    friction velocity u_tau = U sqrt(0.3164 Re^-1/4 / 8): the starting point
    the reference's adaptive runner refines (run_adaptive_case.py:56-204)."""
    ut = (0.3164 * re ** -0.25 / 8.0) ** 0.5
    y = 30.0 / (re * ut)                     # D = U = 1
    return y / (4.4084029504620439e-4 / 0.004)


def config5_context(mesh: str, precision: str, max_outer: int, tol: float):
    from nodals_amd._lib import H8

    h = H8(config5_args(mesh, precision, max_outer, tol))
    h.turb_init(["--sst-wall-sample-scale", repr(yplus30_sample_scale(50000.0)), *CONFIG5_SAMPLING])
    h.wall_init(CONFIG5_WALL)
    return h


def config5_reset(h) -> None:
    """Start state: the potential-flow start of the wall path (computed once
    at wall init), zero pressure, k and omega at their inlet values on the
    vertices."""
    nv, nc = h.info["nv"], h.info["cells"]
    u0 = h.get("wall_u0", 3 * nv).reshape(nv, 3)
    for d in range(3):
        h.set(f"u{d}", np.ascontiguousarray(u0[:, d]))
    h.set("p", np.zeros(nc))
    ns, nvert = h.turb["ns"], h.turb["nvert"]
    for name, v in (("k", h.turb["k_in"]), ("omega", h.turb["omega_in"])):
        f = np.zeros(ns)
        f[:nvert] = v
        h.set(name, f)


def config5_leg(iters: int) -> dict:
    """Config 5: k-omega SST helical pipe (D = 1, coil radius 5D, pitch 2D,
    two turns; 576-quad O-grid section x 60 slabs, 6 tets per hex: 207,360
    tets), Re = 50,000, uniform inlet with 5% intensity and omega from a
    0.07 D length scale, Spalding wall functions with the sample at y+ 30,
    from the potential-flow start, a fixed budget of `iters` SIMPLE
    iterations (the loop does not reach relative changes of 1e-5 within
    thousands of iterations); plain FP64 and FP32 AMG inside the FP64 PCG.
    Reported per form: wall-clock, ms per SIMPLE iteration and its stages,
    pressure iterations, the continuity residual reached, pressure drop,
    mean wall shear, y+; for FP32 the relative L2 difference of every field
    from FP64 and of pressure drop and wall shear."""
    from nodals_amd import meshgen

    d = ROOT / "_meshes" / "helix_207k"
    if not (d / "faces").exists():
        meshgen.write_helix_case(d)
    out = {"iterations": iters, "wall_args": CONFIG5_WALL}
    fields = {}
    for prec in ("fp64", "fp32"):
        h = config5_context(str(d), prec, iters, 1e-5)
        try:
            config5_reset(h)
            t0 = time.perf_counter()
            r = h.run()
            dt = time.perf_counter() - t0
            print(f"[bench] config5 {prec}: {dt:.2f} s, outer={r['outer']} converged={r['converged']} "
                  f"contRel={r['final_rel_cont']:.3e} avgP={r['avg_pressure_its']:.2f} dp={r['pressure_drop']:.6e}",
                  file=sys.stderr, flush=True)
            nf = h.turb["wall_faces"]
            fit = h.get("wall_fit", 3 * nf).reshape(nf, 3)
            shear = h.get("wall_shear", 4 * nf).reshape(nf, 4)
            area = h.get("wall_area", nf)
            out[prec] = {"s": dt, "ms_per_simple": 1e3 * dt / max(r["outer"], 1), "outer": r["outer"],
                         "converged": r["converged"], "final_rel_cont": r["final_rel_cont"],
                         "pressure_its": r["pressure_its"], "avg_pressure_its": r["avg_pressure_its"],
                         "pressure_drop": r["pressure_drop"],
                         "wall_shear_mean": float(shear[:, 3].sum() / area.sum()),
                         "wall_force": [float(x) for x in shear[:, :3].sum(0)],
                         "yplus_mean": float((fit[:, 2] * area).sum() / area.sum()),
                         "utau_mean": float((fit[:, 1] * area).sum() / area.sum()),
                         "stage_ms_per_simple": {k: r[f"{k}_ms"] / max(r["outer"], 1)
                                                 for k in ("assembly", "momentum", "continuity", "pressure", "turb")},
                         "cells": h.info["cells"], "free_velocity_dofs": h.info["nv"], "wall_faces": nf,
                         "potential_start": [float(x) for x in h.get("wall_start", 7)]}
            cur = {f"u{c}": h.get(f"u{c}", h.info["nv"]) for c in range(3)}
            cur["p"] = h.get("p", h.info["cells"])
            cur["k"] = h.get("k", h.turb["ns"])
            cur["omega"] = h.get("omega", h.turb["ns"])
            if prec == "fp64":
                fields = cur
            else:
                out[prec]["rel_l2_vs_fp64"] = {
                    k: float(np.linalg.norm(cur[k] - ref) / max(np.linalg.norm(ref), 1e-300))
                    for k, ref in fields.items()}
        finally:
            h.close()
    for k in ("pressure_drop", "wall_shear_mean"):
        out["fp32"][f"{k}_rel_diff_vs_fp64"] = abs(out["fp32"][k] - out["fp64"][k]) / max(abs(out["fp64"][k]), 1e-300)
    out["speedup_fp32_amg_vs_fp64"] = out["fp64"]["s"] / out["fp32"]["s"]
    return out


def learned_leg(mesh: str, steps: int, epochs: int) -> dict:
    """Learned warm start + FP32 AMG on the same pipe: samples recorded from
    the solver's own converged runs at Re 10, 15, 30, 45 (every 4th SIMPLE
    iteration), trained on the GPU, then the timed Re = 20 solves start each
    pressure solve from the network's prediction (its residual-optimal
    multiple).  Recording and training are reported, not timed."""
    from nodals_amd._lib import Gnn, H8

    h = H8(solver_args(mesh, "fp32"))
    g = Gnn(h, seed=1)
    g.reserve(3000)
    t0 = time.perf_counter()
    for re in (10.0, 15.0, 30.0, 45.0):
        hr = H8(solver_args(mesh, "fp32") + ["--re", str(re)])
        hr.attach_gnn(g, warm_start=False, record_every=4)
        hr.run()
        hr.attach_gnn(None)
        hr.close()
    t1 = time.perf_counter()
    l0 = g.evaluate()
    last = g.train(epochs=epochs, lr=1e-3, seed=2)
    t2 = time.perf_counter()
    h.attach_gnn(g, warm_start="guess")
    times, runs = [], []
    for _ in range(max(1, steps)):
        reset(h)
        ts = time.perf_counter()
        r = h.run()
        times.append(time.perf_counter() - ts)
        runs.append(r)
        print(f"[bench] learned+fp32 step: {times[-1]:.3f} s, outer={r['outer']} pressure its={r['pressure_its']} "
              f"guesses dropped={r['fallbacks']}/{r['warm_starts']}", file=sys.stderr, flush=True)
    h.attach_gnn(None)
    out = {"s": float(np.mean(times)), "outer": runs[-1]["outer"], "pressure_its": runs[-1]["pressure_its"],
           "guesses": runs[-1]["warm_starts"], "guesses_dropped": runs[-1]["fallbacks"],
           "pressure_drop": runs[-1]["pressure_drop"], "samples": g.samples, "record_s": t1 - t0,
           "train_s": t2 - t1, "epochs": epochs, "loss_before": l0, "loss_after": last}
    g.close()
    h.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="time budget of the CPU baseline leg; iterations beyond it are projected")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=500)
    ap.add_argument("--no-poisson", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-cavity", action="store_true")
    ap.add_argument("--no-reference-legs", action="store_true")
    ap.add_argument("--no-turb", action="store_true")
    ap.add_argument("--no-sst-coupled", action="store_true")
    ap.add_argument("--no-helix", action="store_true")
    ap.add_argument("--no-config5", action="store_true")
    ap.add_argument("--config5-iters", type=int, default=120)
    ap.add_argument("--learned-epochs", type=int, default=20)
    ap.add_argument("--kernel-profile", action="store_true",
                    help="only the fine SpMV and the copy calibration kernel (for counter runs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    from nodals_amd import meshgen
    from nodals_amd._lib import H8

    mesh_root = ROOT / "_meshes"
    mesh = mesh_root / "pipe_224x30"
    if not (mesh / "faces").exists():
        meshgen.write_pipe_case(mesh, core=8, ring=5, nslab=30)
    mesh = str(mesh)

    if a.kernel_profile:
        # Only the profiled kernels run: each hot kernel of the config-2 step
        # 20 (+1 warm) times, and the calibration copy 5 (+1) times.
        h = H8(solver_args(mesh, "fp32"))
        h.assemble(0.5)
        h.pressure_refresh()
        out = {"colours": int(h.info["colors"])}
        for k in ("fine_spmv", "mom_gs", "refresh", "assemble"):
            out[f"{k}_ms"], out[f"{k}_bytes"] = h.time_kernel(k, 20)
        out["copy_ms"], out["copy_bytes"] = h.time_kernel("copy", 5)
        print(json.dumps(out))
        h.close()
        return 0

    results = {}
    # One solver context alive at a time.  A context runs the fine Schur
    # refresh on a second stream beside the momentum solve; streams share a
    # small pool of hardware queues, and with another context alive the two
    # streams of the new one can share a queue and lose the overlap.
    for prec in ("fp64", "fp32"):
        h = H8(solver_args(mesh, prec))
        try:
            for _ in range(a.warmup):
                reset(h)
                h.run()
            times, runs = [], []
            for _ in range(a.steps):
                reset(h)
                t0 = time.perf_counter()
                r = h.run()
                times.append(time.perf_counter() - t0)
                runs.append(r)
                print(f"[bench] {prec} step: {times[-1]:.3f} s, outer={r['outer']} converged={r['converged']} "
                      f"avgP={r['avg_pressure_its']:.2f} dp={r['pressure_drop']:.6f}", file=sys.stderr, flush=True)
            results[prec] = {"times": times, "runs": runs}
            if prec != "fp32":
                continue
            info = dict(h.info)
            # Roofline kernel: the FP64 fine Schur SpMV (the Krylov operator).
            # Its launches inside one more full solve of the same workload are
            # bracketed by HIP events on the solver stream (kept out of the
            # timed steps, whose wall clock the extra events would perturb).
            reset(h)
            h.probe_start()
            h.run()
            launches, probe_ms = h.probe_stop()
            ms = probe_ms / max(launches, 1)
            _, by = h.time_kernel("fine_spmv", 1)
            spmv_gbs = by / (ms * 1e-3) / 1e9
            spmv = {"kernel": "k_schur_spmv (fine P0 Schur SpMV, FP64)", "GB_s": spmv_gbs, "frac": spmv_gbs / HBM_PEAK_GBS,
                    "kernel_ms": ms, "launches": launches, "bytes_per_launch": by}
            # Roofline kernel: the momentum Gauss-Seidel colour pass (k_mom_gs,
            # the workload's largest kernel by GPU time).  Its launches inside
            # one more full solve are bracketed by HIP events from their own
            # dispatch packets; a sweep's algorithmic bytes are CSR values and
            # columns, row pointers, diagonal, right-hand side and the
            # velocity read and written once.
            reset(h)
            h.probe_start(2)
            h.run()
            gs_launches, gs_ms = h.probe_stop()
            _, sweep_bytes = h.time_kernel("mom_gs", 1)
            colours = int(info["colors"])
            sweeps = gs_launches / max(colours, 1)
            by = sweep_bytes / max(colours, 1)
            ms = gs_ms / max(gs_launches, 1)
            launches = gs_launches
            achieved = sweep_bytes * sweeps / (gs_ms * 1e-3) / 1e9
            warm_ms, _ = h.time_kernel("fine_spmv", a.spmv_reps)
            spmv["warm_loop_kernel_ms"] = warm_ms
            extra = {}
            for k in ("refresh", "mom_gs", "assemble", "continuity"):
                kms, kby = h.time_kernel(k, 100)
                extra[k] = {"ms": kms, "GB_s": kby / (kms * 1e-3) / 1e9}
            turb = None if a.no_turb else turb_leg(h, 20)
        finally:
            h.close()

    fp32 = results["fp32"]
    fp64 = results["fp64"]
    value = float(np.mean(fp32["times"]))
    outer = int(fp32["runs"][-1]["outer"])
    dp32 = fp32["runs"][-1]["pressure_drop"]
    dp64 = fp64["runs"][-1]["pressure_drop"]
    coupled = None if a.no_sst_coupled else sst_coupled_leg(mesh)
    helix = None if a.no_helix else helix_leg()
    config5 = None if a.no_config5 else config5_leg(a.config5_iters)
    poisson = None if a.no_poisson else poisson_leg(64, max(1, a.steps))
    learned = None if a.no_learned else learned_leg(mesh, a.steps, a.learned_epochs)
    cavity = None if a.no_cavity else cavity_leg(1)
    refl = None if a.no_reference_legs else reference_legs()
    if learned:
        learned["pressure_its_fp32_amg"] = int(fp32["runs"][-1]["pressure_its"])
        learned["speedup_vs_fp64"] = float(np.mean(fp64["times"])) / learned["s"]
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(mesh, int(fp64["runs"][-1]["outer"]), a.cpu_seconds, dp64)
    traffic = traffic_per_launch()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": value * 1e3,
        "ms_per_simple_iteration": value * 1e3 / max(outer, 1),
        "higher_is_better": False,
        "scaling": "none (one GPU)",
        "vs_baseline": None,
        "dtype": "f64 state / f32 AMG",
        "data": "synthetic (generated O-grid pipe mesh)",
        "config": {
            "workload": ("laminar straight pipe D=1 L=10D Re=20, O-grid 224 quads x 30 slabs, 40,320 tets; "
                         "SIMPLE from rest to residual 1e-6 (FGS momentum, SA-AMG FGMRES(30) pressure at the reference's "
                         "default p-rtol 0.5)"),
            "cells": info["cells"], "free_velocity_dofs": info["nv"],
            "fp32_amg_s": value, "fp64_s": float(np.mean(fp64["times"])),
            "speedup_fp32_amg_vs_fp64": float(np.mean(fp64["times"])) / value,
            "outer_iterations_fp32_amg": outer, "outer_iterations_fp64": int(fp64["runs"][-1]["outer"]),
            "avg_pressure_its_fp32_amg": fp32["runs"][-1]["avg_pressure_its"],
            "pressure_drop_fp64": dp64, "pressure_drop_fp32_amg": dp32,
            "pressure_drop_rel_diff": abs(dp32 - dp64) / abs(dp64),
            "hagen_poiseuille_drop": fp64["runs"][-1]["exact_pressure_drop"],
            "setup_ms": fp32["runs"][-1]["setup_ms"],
            "stage_ms_per_simple_fp32_amg": {
                k: fp32["runs"][-1][f"{k}_ms"] / max(outer, 1)
                for k in ("assembly", "momentum", "continuity", "pressure", "pupdate")},
            "kernels": extra,
            "poisson_64": poisson,
            "learned_fp32_amg": learned,
            "cavity_16_fp64": cavity,
            "k_omega_step_40k": turb,
            "sst_coupled_pipe_40k_re200": coupled,
            "helix_207k_laminar_fp32_amg": helix,
            "config5_wall_sst_207k_re50000": config5,
            "reference_published_benchmarks": refl,
            "parallelism": "single GPU",
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic["traffic_bytes_per_launch"],
                     "kernel": "k_mom_gs<32, 16> (momentum Gauss-Seidel colour pass, FP64, fgs1)",
                     "kernel_ms": ms, "launches": launches, "bytes_per_launch": by,
                     "colours": colours, "sweeps": sweeps,
                     "traffic_calibration": None if traffic is None else
                     {k: v for k, v in traffic.items() if k != "traffic_bytes_per_launch"},
                     "fine_spmv": spmv},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
