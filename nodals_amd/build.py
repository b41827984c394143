"""Build recipe for libnodals_hip.so and the nodals_gpu_h8_* executables.

Compiles every HIP source for gfx950 with hipcc, the host setup sources
with -ffp-contract=off (their floating-point results must be reproducible,
see csrc/host/h8_setup.hpp), and links in-tree:

    nodals_amd/libnodals_hip.so
    nodals_amd/bin/nodals_gpu_h8_fp64   (+ _fp32, _bf16 copies)
    nodals_amd/bin/nodals_sst           (the SST case launcher's native solver)

Usage: ``python -m nodals_amd.build`` or ``nodals_amd.build.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libnodals_hip.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{ROOT / 'include'}"]
HOST_FLAGS = COMMON + ["-ffp-contract=off"]
# Packed-FP32 instructions are disabled for all device code.
DEVICE_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                         "-munsafe-fp-atomics"]

HOST_SOURCES = ["host/mesh.cpp", "host/h8_setup.cpp", "host/amg_setup.cpp", "host/turb_setup.cpp", "host/ray_trace.cpp",
                "host/wall_setup.cpp", "host/potential_start.cpp"]
DEVICE_SOURCES = ["device/h8_solver.hip", "device/sys_solver.hip", "device/sst_kernels.hip", "device/gnn.hip", "device/turb_step.hip",
                   "device/wall_solver.hip", "device/hex_kernels.hip", "device/capi.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def _compile(src: str, verbose: bool) -> Path:
    s = CSRC / src
    o = OBJ / (src.replace("/", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + DEVICE_FLAGS + ["-x", "hip", "-c", str(s), "-o", str(o)]
    else:
        cmd = [HIPCC] + HOST_FLAGS + ["-c", str(s), "-o", str(o)]
    out = _run(cmd, verbose)
    if verbose and out.strip():
        print(out)
    return o


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile everything (always) and return the library path."""
    OBJ.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    srcs = HOST_SOURCES + DEVICE_SOURCES
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs], verbose)
    os.replace(tmp, LIB)
    main_o = OBJ / "h8_main.o"
    _run([HIPCC] + HOST_FLAGS + ["-c", str(CSRC / "h8_main.cpp"), "-o", str(main_o)], verbose)
    exe = BIN / "nodals_gpu_h8_fp64"
    _run([HIPCC, "-o", str(exe), str(main_o), f"-L{PKG}", "-lnodals_hip", "-Wl,-rpath,$ORIGIN/.."], verbose)
    for suffix in ("fp32", "bf16"):
        shutil.copy2(exe, BIN / f"nodals_gpu_h8_{suffix}")
    sst_o = OBJ / "sst_main.o"
    _run([HIPCC] + HOST_FLAGS + ["-c", str(CSRC / "sst_main.cpp"), "-o", str(sst_o)], verbose)
    _run([HIPCC, "-o", str(BIN / "nodals_sst"), str(sst_o), f"-L{PKG}", "-lnodals_hip", "-Wl,-rpath,$ORIGIN/.."],
         verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
    print(f"built {LIB}")
