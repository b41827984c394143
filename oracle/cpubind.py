"""ctypes binding of oracle/_cpu/libnodals_cpu.so, the FP64 C++/OpenMP
restatement of the H8 SIMPLE loop (oracle/cpu_h8.cpp).

TEST INFRASTRUCTURE ONLY: loaded by tests/ (checked against the numpy
oracle's ``simple``) and by bench.py's CPU baseline leg; never by the
product package.
"""
from __future__ import annotations

import ctypes as C
import os
import platform
from pathlib import Path

import numpy as np

from . import h8_oracle as O

CPU_LIB = Path(__file__).resolve().parent / "_cpu" / "libnodals_cpu.so"

_P = C.c_void_p


class _Case(C.Structure):
    _fields_ = [("nc", C.c_int32), ("nv", C.c_int32), ("ncolors", C.c_int32), ("pin", C.c_int32)] + [
        (n, _P) for n in ("cell_ref", "cell_slot", "cell_det", "cell_invJ", "bcoef", "fixed_value", "mom_row",
                          "mom_col", "mom_diff", "color", "static_rhs", "fixed_div", "tcent", "f_row", "f_col",
                          "f_diag", "inc_off", "inc_c", "inc_a", "contrib_off", "slot")]


class _Level(C.Structure):
    _fields_ = [("nfine", C.c_int32), ("ncoarse", C.c_int32)] + [
        (n, _P) for n in ("p_row", "p_col", "p_val", "a_row", "a_col", "a_val")]


class _Hier(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("nterm", C.c_int32), ("levels", C.POINTER(_Level)), ("tinv", _P)]


def available() -> bool:
    return CPU_LIB.exists()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def usable_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


class CpuSimple:
    """Holds the case arrays (kept alive for the library) and runs SIMPLE."""

    def __init__(self, case: O.H8Case, jacobi_omega: float = 0.7):
        if not available():
            raise RuntimeError(f"{CPU_LIB} is not built (make -C oracle cpu)")
        L = C.CDLL(str(CPU_LIB))
        L.cpu_h8_open.restype = _P
        L.cpu_h8_open.argtypes = [C.POINTER(_Case), C.POINTER(_Hier), C.c_double]
        L.cpu_h8_close.argtypes = [_P]
        L.cpu_h8_simple.restype = C.c_int
        L.cpu_h8_simple.argtypes = [_P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.c_int, _P, _P, _P]
        L.cpu_h8_threads.restype = C.c_int
        self.L = L
        self.case = case
        T = O.fine_topology(case)
        _, tcent = O.tensors()
        keep = {}

        def arr(name, a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep[name] = a
            return a.ctypes.data

        c = _Case()
        c.nc, c.nv = case.nc, case.nv
        c.ncolors = int(case.color.max()) + 1
        c.pin = int(case.pin)
        c.cell_ref = arr("cell_ref", case.cell_ref, np.int64)
        c.cell_slot = arr("cell_slot", case.cell_slot, np.uint8)
        c.cell_det = arr("cell_det", case.cell_det, np.float64)
        c.cell_invJ = arr("cell_invJ", case.cell_invJ, np.float64)
        c.bcoef = arr("bcoef", case.bcoef, np.float64)
        c.fixed_value = arr("fixed_value", case.fixed_value, np.float64)
        c.mom_row = arr("mom_row", case.mom_row, np.int64)
        c.mom_col = arr("mom_col", case.mom_col, np.int32)
        c.mom_diff = arr("mom_diff", case.mom_diff, np.float64)
        c.color = arr("color", case.color, np.int32)
        c.static_rhs = arr("static_rhs", case.static_rhs, np.float64)
        c.fixed_div = arr("fixed_div", case.fixed_div, np.float64)
        c.tcent = arr("tcent", tcent, np.float64)
        for k, src in (("f_row", "row"), ("f_col", "col"), ("f_diag", "diag_pos"), ("inc_off", "inc_off"),
                       ("inc_c", "inc_c"), ("inc_a", "inc_a"), ("contrib_off", "contrib_off"), ("slot", "slot")):
            setattr(c, k, arr(k, T[src], np.int64))
        H = case.hier
        nl = len(H["A"])
        levels = (_Level * nl)()
        for l in range(nl):
            P, A = H["P"][l].tocsr(), H["A"][l].tocsr()
            lv = levels[l]
            lv.nfine, lv.ncoarse = P.shape
            lv.p_row = arr(f"p_row{l}", P.indptr, np.int64)
            lv.p_col = arr(f"p_col{l}", P.indices, np.int32)
            lv.p_val = arr(f"p_val{l}", P.data, np.float64)
            lv.a_row = arr(f"a_row{l}", A.indptr, np.int64)
            lv.a_col = arr(f"a_col{l}", A.indices, np.int32)
            lv.a_val = arr(f"a_val{l}", A.data, np.float64)
        h = _Hier()
        h.nlevels = nl
        h.nterm = int(H["Tinv"].shape[0])
        h.levels = levels
        h.tinv = arr("tinv", H["Tinv"], np.float64)
        self._keep, self._c, self._h, self._levels = keep, c, h, levels
        self.handle = L.cpu_h8_open(C.byref(c), C.byref(h), jacobi_omega)

    def close(self):
        if getattr(self, "handle", None):
            self.L.cpu_h8_close(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def max_threads(self) -> int:
        return int(self.L.cpu_h8_threads())

    def simple(self, iters: int, alpha_u=0.5, alpha_p=0.5, p_rtol=0.5, p_atol=1e-12, p_max=200, restart=30,
               threads=0, U0=None, p0=None) -> dict:
        """`iters` fixed SIMPLE iterations (FGS momentum, FGMRES + SA V-cycle
        with damped-Jacobi smoothing) from U0 / p0 (default rest)."""
        case = self.case
        U = np.zeros((3, case.nv)) if U0 is None else np.array(U0, dtype=np.float64, order="C")
        p = np.zeros(case.nc) if p0 is None else np.array(p0, dtype=np.float64)
        stats = np.zeros((iters, 3))
        rc = self.L.cpu_h8_simple(self.handle, iters, alpha_u, alpha_p, p_rtol, p_atol, p_max, restart, threads,
                                  U.ctypes.data, p.ctypes.data, stats.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"cpu_h8_simple failed ({rc})")
        return {"U": U, "p": p, "cont": stats[:, 0], "p_its": stats[:, 1].astype(int), "p_rel": stats[:, 2]}
