"""ctypes binding of libnodals_hip.so (include/nodals_hip.h).

The library is built in-tree by ``nodals_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Dict, Optional, Sequence

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libnodals_hip.so"


class NodalsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


class SolveResult(C.Structure):
    _fields_ = [("its", C.c_int32), ("ok", C.c_int32), ("rel", C.c_double), ("r0", C.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {"its": int(self.its), "ok": bool(self.ok), "rel": float(self.rel), "r0": float(self.r0)}


class RunResult(C.Structure):
    _fields_ = [("outer", C.c_int32), ("converged", C.c_int32), ("pass_", C.c_int32), ("reserved", C.c_int32),
                ("final_rel_cont", C.c_double), ("avg_pressure_its", C.c_double), ("pressure_drop", C.c_double),
                ("exact_pressure_drop", C.c_double), ("loop_ms", C.c_double), ("assembly_ms", C.c_double),
                ("momentum_ms", C.c_double), ("continuity_ms", C.c_double), ("pressure_ms", C.c_double),
                ("pupdate_ms", C.c_double), ("setup_ms", C.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib: Optional[C.CDLL] = None

EXPORTS = [
    "nodals_last_error", "nodals_version",
    "nodals_h8_create", "nodals_h8_create_host", "nodals_h8_destroy", "nodals_h8_info", "nodals_h8_assemble", "nodals_h8_momentum_work",
    "nodals_h8_momentum_audit", "nodals_h8_continuity", "nodals_h8_pressure_refresh", "nodals_h8_pressure_solve",
    "nodals_h8_pressure_update", "nodals_h8_run", "nodals_h8_get", "nodals_h8_set", "nodals_h8_fine_spmv",
    "nodals_h8_precond",
    "nodals_h8_time_kernel", "nodals_h8_probe",
    "nodals_sys_create", "nodals_sys_create_host", "nodals_sys_get", "nodals_sys_destroy", "nodals_sys_info",
    "nodals_sys_spmv", "nodals_sys_precond",
    "nodals_sys_solve", "nodals_sys_solve_timed", "nodals_sys_set_rhs", "nodals_sys_time_spmv",
]


def lib() -> C.CDLL:
    """Load libnodals_hip.so (built in-tree).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m nodals_amd.build` (no CPU fallback exists)")
    L = C.CDLL(str(LIB_PATH))
    L.nodals_last_error.restype = C.c_char_p
    L.nodals_version.restype = C.c_char_p
    vp = C.c_void_p
    L.nodals_h8_create.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(vp)]
    L.nodals_h8_create_host.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(vp)]
    L.nodals_h8_destroy.argtypes = [vp]
    L.nodals_h8_destroy.restype = None
    L.nodals_h8_info.argtypes = [vp, C.POINTER(C.c_int64)]
    L.nodals_h8_assemble.argtypes = [vp, C.c_double, C.c_double]
    L.nodals_h8_momentum_work.argtypes = [vp, C.c_char_p, C.c_double, C.c_int32]
    L.nodals_h8_momentum_audit.argtypes = [vp, C.POINTER(C.c_double)]
    L.nodals_h8_continuity.argtypes = [vp, C.POINTER(C.c_double)]
    L.nodals_h8_pressure_refresh.argtypes = [vp]
    L.nodals_h8_pressure_solve.argtypes = [vp, C.c_char_p, C.c_double, C.c_double, C.c_int32, C.POINTER(SolveResult)]
    L.nodals_h8_pressure_update.argtypes = [vp, C.c_double]
    L.nodals_h8_run.argtypes = [vp, C.POINTER(RunResult)]
    L.nodals_h8_get.argtypes = [vp, C.c_char_p, vp, C.c_int64]
    L.nodals_h8_set.argtypes = [vp, C.c_char_p, vp, C.c_int64]
    L.nodals_h8_fine_spmv.argtypes = [vp, vp, vp]
    L.nodals_h8_precond.argtypes = [vp, vp, vp]
    L.nodals_h8_time_kernel.argtypes = [vp, C.c_char_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.nodals_h8_probe.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.nodals_sys_create.argtypes = [C.c_int32, vp, vp, vp, C.c_char_p, C.POINTER(vp)]
    L.nodals_sys_create_host.argtypes = [C.c_int32, vp, vp, vp, C.c_char_p, C.POINTER(vp)]
    L.nodals_sys_get.argtypes = [vp, C.c_char_p, vp, C.c_int64]
    L.nodals_sys_destroy.argtypes = [vp]
    L.nodals_sys_destroy.restype = None
    L.nodals_sys_info.argtypes = [vp, C.POINTER(C.c_int64)]
    L.nodals_sys_spmv.argtypes = [vp, vp, vp]
    L.nodals_sys_precond.argtypes = [vp, vp, vp]
    L.nodals_sys_solve.argtypes = [vp, vp, vp, C.c_char_p, C.c_double, C.c_int32, C.c_int32, C.POINTER(SolveResult)]
    L.nodals_sys_solve_timed.argtypes = [vp, C.c_char_p, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(C.c_double), C.POINTER(SolveResult)]
    L.nodals_sys_set_rhs.argtypes = [vp, vp]
    L.nodals_sys_time_spmv.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise NodalsError(rc, lib().nodals_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


_DTYPES = {
    "mom_row": np.int64, "mom_col": np.int32, "color": np.int32, "g2free": np.int32, "fine_row": np.int32,
    "fine_col": np.int32, "cell_ref": np.int32, "cell_slot": np.uint8,
}


class H8:
    """H8 laminar SIMPLE solver context on the GPU (argv as the H8 binary)."""

    def __init__(self, argv: Sequence[str], device: bool = True):
        L = lib()
        args = ["nodals_gpu_h8"] + [str(a) for a in argv]
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        h = C.c_void_p()
        create = L.nodals_h8_create if device else L.nodals_h8_create_host
        _check(create(len(args), arr, C.byref(h)))
        self._h = h
        info = (C.c_int64 * 8)()
        _check(L.nodals_h8_info(self._h, info))
        self.info = {k: int(v) for k, v in zip(
            ["cells", "nv", "mom_nnz", "colors", "fine_nnz", "amg_levels", "terminal", "nfixed"], info)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nodals_h8_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # solver calls -------------------------------------------------------
    def assemble(self, alpha_u: float, rau_scale: float = 1.0) -> None:
        _check(lib().nodals_h8_assemble(self._h, alpha_u, rau_scale))

    def momentum_work(self, mode: str = "fgs1", omega: float = 1.0, outer: int = 1) -> None:
        _check(lib().nodals_h8_momentum_work(self._h, mode.encode(), omega, outer))

    def momentum_audit(self) -> np.ndarray:
        r = (C.c_double * 3)()
        _check(lib().nodals_h8_momentum_audit(self._h, r))
        return np.array(r[:])

    def continuity(self) -> float:
        n = C.c_double()
        _check(lib().nodals_h8_continuity(self._h, C.byref(n)))
        return n.value

    def pressure_refresh(self) -> None:
        _check(lib().nodals_h8_pressure_refresh(self._h))

    def pressure_solve(self, solver: str, rtol: float, atol: float, maxit: int) -> Dict[str, float]:
        r = SolveResult()
        _check(lib().nodals_h8_pressure_solve(self._h, solver.encode(), rtol, atol, maxit, C.byref(r)))
        return r.as_dict()

    def pressure_update(self, alpha_p: float) -> None:
        _check(lib().nodals_h8_pressure_update(self._h, alpha_p))

    def run(self) -> Dict[str, float]:
        r = RunResult()
        _check(lib().nodals_h8_run(self._h, C.byref(r)))
        return r.as_dict()

    def fine_spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().nodals_h8_fine_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def precond(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        _check(lib().nodals_h8_precond(self._h, _ptr(r), _ptr(z)))
        return z

    def time_kernel(self, name: str, reps: int):
        ms, by = C.c_double(), C.c_double()
        _check(lib().nodals_h8_time_kernel(self._h, name.encode(), reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def probe_start(self):
        _check(lib().nodals_h8_probe(self._h, 1, None, None))

    def probe_stop(self):
        """(launches, summed ms) of the fine-SpMV launches since probe_start."""
        n, ms = C.c_int64(), C.c_double()
        _check(lib().nodals_h8_probe(self._h, 0, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    # host access --------------------------------------------------------
    def get(self, name: str, count: int, dtype=None) -> np.ndarray:
        dt = dtype or _DTYPES.get(name, np.float64)
        out = np.empty(count, dtype=dt)
        _check(lib().nodals_h8_get(self._h, name.encode(), _ptr(out), count))
        return out

    def set(self, name: str, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().nodals_h8_set(self._h, name.encode(), _ptr(v), v.size))


class System:
    """Explicit FP64 SPD CSR system with a device SA hierarchy."""

    def __init__(self, row: np.ndarray, col: np.ndarray, val: np.ndarray, options: str = "", device: bool = True):
        self._row = np.ascontiguousarray(row, dtype=np.int64)
        self._col = np.ascontiguousarray(col, dtype=np.int32)
        self._val = np.ascontiguousarray(val, dtype=np.float64)
        n = self._row.size - 1
        h = C.c_void_p()
        create = lib().nodals_sys_create if device else lib().nodals_sys_create_host
        _check(create(n, _ptr(self._row), _ptr(self._col), _ptr(self._val), options.encode(), C.byref(h)))
        self._h = h
        self.n = n
        info = (C.c_int64 * 16)()
        _check(lib().nodals_sys_info(self._h, info))
        self.levels = int(info[2])
        self.terminal = int(info[3])
        self.level_rows = [int(info[4 + l]) for l in range(min(self.levels, 12))]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nodals_sys_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().nodals_sys_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def get(self, name: str, count: int, dtype=np.float64) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        _check(lib().nodals_sys_get(self._h, name.encode(), _ptr(out), count))
        return out

    def hierarchy(self) -> dict:
        """Host SA hierarchy as scipy CSR matrices (oracle V-cycle input)."""
        import scipy.sparse as sp

        H = {"P": [], "A": []}
        nfine = self.n
        for l in range(self.levels):
            pr = self.get(f"P{l}_row", nfine + 1, np.int64)
            pc = self.get(f"P{l}_col", int(pr[-1]), np.int32)
            pv = self.get(f"P{l}_val", int(pr[-1]))
            nc = self.level_rows[l]
            ar = self.get(f"A{l}_row", nc + 1, np.int64)
            ac = self.get(f"A{l}_col", int(ar[-1]), np.int32)
            av = self.get(f"A{l}_val", int(ar[-1]))
            H["P"].append(sp.csr_matrix((pv, pc, pr), shape=(nfine, nc)))
            H["A"].append(sp.csr_matrix((av, ac, ar), shape=(nc, nc)))
            nfine = nc
        H["Tinv"] = self.get("terminal_inv", nfine * nfine).reshape(nfine, nfine)
        H["fine_lambda"] = float(self.get("fine_lambda", 1)[0])
        H["level_lambda"] = self.get("level_lambda", self.levels)
        return H

    def precond(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        _check(lib().nodals_sys_precond(self._h, _ptr(r), _ptr(z)))
        return z

    def solve(self, b: np.ndarray, solver: str = "fgmres", rtol: float = 1e-8, restart: int = 30, maxit: int = 500):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        r = SolveResult()
        _check(lib().nodals_sys_solve(self._h, _ptr(b), _ptr(x), solver.encode(), rtol, restart, maxit, C.byref(r)))
        return x, r.as_dict()

    def set_rhs(self, b: np.ndarray) -> None:
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().nodals_sys_set_rhs(self._h, _ptr(b)))

    def solve_timed(self, solver: str = "fgmres", rtol: float = 1e-8, restart: int = 30, maxit: int = 500,
                    reps: int = 1):
        ms = C.c_double()
        r = SolveResult()
        _check(lib().nodals_sys_solve_timed(self._h, solver.encode(), rtol, restart, maxit, reps, C.byref(ms),
                                            C.byref(r)))
        return ms.value, r.as_dict()

    def time_spmv(self, level: int, reps: int):
        ms, by = C.c_double(), C.c_double()
        _check(lib().nodals_sys_time_spmv(self._h, level, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value
